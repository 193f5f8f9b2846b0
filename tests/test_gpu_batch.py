"""Batched solves: many systems of uneven size, mesh and stencil order in one launch per stage.

Each system's batched result must equal what its own context computes alone, bit for bit in
both precisions (the mesh accumulates in fixed point, and every per-particle sum has a fixed
order), and match the oracle (FP64: 1e-10 relative RMS).  The captured hipGraph is replayed several times with
new positions and host work in between, and every replay is checked against eager results.
"""
import numpy as np
import pytest

from conftest import random_electrolyte, rel_rms
from esp_cases import make_params, oracle_for
from esp_hoomd_amd import _lib
from esp_hoomd_amd.batch import Batch
from esp_hoomd_amd.nlist import build_full_nlist
from esp_hoomd_amd.solver import DeviceSystem, Solver

pytestmark = pytest.mark.gpu

# (N, L, mesh, P, c, rc): uneven sizes, meshes (incl. non-cubic) and the batched orders
CASES = [
    (100, 8.0, 32, 8, 12.0, 3.0),
    (257, 10.0, (24, 30, 20), 6, 10.0, 3.5),
    (64, 6.0, 16, 4, 6.0, 2.5),
    (500, 12.0, 32, 12, 14.0, 3.0),
    (333, 11.0, 36, 10, 12.0, 3.2),
    (1, 5.0, 16, 4, 6.0, 2.0),
    # xy planes are transformed in one workgroup when they fit its LDS: odd ny (a half-empty
    # line pair per plane), and two members that take the separate x and y passes instead --
    # a 64^3 FP64 plane, and a 120 x 75 plane in either precision (odd ny there too)
    (150, 9.0, (20, 25, 18), 8, 10.0, 3.0),
    (40, 7.0, 64, 4, 6.0, 2.5),
    (60, 8.0, (120, 75, 16), 6, 10.0, 2.5),
]


def _systems(precision, seed0=0, cases=CASES):
    out = []
    for k, (N, L, mesh, P, c, rc) in enumerate(cases):
        pos, q = random_electrolyte(N, L, seed=seed0 + 17 * k + N)
        if N == 1:
            q = np.array([1.0])
        box = (L, L, L)
        params = make_params(mesh, P, c, rc, precision=precision)
        sysd = DeviceSystem(pos, q, box, r_list=min(rc + 0.3, 0.499 * L))
        out.append((Solver(params), sysd, pos, q, box, params))
    return out


def _alone(solver, sysd):
    return solver.compute(sysd)


def _check_equal(a, b, exact, tol=1e-12):
    if exact:
        np.testing.assert_array_equal(a["force"], b["force"])
        np.testing.assert_array_equal(a["virial"], b["virial"])
        np.testing.assert_array_equal(a["ext"], b["ext"])
    else:
        assert rel_rms(a["force"][:, :3], b["force"][:, :3]) < tol
        assert rel_rms(a["force"][:, 3], b["force"][:, 3]) < tol
        assert rel_rms(a["virial"], b["virial"]) < tol
        assert rel_rms(a["ext"], b["ext"]) < tol


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_each_system_alone(precision):
    systems = _systems(precision)
    print(f"batch of {len(systems)} systems, precision {precision}", flush=True)
    alone = [_alone(s, d) for s, d, *_ in systems]
    batch = Batch()
    res = batch.compute([(s, d) for s, d, *_ in systems])
    for k, (a, r) in enumerate(zip(alone, res)):
        print(f"system {k}", flush=True)
        _check_equal(r, a, exact=True)


@pytest.mark.parametrize("precision", [0, 1])
def test_prepared_solvers_give_the_same_bits(precision):
    # set-up done ahead (esp_prepare) leaves nothing different: alone and batched results
    # match solvers that set themselves up in their first computation
    fresh = _systems(precision, seed0=3, cases=CASES[:3])
    early = _systems(precision, seed0=3, cases=CASES[:3])
    for s, d, *_ in early:
        s.prepare(d)
    for (s0, d0, *_), (s1, d1, *_) in zip(fresh, early):
        _check_equal(s1.compute(d1), s0.compute(d0), exact=True)
    early = _systems(precision, seed0=3, cases=CASES[:3])
    for s, d, *_ in early:
        s.prepare(d)
    res = Batch().compute([(s, d) for s, d, *_ in early])
    for (s0, d0, *_), r in zip(fresh, res):
        _check_equal(r, s0.compute(d0), exact=True)


def test_batch_fp64_matches_oracle():
    systems = _systems(0, seed0=5)
    res = Batch().compute([(s, d) for s, d, *_ in systems])
    for (solver, sysd, pos, q, box, params), r in zip(systems, res):
        if len(pos) < 2:
            continue
        ref = oracle_for(solver, box, params).compute(pos, q)
        assert rel_rms(r["force"][:, :3], ref["force"][:, :3]) < 1e-10
        assert rel_rms(r["virial"], ref["virial"]) < 1e-10
        assert abs(r["ext"][0] - ref["ext_energy"]) <= 1e-10 * abs(ref["ext_energy"])
        assert rel_rms(r["ext"][1:7], ref["ext_virial"]) < 1e-10


@pytest.mark.parametrize("use_graph", [True, False])
def test_graph_replays_follow_new_positions(use_graph):
    systems = _systems(1, seed0=9, cases=CASES[:4])
    batch = Batch(use_graph=use_graph)
    rng = np.random.default_rng(3)
    pairs = [(s, d) for s, d, *_ in systems]
    for rep in range(4):
        # new positions in the same device arrays, then eager host work between replays
        for s, d, pos, q, box, params in systems:
            pos[:] = rng.uniform(-box[0] / 2, box[0] / 2, pos.shape)
            d.set_positions(pos)
            d.set_nlist(*build_full_nlist(pos, np.asarray(box), min(params.rcut + 0.3, 0.499 * box[0])))
        print(f"replay {rep}", flush=True)
        res = batch.compute(pairs)
        alone = [_alone(s, d) for s, d in pairs]
        for r, a in zip(res, alone):
            _check_equal(r, a, exact=True)
    st = batch.stats()
    if use_graph:
        # the neighbour lists were reallocated each time (new pointers, same shape):
        # descriptors are refreshed, the graph is captured only when the shape changes
        assert st["replays"] == 4
        assert st["captures"] <= 4
    else:
        assert st["launches"] == 4


def test_shape_change_recaptures_and_stays_correct():
    systems = _systems(1, seed0=21)
    batch = Batch()
    pairs = [(s, d) for s, d, *_ in systems]
    r1 = batch.compute(pairs[:3])
    r2 = batch.compute(pairs[2:])
    r3 = batch.compute(pairs[:3])
    st = batch.stats()
    assert st["captures"] == 3
    for r, (s, d) in zip(r1, pairs[:3]):
        _check_equal(r, _alone(s, d), exact=True)
    for r, (s, d) in zip(r2, pairs[2:]):
        _check_equal(r, _alone(s, d), exact=True)
    for a, b in zip(r1, r3):
        _check_equal(a, b, exact=True)


def test_batch_flags_pair_only_and_mesh_only():
    systems = _systems(1, seed0=33, cases=CASES[:3])
    pairs = [(s, d) for s, d, *_ in systems]
    for flags in (_lib.ESP_COMPUTE_PAIR, _lib.ESP_COMPUTE_MESH | _lib.ESP_COMPUTE_ENERGY_VIRIAL):
        res = Batch().compute(pairs, flags=flags)
        for r, (s, d) in zip(res, pairs):
            s.enqueue(d, flags=flags)
            _lib.check(_lib.load().esp_device_synchronize())
            a = {"force": d.force.numpy(), "virial": d.virial.numpy()[:, : d.N], "ext": d.ext.numpy()[:7]}
            np.testing.assert_array_equal(r["force"], a["force"])
            np.testing.assert_array_equal(r["virial"], a["virial"])


def test_batch_rejects_bad_input():
    systems = _systems(1, cases=CASES[:2])
    s0, d0 = systems[0][:2]
    s1, d1 = systems[1][:2]
    with pytest.raises(_lib.EspError) as ei:
        Batch().compute([(s0, d0), (s0, d1)])
    assert ei.value.code == _lib.ESP_ERR_ARG
    fp64 = _systems(0, cases=CASES[:1])[0]
    with pytest.raises(_lib.EspError) as ei:
        Batch().compute([(s0, d0), (fp64[0], fp64[1])])
    assert ei.value.code == _lib.ESP_ERR_ARG
    odd = Solver(make_params(16, 5, 6.0, 2.5, precision=1))
    with pytest.raises(_lib.EspError) as ei:
        Batch().compute([(odd, d1)])
    assert ei.value.code == _lib.ESP_ERR_UNSUPPORTED
