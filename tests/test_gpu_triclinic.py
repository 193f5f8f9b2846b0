"""Tilted (triclinic) boxes on the GPU: HOOMD's BoxDim with tilt factors xy, xz, yz, the
reference's makeFraction (ESPForceComputeGPU.cu:228-231, 320) for spreading and gathering.

FP64 against the oracle at 1e-10 (forces, per-particle energies and virials, external energy
and virial), mixed precision within 3e-6, both against the general-box Ewald sum, a tilted
system alone and in a batch, the HOOMD evaluator loop's minimum image, and a box whose tilt is a
whole lattice vector against the same system written orthorhombic."""
import numpy as np
import pytest

import oracle as O
from conftest import random_electrolyte, random_electrolyte_box, rel_rms
from esp_cases import make_params, oracle_for
from esp_hoomd_amd import _lib
from esp_hoomd_amd.batch import Batch
from esp_hoomd_amd.solver import DeviceSystem, Solver, potential_pair_forces

pytestmark = pytest.mark.gpu

BOXES = [(9.0, 9.5, 10.0, 0.3, -0.2, 0.15), (10.0, 10.0, 10.0, -0.45, 0.3, 0.5)]


def _run(params, pos, q, box, exclusions=None, skin=0.3):
    solver = Solver(params)
    r_list = min(params.rcut + skin, 0.99 * O.box_half_width(box))
    sysd = DeviceSystem(pos, q, box, r_list=r_list, exclusions=exclusions)
    res = solver.compute(sysd)
    res["energy"] = res["force"][:, 3].sum() + res["ext"][0]
    res["total_virial"] = res["virial"].sum(axis=1) + res["ext"][1:7]
    return solver, sysd, res


def _compare(res, ref, tol):
    assert rel_rms(res["force"][:, :3], ref["force"][:, :3]) < tol
    assert rel_rms(res["force"][:, 3], ref["force"][:, 3]) < tol
    assert rel_rms(res["virial"], ref["virial"]) < tol
    assert abs(res["ext"][0] - ref["ext_energy"]) <= tol * abs(ref["ext_energy"])
    assert rel_rms(res["ext"][1:7], ref["ext_virial"]) < tol
    assert abs(res["energy"] - ref["energy"]) <= tol * abs(ref["energy"])


@pytest.mark.parametrize("box", BOXES)
def test_fp64_tilted_box_matches_oracle(box):
    pos, q = random_electrolyte_box(300, box, seed=4)
    q[5] = 2.0  # net charge: background term
    excl = [(0, 1), (2, 3), (10, 200)]
    params = make_params((36, 36, 40), 8, 12.0, 3.0)
    print(f"fp64 tilted box {box}", flush=True)
    solver, _, res = _run(params, pos, q, box, exclusions=excl)
    ref = oracle_for(solver, box, params).compute(pos, q, exclusions=excl)
    _compare(res, ref, 1e-10)


@pytest.mark.parametrize("box", BOXES)
def test_mixed_tilted_box_close_to_fp64_oracle(box):
    pos, q = random_electrolyte_box(400, box, seed=6)
    params = make_params((36, 36, 40), 10, 12.0, 3.0, precision=1)
    print(f"mixed tilted box {box}", flush=True)
    solver, _, res = _run(params, pos, q, box)
    ref = oracle_for(solver, box, params).compute(pos, q)
    err = rel_rms(res["force"][:, :3], ref["force"][:, :3])
    print(f"mixed tilted force err {err:.2e}", flush=True)
    assert err < 3e-6
    assert abs(res["energy"] - ref["energy"]) < 1e-5 * abs(ref["energy"])


@pytest.mark.parametrize("precision", [0, 1])
def test_tilted_box_against_ewald(precision):
    box = BOXES[0]
    pos, q = random_electrolyte_box(500, box, seed=9, min_dist=0.6)
    fe, ee, we = O.ewald_box_auto(pos, q, box)
    params = make_params(48, 10, 14.0, 4.0, precision=precision)
    _, _, res = _run(params, pos, q, box)
    err = rel_rms(res["force"][:, :3], fe)
    print(f"tilted precision={precision} error vs Ewald {err:.2e}", flush=True)
    assert err < 1e-5
    assert abs(res["energy"] - ee) < 1e-5 * abs(ee)


def test_whole_lattice_vector_tilt_is_the_same_system():
    """xy = 1 with Lx = Ly (a2' = a2 + a1) describes the orthorhombic system again; the mesh is
    sheared, so the two agree to the method's accuracy, each against Ewald."""
    L = 10.0
    pos, q = random_electrolyte(400, L, seed=13, min_dist=0.6)
    fe, ee, _ = O.ewald_auto(pos, q, (L, L, L), rcut=4.9)
    params = make_params(48, 10, 14.0, 3.4)
    _, _, orth = _run(params, pos, q, (L, L, L))
    _, _, tilt = _run(params, pos, q, (L, L, L, 1.0, 0.0, 0.0))
    e_orth = rel_rms(orth["force"][:, :3], fe)
    e_tilt = rel_rms(tilt["force"][:, :3], fe)
    print(f"lattice-equivalent tilt: orthorhombic {e_orth:.2e}, tilted {e_tilt:.2e}", flush=True)
    assert e_orth < 1e-5 and e_tilt < 1e-5
    assert abs(tilt["energy"] - ee) < 1e-5 * abs(ee)


@pytest.mark.parametrize("precision", [0, 1])
def test_tilted_system_in_a_batch_equals_alone(precision):
    entries = []
    for k, box in enumerate(BOXES + [(8.0, 8.0, 8.0)]):
        pos, q = random_electrolyte_box(150 + 50 * k, box, seed=20 + k)
        params = make_params(32, 8, 12.0, 3.0, precision=precision)
        sysd = DeviceSystem(pos, q, box, r_list=min(3.3, 0.99 * O.box_half_width(box)))
        entries.append((Solver(params), sysd))
    alone = [s.compute(d) for s, d in entries]
    alone = [{k: v.copy() for k, v in a.items()} for a in alone]
    res = Batch().compute(entries)
    for a, r in zip(alone, res):
        np.testing.assert_array_equal(r["force"], a["force"])
        np.testing.assert_array_equal(r["virial"], a["virial"])
        np.testing.assert_array_equal(r["ext"], a["ext"])


@pytest.mark.parametrize("precision", [_lib.ESP_PREC_FP64, _lib.ESP_PREC_MIXED])
def test_evaluator_loop_in_tilted_box_matches_the_pair_pass(precision):
    box = BOXES[1]
    rc = 3.0
    pos, q = random_electrolyte_box(1200, box, seed=31)
    solver = Solver(make_params(32, 8, 12.0, rc, precision=precision))
    sysd = DeviceSystem(pos, q, box, r_list=rc + 0.3)
    ref = solver.pair_forces(sysd)
    ref = {k: v.copy() for k, v in ref.items()}
    got = potential_pair_forces(sysd, solver.pair_param(), rc)
    tol = 1e-10 if precision == _lib.ESP_PREC_FP64 else 2e-6
    assert rel_rms(got["force"][:, :3], ref["force"][:, :3]) < tol
    assert rel_rms(got["virial"], ref["virial"]) < tol
