"""GPU parity: the HIP path against the CPU oracle on identical seeded inputs.

FP64 mode: forces, per-particle energies and virials, external energy and virial within
1e-10 relative RMS, and bit-for-bit reproducible (the mesh accumulates in int64 fixed point).
Mixed mode: within 2e-6 of the FP64 oracle at the same parameters.
"""
import numpy as np
import pytest

from conftest import random_electrolyte, rel_rms
from esp_cases import make_params, oracle_for, run_gpu

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def _compare(res, ref, tol):
    assert rel_rms(res["force"][:, :3], ref["force"][:, :3]) < tol
    assert rel_rms(res["force"][:, 3], ref["force"][:, 3]) < tol
    assert rel_rms(res["virial"], ref["virial"]) < tol
    assert abs(res["ext"][0] - ref["ext_energy"]) <= tol * abs(ref["ext_energy"])
    assert rel_rms(res["ext"][1:7], ref["ext_virial"]) < tol
    assert abs(res["energy"] - ref["energy"]) <= tol * abs(ref["energy"])


@pytest.mark.parametrize("N,L,mesh,P,c,rc", [
    (100, 8.0, 32, 8, 12.0, 3.0),
    (300, 10.0, (24, 30, 20), 7, 10.0, 3.5),
    (64, 6.0, 16, 4, 6.0, 2.5),
    (500, 12.0, 32, 12, 14.0, 3.0),
])
def test_fp64_pswf_matches_oracle(N, L, mesh, P, c, rc):
    pos, q = random_electrolyte(N, L, seed=N)
    box = (L, L, L)
    params = make_params(mesh, P, c, rc)
    solver, _, res = run_gpu(params, pos, q, box)
    ref = oracle_for(solver, box, params).compute(pos, q)
    print(f"case N={N} mesh={mesh} P={P}", flush=True)
    _compare(res, ref, FP64_TOL)


def test_fp64_gaussian_split_matches_oracle():
    pos, q = random_electrolyte(200, 9.0, seed=7)
    box = (9.0, 9.0, 9.0)
    params = make_params(36, 9, None, 4.0, split=1, kappa=1.6)
    solver, _, res = run_gpu(params, pos, q, box)
    ref = oracle_for(solver, box, params).compute(pos, q)
    _compare(res, ref, FP64_TOL)


def test_fp64_rectangular_box_and_exclusions():
    pos, q = random_electrolyte(150, 9.0, seed=3)
    box = (9.0, 10.0, 11.0)
    excl = [(0, 1), (2, 3), (4, 9), (10, 100)]
    params = make_params((30, 32, 36), 6, 10.0, 3.0)
    solver, _, res = run_gpu(params, pos, q, box, exclusions=excl)
    ref = oracle_for(solver, box, params).compute(pos, q, exclusions=excl)
    _compare(res, ref, FP64_TOL)


def test_fp64_group_and_zero_charges():
    pos, q = random_electrolyte(120, 8.0, seed=5)
    q[::7] = 0.0
    q -= q.mean()  # keep neutral with fractional charges
    box = (8.0, 8.0, 8.0)
    params = make_params(32, 8, 12.0, 3.0)
    solver, _, res = run_gpu(params, pos, q, box, group=np.random.default_rng(0).permutation(120).astype(np.uint32))
    ref = oracle_for(solver, box, params).compute(pos, q)
    _compare(res, ref, FP64_TOL)
    qs, q2s = solver.charge_sums()
    assert abs(qs - q.sum()) < 1e-12 and abs(q2s - (q * q).sum()) < 1e-10


@pytest.mark.parametrize("precision,tol", [(0, FP64_TOL), (1, 3e-6)])
def test_screening_alpha_matches_oracle(precision, tol):
    """Debye screening alpha > 0 (the influence function carries exp(-alpha |k|),
    ESPForceComputeGPU.cu:199) in a rectangular box with a net charge: forces, energies and
    virials against the oracle at the same alpha, and the screening changes the result."""
    pos, q = random_electrolyte(200, 9.0, seed=17)
    q[3] = 2.0
    box = (9.0, 9.5, 10.0)
    params = make_params((30, 32, 32), 8, 12.0, 3.0, precision=precision, alpha=0.35)
    solver, _, res = run_gpu(params, pos, q, box)
    ref = oracle_for(solver, box, params).compute(pos, q)
    err = rel_rms(res["force"][:, :3], ref["force"][:, :3])
    print(f"alpha=0.35 precision={precision} force err {err:.2e}", flush=True)
    if precision == 0:
        _compare(res, ref, tol)
    else:
        assert err < tol
        assert abs(res["energy"] - ref["energy"]) < 1e-5 * abs(ref["energy"])
    params0 = make_params((30, 32, 32), 8, 12.0, 3.0, precision=precision)
    _, _, res0 = run_gpu(params0, pos, q, box)
    assert rel_rms(res0["force"][:, :3], res["force"][:, :3]) > 1e-3


def test_fp64_non_neutral_background():
    pos, q = random_electrolyte(90, 8.0, seed=11)
    q[0] = 2.0  # net charge
    box = (8.0, 8.0, 8.0)
    params = make_params(32, 8, 12.0, 3.0)
    solver, _, res = run_gpu(params, pos, q, box)
    ref = oracle_for(solver, box, params).compute(pos, q)
    _compare(res, ref, FP64_TOL)


@pytest.mark.parametrize("P", [6, 8, 10])
def test_mixed_precision_close_to_fp64_oracle(P):
    pos, q = random_electrolyte(400, 12.0, seed=P)
    box = (12.0, 12.0, 12.0)
    params = make_params(40, P, 12.0, 3.5, precision=1)
    solver, _, res = run_gpu(params, pos, q, box)
    ref = oracle_for(solver, box, params).compute(pos, q)
    err = rel_rms(res["force"][:, :3], ref["force"][:, :3])
    print(f"mixed P={P} force err {err:.2e}", flush=True)
    assert err < 2e-6
    assert abs(res["energy"] - ref["energy"]) < 1e-5 * abs(ref["energy"])


@pytest.mark.parametrize("P,c,sigma", [(10, 11.0, 1.15), (8, 11.4, 1.44), (12, 14.0, 1.2)])
def test_mixed_precision_at_low_oversampling(P, c, sigma):
    """At small mesh oversampling sigma the window deconvolution reaches ~1e16 in the mesh
    corners; the reciprocal cutoff at 1.2 kappa keeps FP32 round-off from being amplified
    there.  Mixed results stay within 3e-6 of the FP64 oracle (same cutoff)."""
    L, rc = 24.0, 4.5
    kappa = c / rc
    from esp_hoomd_amd.accuracy import fft_size_at_least
    n = fft_size_at_least(int(round(L * sigma * kappa / np.pi)))
    pos, q = random_electrolyte(1382, L, seed=P)
    box = (L, L, L)
    params = make_params(n, P, c, rc, precision=1)
    solver, _, res = run_gpu(params, pos, q, box)
    ref = oracle_for(solver, box, params).compute(pos, q)
    err = rel_rms(res["force"][:, :3], ref["force"][:, :3])
    print(f"mixed low-sigma P={P} n={n} force err {err:.2e}", flush=True)
    assert err < 3e-6
    assert abs(res["energy"] - ref["energy"]) < 1e-5 * abs(ref["energy"])


def test_empty_and_tiny_systems():
    params = make_params(16, 4, 6.0, 2.5)
    pos, q = random_electrolyte(2, 6.0, seed=1)
    solver, _, res = run_gpu(params, pos, q, (6.0, 6.0, 6.0))
    ref = oracle_for(solver, (6.0, 6.0, 6.0), params).compute(pos, q)
    _compare(res, ref, FP64_TOL)
    # all charges zero: forces and energies vanish
    solver2, _, res2 = run_gpu(params, pos, np.zeros(2), (6.0, 6.0, 6.0))
    assert np.all(res2["force"] == 0.0) and res2["ext"][0] == 0.0


def test_reference_table_lookup_on_device_matches_golden():
    """An external PSWF_Coeffs.h-format table is evaluated with the reference lookup
    (EvaluatorPairPSWF.h:113-143) by the device pair kernel: isolated like-charge pairs at the
    golden distances give E_i = L(r)/2 and F_x = -(-dL/dr)."""
    import json
    from pathlib import Path
    from esp_hoomd_amd.solver import DeviceSystem
    d = json.loads((Path(__file__).resolve().parent / "golden" / "pswf_ref_lookup.json").read_text())
    n = d["n_segs"]
    scr = np.zeros((n, 5)); dscr = np.zeros((n, 5))
    rlo = np.arange(n) / n; dri = np.full(n, float(n))
    for s, v in d["segments"].items():
        s = int(s)
        scr[s], dscr[s], rlo[s], dri[s] = v["screen"], v["dscreen"], v["r_lo"], v["dr_inv"]
    looks = [x for x in d["lookups"] if 1e-6 < x["r"] < 1.0]
    K = len(looks)
    m = int(np.ceil(K ** (1 / 3)))
    L = 3.0 * m
    pos = []
    for k, x in enumerate(looks):
        base = np.array([k % m, (k // m) % m, k // (m * m)]) * 3.0 - L / 2 + 0.5
        pos += [base, base + np.array([x["r"], 0, 0])]
    pos = np.array(pos)
    q = np.ones(2 * K)
    params = make_params(16, 4, 9.0, 1.0)
    from esp_hoomd_amd.solver import Solver
    solver = Solver(params)
    solver.set_split_table(scr, dscr, rlo, dri)
    sysd = DeviceSystem(pos, q, (L, L, L), r_list=1.1)
    res = solver.pair_forces(sysd)
    # the distance the kernel sees is fl(base + r) - base; evaluate the (golden-pinned) oracle
    # lookup there and also check closeness to the golden value at r itself
    import oracle as O
    r_eff = pos[1::2, 0] - pos[0::2, 0]
    Lr, nd, _, _ = O.split_eval(scr, dscr, rlo, dri, 1.0, r_eff)
    np.testing.assert_allclose(res["force"][0::2, 3], 0.5 * Lr, rtol=1e-14, atol=1e-24)
    np.testing.assert_allclose(res["force"][0::2, 0], -nd, rtol=1e-13, atol=1e-12)
    for k, x in enumerate(looks):
        assert abs(res["force"][2 * k, 3] - 0.5 * x["L"]) <= 1e-9 * abs(x["L"]) + 1e-20


@pytest.mark.parametrize("tile", [4, 10, 12])
def test_mixed_result_is_independent_of_the_tile_edge(tile):
    """Mixed precision accumulates the mesh in fixed point and gathers each particle from the
    same potential values in the same order, so any tile edge gives the same bits."""
    N, L = 600, 12.0
    pos, q = random_electrolyte(N, L, seed=31)
    _, _, ref = run_gpu(make_params(40, 8, 12.0, 3.5, precision=1), pos, q, (L, L, L))
    print(f"tile {tile}", flush=True)
    _, _, res = run_gpu(make_params(40, 8, 12.0, 3.5, precision=1, tile=tile), pos, q, (L, L, L))
    assert np.array_equal(res["force"], ref["force"])
    assert np.array_equal(res["virial"], ref["virial"])
    assert np.array_equal(res["ext"], ref["ext"])


def test_tile_edges_outside_the_kernels_are_refused():
    pos, q = random_electrolyte(100, 8.0, seed=2)
    with pytest.raises(Exception):
        run_gpu(make_params(32, 8, 12.0, 3.0, precision=0, tile=16), pos, q, (8.0, 8.0, 8.0))  # FP64 box 23^3 > 64 KiB
    with pytest.raises(Exception):
        run_gpu(make_params(32, 8, 12.0, 3.0, precision=1, tile=17), pos, q, (8.0, 8.0, 8.0))


def test_mixed_fixed_point_follows_small_charges():
    """The fixed-point mesh scale follows the largest |q| (2^23 / 2^ceil(log2 max|q|)): a system of
    +-0.01 charges keeps the relative accuracy of the same system with +-1 charges."""
    N, L = 800, 12.0
    pos, q = random_electrolyte(N, L, seed=17)
    errs = {}
    for scale in (1.0, 0.01):
        p64 = make_params(32, 8, 12.0, 3.5, precision=0)
        p32 = make_params(32, 8, 12.0, 3.5, precision=1)
        _, _, ref = run_gpu(p64, pos, q * scale, (L, L, L))
        solver, _, res = run_gpu(p32, pos, q * scale, (L, L, L))
        errs[scale] = rel_rms(res["force"][:, :3], ref["force"][:, :3])
        fx, lowered = solver.fixed_point_scale()
        print(f"charges x{scale}: mixed vs FP64 {errs[scale]:.3e}, fixed-point scale 2^{np.log2(fx):.0f}", flush=True)
        assert fx == (2.0 ** 23 if scale == 1.0 else 2.0 ** 29) and not lowered
    assert errs[0.01] < 1e-6
    assert errs[0.01] < 3 * errs[1.0]


def test_mixed_fixed_point_guards_dense_high_charge_tiles():
    """50,000 particles in a 10^3 box with 16-cell tiles put ~6,000 charges in a tile, whose int32
    LDS cells the 2^23 scale would overflow: the scale is lowered and the result stays close to
    FP64."""
    N, L = 50_000, 10.0
    rng = np.random.default_rng(5)
    pos = rng.uniform(-L / 2, L / 2, (N, 3))
    q = np.where(np.arange(N) % 2 == 0, 2.0, -2.0)
    p64 = make_params(32, 4, 8.0, 1.0, precision=0)
    p32 = make_params(32, 4, 8.0, 1.0, precision=1, tile=16)
    _, _, ref = run_gpu(p64, pos, q, (L, L, L), skin=0.1)
    solver, _, res = run_gpu(p32, pos, q, (L, L, L), skin=0.1)
    fx, lowered = solver.fixed_point_scale()
    err = rel_rms(res["force"][:, :3], ref["force"][:, :3])
    print(f"dense tiles: scale 2^{np.log2(fx):.0f} lowered={lowered}, mixed vs FP64 {err:.3e}", flush=True)
    assert lowered and fx < 2.0 ** 22
    assert err < 1e-4


def _permuted_system(pos, q, box, r_list, perm):
    """The same system with particle k = old particle perm[k]; every neighbour row keeps its
    order (mapped to the new indices), as a caller that relabels particles would pass it."""
    from esp_hoomd_amd.nlist import build_full_nlist
    from esp_hoomd_amd.solver import DeviceSystem
    n, lst, head = build_full_nlist(pos, box, r_list)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    n_new = n[perm]
    rows = [inv[lst[int(head[i]):int(head[i]) + int(n[i])]] for i in perm]
    head_new = np.concatenate([[0], np.cumsum(n_new)[:-1]]).astype(np.uint64)
    sysd = DeviceSystem(pos[perm], q[perm], box)
    sysd.set_nlist(n_new, np.concatenate(rows).astype(np.uint32), head_new)
    return sysd, inv


def test_fp64_is_bitwise_reproducible_and_independent_of_particle_order():
    """The FP64 mesh is summed in int64 fixed point and every particle's force, energy and
    virial is reduced by one thread in its neighbour-row order, so the same input gives the same
    bits on every run, and relabelling the particles (rows kept in order) permutes the outputs
    exactly -- the external energy and virial included."""
    from esp_hoomd_amd.solver import Solver
    N, L = 3000, 16.0
    pos, q = random_electrolyte(N, L, seed=41)
    params = make_params(40, 8, 12.0, 3.5, precision=0)
    solver, _, a = run_gpu(params, pos, q, (L, L, L), skin=0.0)
    _, _, b = run_gpu(params, pos, q, (L, L, L), skin=0.0)
    for k in ("force", "virial", "ext"):
        assert np.array_equal(a[k], b[k]), k
    fx, lowered = solver.fixed_point_scale()
    print(f"FP64 fixed-point scale 2^{np.log2(fx):.0f}, lowered={lowered}", flush=True)
    assert fx == 2.0 ** 50 and not lowered
    perm = np.random.default_rng(42).permutation(N)
    sysd, inv = _permuted_system(pos, q, (L, L, L), 3.5, perm)
    c = Solver(params).compute(sysd)
    assert np.array_equal(c["force"][inv], a["force"])
    assert np.array_equal(c["virial"][:, inv], a["virial"])
    assert np.array_equal(c["ext"], a["ext"])


@pytest.mark.parametrize("tile", [2, 4, 12])
def test_fp64_result_is_independent_of_the_tile_edge(tile):
    """Each particle's stencil products are the same whatever tile holds it, and their int64 sums
    do not depend on how tiles split them, so FP64 results match the default tiling bit for bit."""
    N, L = 600, 12.0
    pos, q = random_electrolyte(N, L, seed=31)
    _, _, ref = run_gpu(make_params(40, 8, 12.0, 3.5, precision=0, tile=8), pos, q, (L, L, L))
    print(f"tile {tile}", flush=True)
    _, _, res = run_gpu(make_params(40, 8, 12.0, 3.5, precision=0, tile=tile), pos, q, (L, L, L))
    assert np.array_equal(res["force"], ref["force"])
    assert np.array_equal(res["ext"], ref["ext"])


def test_fp64_fixed_point_guards_dense_high_charge_tiles():
    """60,000 charges of +-2 packed into a 2^3 cube put at least 7,500 into one tile; over the
    tile boxes that overlap a mesh cell the int64 cells could overflow at the 2^49 scale of
    |q| = 2, so the scan lowers the scale, and the result keeps FP64 accuracy against the
    oracle."""
    N, L = 60_000, 10.0
    rng = np.random.default_rng(5)
    pos = rng.uniform(0.0, 2.0, (N, 3))
    q = np.where(np.arange(N) % 2 == 0, 2.0, -2.0)
    p64 = make_params(32, 4, 3.0, 0.15, precision=0)
    solver, _, res = run_gpu(p64, pos, q, (L, L, L), skin=0.0)
    fx, lowered = solver.fixed_point_scale()
    print(f"FP64 dense tiles: scale 2^{np.log2(fx):.0f} lowered={lowered}", flush=True)
    assert lowered and fx < 2.0 ** 49
    ref = oracle_for(solver, (L, L, L), p64).compute(pos, q)
    assert rel_rms(res["force"][:, :3], ref["force"][:, :3]) < FP64_TOL
    assert rel_rms(res["ext"][1:7], ref["ext_virial"]) < FP64_TOL
