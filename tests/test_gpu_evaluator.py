"""include/EvaluatorPairPSWF.h on the GPU: HOOMD's generic per-particle pair loop over the
evaluator (esp_potential_pair_forces, the PotentialPairGPU<EvaluatorPairPSWF> path of
module.cc:79) against the library's own pair pass (esp_pair_forces) on the same neighbour list."""
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import random_electrolyte, rel_rms
from esp_cases import make_params
from esp_hoomd_amd import _lib
from esp_hoomd_amd.solver import DeviceSystem, Solver, potential_pair_forces

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [_lib.ESP_PREC_FP64, _lib.ESP_PREC_MIXED])
def test_evaluator_loop_matches_the_pair_pass(precision):
    N, L, rc = 2000, 14.0, 3.0
    pos, q = random_electrolyte(N, L, seed=12, q2_frac=0.2)
    solver = Solver(make_params(32, 8, 12.0, rc, precision=precision))
    sysd = DeviceSystem(pos, q, (L, L, L), r_list=rc + 0.3)
    ref = solver.pair_forces(sysd)
    ref = {k: v.copy() for k, v in ref.items()}
    got = potential_pair_forces(sysd, solver.pair_param(), rc)
    # the pair pass reads a compact refit of the same split (and FP32 arithmetic in mixed
    # precision); the evaluator the 4096-segment FP64 table
    tol = 1e-10 if precision == _lib.ESP_PREC_FP64 else 2e-6
    assert rel_rms(got["force"][:, :3], ref["force"][:, :3]) < tol
    assert rel_rms(got["force"][:, 3], ref["force"][:, 3]) < tol
    assert rel_rms(got["virial"], ref["virial"]) < tol


def test_evaluator_loop_on_the_reference_table_matches_the_pair_pass():
    """With the reference's PSWF_Coeffs.h arrays loaded, both paths evaluate the reference's
    lookup verbatim: equal up to the order of the per-particle sums."""
    d = json.loads((Path(__file__).resolve().parent / "golden" / "pswf_ref_lookup.json").read_text())
    n = d["n_segs"]
    scr = np.zeros((n, 5)); dscr = np.zeros((n, 5))
    rlo = np.arange(n) / n; dri = np.full(n, float(n))
    for s, v in d["segments"].items():
        s = int(s)
        scr[s], dscr[s], rlo[s], dri[s] = v["screen"], v["dscreen"], v["r_lo"], v["dr_inv"]
    # pairs only inside the first 32 segments (r < 32 / 4096), where the golden table has data
    N, L = 400, 6.0
    rng = np.random.default_rng(3)
    centres = rng.uniform(-L / 2 + 0.5, L / 2 - 0.5, (N // 2, 3))
    off = rng.normal(size=(N // 2, 3))
    off *= (rng.uniform(1e-4, 7.5e-3, N // 2) / np.linalg.norm(off, axis=1))[:, None]
    pos = np.concatenate([centres, centres + off])
    q = rng.choice([-1.0, 1.0, 2.0], N)
    solver = Solver(make_params(16, 4, 9.0, 1.0))
    solver.set_split_table(scr, dscr, rlo, dri)
    sysd = DeviceSystem(pos, q, (L, L, L), r_list=0.0078)
    ref = {k: v.copy() for k, v in solver.pair_forces(sysd).items()}
    got = potential_pair_forces(sysd, solver.pair_param(), 1.0)
    assert np.abs(ref["force"][:, :3]).max() > 0
    assert rel_rms(got["force"][:, :3], ref["force"][:, :3]) < 1e-14
    assert rel_rms(got["force"][:, 3], ref["force"][:, 3]) < 1e-14
    assert rel_rms(got["virial"], ref["virial"]) < 1e-14


def test_evaluator_loop_rejects_bad_input():
    pos, q = random_electrolyte(10, 5.0, seed=1)
    solver = Solver(make_params(16, 4, 8.0, 2.0))
    sysd = DeviceSystem(pos, q, (5.0, 5.0, 5.0), r_list=2.2)
    p = solver.pair_param()
    p.n_segs = 0
    with pytest.raises(_lib.EspError):
        potential_pair_forces(sysd, p, 2.0)
