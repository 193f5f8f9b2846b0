"""include/EvaluatorPairPSWF.h on the CPU (HOOMD's PotentialPair<EvaluatorPairPSWF> path): the
evaluator, compiled by g++ with its pybind11 dict constructor, against the oracle's lookup
(itself pinned bit-exactly to the reference's PSWF_Coeffs.h data) and the reference's return
rules (EvaluatorPairPSWF.h:274-295)."""
import importlib
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle as O

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = Path(__file__).resolve().parent / "golden" / "pswf_ref_lookup.json"


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    out = tmp_path_factory.mktemp("evaluator_probe")
    incs = subprocess.run([sys.executable, "-m", "pybind11", "--includes"], capture_output=True, text=True, check=True).stdout.split()
    suffix = subprocess.run([sys.executable, "-c", "import sysconfig; print(sysconfig.get_config_var('EXT_SUFFIX'))"],
                            capture_output=True, text=True, check=True).stdout.strip()
    so = out / f"evaluator_probe{suffix}"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", "-Wextra", *incs, f"-I{ROOT / 'include'}",
           str(Path(__file__).resolve().parent / "evaluator_probe.cpp"), "-o", str(so)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    sys.path.insert(0, str(out))
    try:
        return importlib.import_module("evaluator_probe")
    finally:
        sys.path.remove(str(out))


def _reference_table():
    """The reference's PSWF_Coeffs.h data (rcut = 1) as interleaved records; segments absent
    from the golden file are left zero (no lookup below falls there)."""
    d = json.loads(GOLDEN.read_text())
    n = d["n_segs"]
    scr = np.zeros((n, 5))
    dscr = np.zeros((n, 5))
    rlo = np.arange(n) / n
    dri = np.full(n, float(n))
    for s, v in d["segments"].items():
        s = int(s)
        scr[s], dscr[s], rlo[s], dri[s] = v["screen"], v["dscreen"], v["r_lo"], v["dr_inv"]
    rec = np.zeros((n, 12))
    rec[:, 0], rec[:, 1], rec[:, 2:7], rec[:, 7:12] = rlo, dri, scr, dscr
    return d, rec, scr, dscr, rlo, dri


def test_evaluator_matches_the_reference_lookup(probe):
    d, rec, scr, dscr, rlo, dri = _reference_table()
    looks = [x for x in d["lookups"] if 1e-6 < x["r"] < 1.0]
    r = np.array([x["r"] for x in looks])
    rsq = r * r
    params = {"rcut": 1.0, "kappa": 9.0, "alpha": 0.0, "V_cut": 0.0, "table_ptr": probe.address_of(rec), "n_segs": len(rec)}
    ok, fdivr, eng = probe.eval_pairs(params, 1.0, False, rsq, np.full(len(r), -2.0))
    assert ok.all()
    reff = np.sqrt(rsq)
    Lr, nd, _, _ = O.split_eval(scr, dscr, rlo, dri, 1.0, reff)
    np.testing.assert_array_equal(eng, -2.0 * Lr)
    np.testing.assert_allclose(fdivr, -2.0 * nd / reff, rtol=2e-16, atol=0)
    for k, x in enumerate(looks):  # golden values at r itself
        assert abs(eng[k] + 2.0 * x["L"]) <= 1e-9 * abs(x["L"])


def test_evaluator_rules(probe):
    _, rec, *_ = _reference_table()
    base = {"rcut": 1.0, "table_ptr": probe.address_of(rec), "n_segs": len(rec)}
    rsq = np.array([0.25, 0.25, 1.0, 1.1, 0.25])
    qq = np.array([1.0, 0.0, 1.0, 1.0, 1.0])
    ok, f, e = probe.eval_pairs({**base, "V_cut": 0.125}, 1.0, True, rsq, qq)
    # rsq >= rcutsq and qiqj == 0 are rejected (outputs untouched)
    assert list(ok) == [1, 0, 0, 0, 1]
    assert f[1] == -7.0 and e[3] == -7.0
    ok2, _, e2 = probe.eval_pairs({**base, "V_cut": 0.125}, 1.0, False, rsq, qq)
    assert e[0] == e2[0] - 0.125  # energy shift subtracts qiqj V_cut
    # rcut below 1e-12: no interaction (EvaluatorPairPSWF.h:276-277)
    ok3, _, _ = probe.eval_pairs({**base, "rcut": 1e-13}, 1.0, False, rsq, qq)
    assert not ok3.any()
    facts = probe.facts({**base, "rcut": 0.0})
    assert facts["name"] == "pswf" and facts["needs_charge"] and not facts["needs_diameter"]
    assert facts["useless"] and facts["shape_spec_throws"]
    assert facts["lrc_energy"] == 0.0 and facts["lrc_pressure"] == 0.0
    assert facts["param_size"] == 64 and facts["param_align"] == 16


def test_param_dict_round_trip(probe):
    _, rec, *_ = _reference_table()
    p = {"rcut": 3.5, "kappa": 1.4, "alpha": 0.0, "V_cut": 0.0, "table_ptr": probe.address_of(rec),
         "table_device_ptr": 4096, "n_segs": len(rec)}
    assert probe.round_trip(p) == p
    # keys default to zero (EvaluatorPairPSWF.h:204-212)
    q = probe.round_trip({"rcut": 2.0})
    assert q["kappa"] == 0.0 and q["V_cut"] == 0.0 and q["n_segs"] == 0 and q["table_ptr"] == 0
