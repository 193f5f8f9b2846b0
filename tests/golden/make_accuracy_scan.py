"""Write tests/golden/accuracy_scan.json: the measurements behind the autotuner's accuracy model
(esp_hoomd_amd/tune.py: SPLIT_FIT, WINDOW_ERROR, CUT_FACTOR), taken with the CPU oracle's FP64
ESP against its FP64 direct Ewald sum.

Systems: random neutral +-1 electrolytes at number density 0.1 (N = 1,728, L = 25.85; seeds 0
and 1), the model's reference structure.  Each ESP configuration uses the library's own tables
(split, window) for its bandwidth c = kappa r_c, order P and oversampling sigma = pi / (h kappa):

* points: c = 6 ... 20, P = 4 ... 12 and sigma on the model's grid, with the library's default
  reciprocal cutoff |k| <= 1.2 kappa (how the solver runs);
* cut: the same cutoff against none at P = 12, sigma = 2, c = 8 ... 16.

The relative RMS force error of each configuration (RMS over the two systems) is recorded.  Run from the repository root (host only, no GPU):
    python tests/golden/make_accuracy_scan.py
"""
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import oracle as O  # noqa: E402
from esp_hoomd_amd import solver as S  # noqa: E402

OUT = Path(__file__).resolve().parent / "accuracy_scan.json"
N = 1728
RHO = 0.1
L = (N / RHO) ** (1.0 / 3.0)
RC = 5.0
SIGMAS = (1.15, 1.30, 1.50, 1.75, 2.00)


def system(seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-L / 2, L / 2, (N, 3))
    q = rng.permutation(np.where(np.arange(N) % 2 == 0, 1.0, -1.0))
    f, _, _ = O.ewald_auto(pos, q, (L, L, L), tol=1e-15, energy=False)
    return pos, q, f


def esp_error(pos, q, fe, c, P, sigma, k_cut_factor=0.0):
    """Relative RMS force error of ESP(c, P, sigma) on this system; the mesh is the smallest
    FFT size with oversampling >= sigma (the achieved sigma is returned)."""
    kappa = c / RC
    n = math.ceil(L * sigma * kappa / math.pi)
    while True:  # 2-3-5 smooth
        k = n
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            break
        n += 1
    h = L / n
    sig = math.pi / (h * kappa)
    D, cw = S.default_window(P, sig)
    w = S.build_window(P, D, cw)
    scr, dscr, rlo, dri = S.build_split_table(0, kappa, RC, 4096)
    o = O.OracleESP((L, L, L), (n, n, n), P, w, scr, dscr, rlo, dri, RC, kappa=kappa, k_cut=k_cut_factor * kappa)
    f = o.compute(pos, q)["force"][:, :3]
    return float(np.sqrt(((f - fe) ** 2).sum() / (fe ** 2).sum())), sig


def main():
    t0 = time.time()
    systems = [system(s) for s in (0, 1)]

    def err(c, P, sigma, cut):
        vals = [esp_error(pos, q, fe, c, P, sigma, cut) for pos, q, fe in systems]
        return float(np.sqrt(np.mean([v[0] ** 2 for v in vals]))), vals[0][1]

    points = []
    for c in range(6, 21, 2):
        for P in (4, 6, 8, 10, 12):
            for sigma in SIGMAS:
                e, sig = err(float(c), P, sigma, 1.2)
                points.append({"c": c, "P": P, "sigma": sig, "error": e})
                print(f"c={c} P={P} sigma={sig:.3f}: {e:.3e}", flush=True)
    cut = []
    for c in (8, 10, 12, 14, 16):
        e_none, _ = err(float(c), 12, 2.0, 0.0)
        e_cut, _ = err(float(c), 12, 2.0, 1.2)
        cut.append({"c": c, "error_no_cut": e_none, "error_cut": e_cut, "ratio": e_cut / e_none})
        print(f"cut c={c}: {e_cut / e_none:.3f}", flush=True)
    data = {"N": N, "L": L, "rc": RC, "seeds": [0, 1], "k_cut_factor": 1.2, "points": points, "cut": cut,
            "seconds": time.time() - t0}
    OUT.write_text(json.dumps(data, indent=1))
    print(f"wrote {OUT} in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
