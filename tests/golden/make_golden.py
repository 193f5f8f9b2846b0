"""Writes pswf_ref_lookup.json: golden vectors of the reference's S(r)/S'(r) table lookup.

Reads the numeric arrays of the reference's generated table header src/PSWF_Coeffs.h as
text data (PSWF_SCREEN_COEFFS, PSWF_DSCREEN_COEFFS, PSWF_R_LO, PSWF_DR_INV; lines 24-9599),
keeps a sample of segments, and evaluates the lookup of EvaluatorPairPSWF.h:113-143
(clamp s to [0, 1 - 1e-9], seg = min(floor(s * 4096), 4095), t = (r - r_lo) * dr_inv,
Horner c4 -> c0, L = S / (4 pi r), -dL/dr = S / (4 pi r^2) - S' / (4 pi r)) at r_c = 1 with
numpy.polyval.  Run where the reference sources are present (they are not needed by the tests):

    python tests/golden/make_golden.py /root/reference/src/PSWF_Coeffs.h
"""
import json
import re
import sys
from pathlib import Path

import numpy as np

INV_4PI = 0.079577471545947673


def parse_array(text, name):
    m = re.search(r"inline constexpr double " + name + r"\[(\d+)\] = \{(.*?)\};", text, re.S)
    vals = np.array([float(v) for v in m.group(2).replace("\n", " ").split(",") if v.strip()])
    assert len(vals) == int(m.group(1))
    return vals


def main(path):
    text = Path(path).read_text()
    n_segs = int(re.search(r"PSWF_N_SEGS = (\d+);", text).group(1))
    rcut = float(re.search(r"PSWF_RCUT = ([0-9.eE+-]+);", text).group(1))
    scr = parse_array(text, "PSWF_SCREEN_COEFFS").reshape(n_segs, 5)
    dscr = parse_array(text, "PSWF_DSCREEN_COEFFS").reshape(n_segs, 5)
    rlo = parse_array(text, "PSWF_R_LO")
    dri = parse_array(text, "PSWF_DR_INV")
    rng = np.random.default_rng(2026)
    segs = sorted(set([0, 1, 2, 100, 1000, 2047, 2048, 3000, 4094, 4095] + list(rng.integers(0, n_segs, 22))))
    rs = []
    for s in segs:
        for t in (0.0, 0.37, 0.5, 0.999):
            rs.append((rlo[s] + t / dri[s]) * rcut)
    rs = [r for r in rs if r >= 1e-12] + [5e-13, 1.0, 1.0 - 1e-12, 1.5, 2.44140625e-04]
    out = []
    for r in rs:
        if r < 1e-12 or r > rcut:
            out.append(dict(r=r, S=0.0, dS=0.0, L=0.0, negdLdr=0.0))
            continue
        s = min(max(r / rcut, 0.0), 1.0 - 1e-9)
        seg = min(int(s * n_segs), n_segs - 1)
        t = (r - rlo[seg]) * dri[seg]
        S = float(np.polyval(scr[seg][::-1], t))
        dS = float(np.polyval(dscr[seg][::-1], t))
        L = S * INV_4PI / r
        nd = S * INV_4PI / r / r - dS * INV_4PI / r
        out.append(dict(r=r, seg=seg, S=S, dS=dS, L=L, negdLdr=nd))
    used = sorted(set(segs) | {o["seg"] for o in out if "seg" in o})
    fixture = dict(
        source="reference src/PSWF_Coeffs.h (epsilon=1e-4, c=9, r_cut=1, n_segs=4096, P=6)",
        n_segs=n_segs, rcut=rcut,
        segments={str(s): dict(screen=scr[s].tolist(), dscreen=dscr[s].tolist(), r_lo=rlo[s], dr_inv=dri[s]) for s in used},
        S0=float(scr[0][0]), S_end=float(np.polyval(scr[-1][::-1], 1.0)),
        lookups=out,
    )
    dst = Path(__file__).resolve().parent / "pswf_ref_lookup.json"
    dst.write_text(json.dumps(fixture, indent=1))
    print(f"wrote {dst} ({len(out)} lookups over {len(used)} segments)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/PSWF_Coeffs.h")
