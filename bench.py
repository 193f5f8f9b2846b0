#!/usr/bin/env python3
"""ESP long-range electrostatics benchmark on one MI355X.

A step is one full ESP force evaluation (pair correction over the neighbour list, tile
binning, PSWF spreading, R2C FFT, influence multiply with the FP64 energy/virial reduction,
C2R FFT, gather) of one system whose inputs are resident in HBM.  The default workload is
the 1,000,000-particle primitive-model electrolyte (q = +-1, number density 0.1, jittered
simple-cubic lattice) with mesh 128^3, P = 10, r_c = 3.5 in FP64.

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s at 1e-5 rel force err; batched solves/s + p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def lattice_electrolyte(N, rho, seed_pos=2, seed_q=3, jitter=0.1, q2_frac=0.0):
    """Seeded jittered simple-cubic lattice of N = m^3 sites at number density rho,
    box centred on the origin, neutral +-1 charges in a random permutation."""
    m = int(round(N ** (1.0 / 3.0)))
    if m ** 3 != N:
        raise ValueError("N must be a perfect cube")
    L = (N / rho) ** (1.0 / 3.0)
    a = L / m
    g = (np.arange(m) + 0.5) * a - L / 2
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    rng = np.random.default_rng(seed_pos)
    pos += rng.uniform(-jitter * a, jitter * a, pos.shape)
    q = np.where(np.arange(N) < N // 2, 1.0, -1.0)
    if N % 2:
        q[-1] = 0.0
    q = np.random.default_rng(seed_q).permutation(q)
    return pos, q, L


def kernel_bytes(stage, N, n_charged, n_pairs, M, Mh, R, want_ev=True):
    """Algorithmic HBM bytes of one launch of each stage (reads + writes of its operands,
    counted once; neighbour-position gathers excluded)."""
    rec = 4 * R + 8
    return {
        "pair": N * (32 + 8 + 32 + 48 + 12) + 4 * n_pairs,
        "bin": N * (32 + 8 + 8),
        "scan": 0,
        "scatter": N * (32 + 8 + 8) + n_charged * rec,
        "spread": n_charged * rec + M * 8,
        "fft_x_r2c": M * 8 * 2 + Mh * 2 * R,
        "fft_y": Mh * 4 * R,
        "fft_z_convolve": Mh * 4 * R + Mh * 8 * (2 if want_ev else 1),
        "ifft_y": Mh * 4 * R,
        "ifft_x_c2r": Mh * 2 * R + M * R,
        "gather": n_charged * (rec + 64) + M * R,
        "finalize": 0,
    }[stage]


def stage_kernel(stage, precision, P):
    """Device kernel name that one stage launches (as rocprofv3 prints it, template
    arguments dropped)."""
    return {
        "pair": "k_pair", "bin": "k_bin", "scan": "k_scan", "scatter": "k_scatter",
        "spread": "k_spread_own" if precision == "fp64" else "k_spread_tp",
        "fft_x_r2c": "k_fft_x_r2c", "fft_y": "k_fft_y", "fft_z_convolve": "k_fft_z_conv",
        "ifft_y": "k_fft_y", "ifft_x_c2r": "k_fft_x_c2r", "gather": "k_gather_tp", "finalize": "k_finalize",
    }[stage]


def measured_traffic(workload_key, kernel):
    """HBM bytes per launch of `kernel` for `workload_key`, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary kept in profiles/hbm_traffic.txt (lines
    'TRAFFIC <workload_key> <kernel> <bytes_per_launch>'), or None when not measured."""
    f = ROOT / "profiles" / "hbm_traffic.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        parts = line.split()
        if len(parts) == 4 and parts[0] == "TRAFFIC" and parts[1] == workload_key and parts[2] == kernel:
            return float(parts[3])
    return None


def copy_calibration(mib, steps):
    """Launch the library's plain record-copy kernel `steps` times over `mib` MiB (64 bytes of
    HBM traffic per 32-byte record): the known-size kernel that scales the FETCH_SIZE and
    WRITE_SIZE counters to bytes."""
    from esp_hoomd_amd import _lib
    from esp_hoomd_amd.device import DeviceArray, Stream
    n_rec = int(mib * 2 ** 20) // 32
    src = DeviceArray((n_rec, 4), np.float64, zero=True)
    dst = DeviceArray((n_rec, 4), np.float64)
    stream = Stream()
    for _ in range(steps):
        _lib.check(_lib.load().esp_copy_records(dst._ptr, src._ptr, n_rec, stream.handle))
    stream.synchronize()
    print(json.dumps({"copy_calibration": {"kernel": "k_copy4", "records": n_rec, "read_bytes": n_rec * 32,
                                           "write_bytes": n_rec * 32, "launches": steps}}))
    return 0


def cpu_baseline(h, rc, P, kappa, precision_name, n_mesh=48, reps=2):
    """The oracle (FP64 CPU ESP) on a same-density sample with the same h, r_c, P."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from esp_hoomd_amd import solver as S

    L = n_mesh * h
    m = int(round((0.1 * L ** 3) ** (1.0 / 3.0)))
    N = m ** 3
    pos, q, _ = lattice_electrolyte(N, N / L ** 3)
    sigma = np.pi / (h * kappa)
    D, cw = S.default_window(P, sigma)
    w = S.build_window(P, D, cw)
    scr, dscr, rlo, dri = S.build_split_table(0, kappa, rc, 4096)
    o = O.OracleESP((L, L, L), (n_mesh,) * 3, P, w, scr, dscr, rlo, dri, rc, k_cut=1.2 * kappa)
    o.compute(pos[:64], q[:64])  # warm the library
    t0 = time.perf_counter()
    for _ in range(reps):
        o.compute(pos, q)
    dt = (time.perf_counter() - t0) / reps
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": N / dt,
        "unit": "particle-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"oracle FP64 ESP (brute-force pair) on N={N} at density 0.1, mesh {n_mesh}^3 (h={h:.4f}), "
                  f"P={P}, r_c={rc}, mean of {reps} evaluations ({dt:.2f} s each); GPU run is {precision_name}",
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--N", type=int, default=1_000_000)
    ap.add_argument("--mesh", type=int, default=128)
    ap.add_argument("--order", type=int, default=10)
    ap.add_argument("--rcut", type=float, default=3.5)
    ap.add_argument("--kappa", type=float, default=0.0, help="0: sigma=1.3 oversampling of the band limit")
    ap.add_argument("--skin", type=float, default=0.4)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="MIB",
                    help="only run the record-copy kernel over MIB MiB, --steps times (counter calibration)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py: only single-GPU runs are supported (got --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not part of this build", file=sys.stderr)
        return 2

    if args.calibrate_copy > 0:
        return copy_calibration(args.calibrate_copy, args.steps)

    from esp_hoomd_amd import _lib
    from esp_hoomd_amd.device import Event, Stream, synchronize
    from esp_hoomd_amd.solver import DeviceSystem, Params, Solver

    rho = 0.1
    pos, q, L = lattice_electrolyte(args.N, rho)
    h = L / args.mesh
    kappa = args.kappa if args.kappa > 0 else np.pi / (1.3 * h)
    prec = _lib.ESP_PREC_FP64 if args.precision == "fp64" else _lib.ESP_PREC_MIXED
    params = Params(nx=args.mesh, ny=args.mesh, nz=args.mesh, order=args.order, kappa=kappa, rcut=args.rcut,
                    precision=prec)
    t0 = time.perf_counter()
    sysd = DeviceSystem(pos, q, (L, L, L), r_list=args.rcut + args.skin)
    n_pairs = int(sysd.n_neigh.numpy().astype(np.int64).sum())
    setup_s = time.perf_counter() - t0
    solver = Solver(params)
    stream = Stream()
    for _ in range(args.warmup):
        solver.enqueue(sysd, stream=stream)
    stream.synchronize()

    solver.set_profiling(True)
    solver.reset_stage_times()
    e0, e1 = Event(), Event()
    synchronize()
    e0.record(stream)
    for k in range(args.steps):
        solver.enqueue(sysd, timestep=k, stream=stream)
    e1.record(stream)
    stream.synchronize()
    synchronize()
    ms_total = e0.elapsed_ms(e1)
    stages = solver.stage_times()
    solver.set_profiling(False)

    ms_per_step = ms_total / args.steps
    value = args.N / (ms_per_step * 1e-3)
    R = 8 if prec == _lib.ESP_PREC_FP64 else 4
    M = args.mesh ** 3
    Mh = (args.mesh // 2 + 1) * args.mesh * args.mesh
    n_charged = int(np.count_nonzero(q))
    avg = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in stages.items()}
    dom = max(avg, key=avg.get)
    b = kernel_bytes(dom, args.N, n_charged, n_pairs, M, Mh, R)
    achieved = b / (avg[dom] * 1e-3) / 1e9 if avg[dom] > 0 else 0.0
    workload_key = f"N{args.N}_mesh{args.mesh}_P{args.order}_rc{args.rcut}_{args.precision}"
    dom_kernel = stage_kernel(dom, args.precision, args.order)
    traffic = measured_traffic(workload_key, dom_kernel)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == _lib.ESP_PREC_FP64 else "f32 (FP64 influence/reductions)",
        "data": "synthetic",
        "config": {
            "workload": f"plain ESP, 1 system: N={args.N} primitive-model electrolyte (q=+-1, density 0.1, "
                        f"jittered simple-cubic lattice, not MD-equilibrated), L={L:.3f}, mesh {args.mesh}^3, "
                        f"P={args.order}, r_c={args.rcut}, kappa={kappa:.4f} (c={kappa * args.rcut:.3f}), "
                        f"skin {args.skin}",
            "precision": args.precision,
            "neighbours_per_particle": n_pairs / args.N,
        },
        "roofline": {
            "kernel": f"{dom} ({dom_kernel})",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes": b,
            "avg_launch_ms": avg[dom],
        },
        "stages_ms": {k: round(v, 5) for k, v in avg.items()},
        "setup_s": setup_s,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(h, args.rcut, args.order, kappa, args.precision)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
