#!/usr/bin/env python3
"""ESP long-range electrostatics benchmark on one MI355X.

Metric: particle-steps/s at 1e-5 relative RMS force error; batched solves/s and served
p50/p99 step latency beside it.  A step is one full ESP force evaluation (pair correction over
the neighbour list, tile binning, PSWF spreading, R2C FFT, influence multiply with the FP64
energy/virial reduction, C2R FFT, gather) of a system whose inputs are resident in HBM.

The 1,000,000-particle electrolyte (density 0.1, +-1) starts from a jittered simple-cubic
lattice and is equilibrated by 1,000 Langevin steps (T = 1, gamma = 1, seed 4) on the GPU before
anything is measured.  Legs (each reports its measured force error against the tight FP64
reference of esp_hoomd_amd.accuracy, computed on the GPU for the same positions; the test
suite ties that reference to the direct Ewald sum at this size):

* literal: the parameters of the reference's GPU configuration (mesh 128^3, P = 10, r_c = 3.5)
  in FP64;
* fp64 / mixed: parameters autotuned (and error-verified) for 1e-5, in FP64 and in mixed
  precision -- the mixed leg is the headline `value`;
* md: NVE MD steps with the headline configuration plus WCA (integration, neighbour-list
  upkeep and rebuilds included), wall-clock timed;
* shuffled: the headline configuration with the particles in a random order;
* batched: 512 independent electrolytes, N log-uniform in [500, 50,000], +-1 with 10 % +-2,
  density 0.1, per-system parameters from the autotuner (meshes 16^3-64^3, P in {6, 8, 10}),
  mixed precision, one launch per stage replayed from a captured hipGraph;
* served: a Poisson stream of MD requests (N log-uniform in [1e3, 2e5], density U[0.05, 0.2],
  10-1000 steps) through the continuous-batching server: first offered faster than it can serve
  (12 s; its capacity in particle-steps/s over the ticks after the first 3 s, in which it fills
  up from empty), then offered at 70 % of that capacity for a fixed
  wall time (latencies, queue delays); every finished request's force error is measured
  afterwards.

`cpu_baseline` times the oracle (FP64 CPU ESP on OpenMP) on the same equilibrated system with
the headline parameters.

`roofline` names the dominant kernel (from a per-stage HIP-event pass outside the timed region)
and divides its algorithmic bytes by its average launch time; for the pair kernel that time
comes from HIP events on the kernel's own stream inside the timed region, which carries no
other markers.

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s at 1e-5 rel force err; batched solves/s + p50/p99 latency"
HBM_PEAK_GBS = 8000.0
TARGET = 1e-5


def progress(msg):
    """One line per finished piece of work on stderr (the JSON result goes to stdout)."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def kernel_bytes(stage, N, n_charged, n_pairs, M, Mh, R, want_ev=True):
    """Algorithmic HBM bytes of one launch of each stage (reads + writes of its operands,
    counted once; neighbour-position gathers excluded)."""
    rec = 4 * R + 8
    return {
        "pair": N * (4 * R + 8 + 32 + 48 + 12) + 4 * n_pairs,  # own record: 32 B FP64, 16 B mixed
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


def stage_kernel(stage, precision):
    """Device kernel that one stage launches (as rocprofv3 names it, template arguments dropped)."""
    return {
        "pair": "k_pair", "bin": "k_bin", "scan": "k_scan", "scatter": "k_scatter",
        "spread": "k_spread_tp",
        "fft_x_r2c": "k_fft_x_r2c", "fft_y": "k_fft_y", "fft_z_convolve": "k_fft_z_conv",
        "ifft_y": "k_fft_y", "ifft_x_c2r": "k_fft_x_c2r", "gather": "k_gather_tp", "finalize": "k_finalize",
    }[stage]


def workload_key(N, params, precision):
    return f"N{N}_mesh{params.nx}_P{params.order}_{precision}"


def measured_traffic(key, kernel):
    """HBM bytes per launch of `kernel` for workload `key`, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary kept in profiles/hbm_traffic.txt (lines
    'TRAFFIC <workload_key> <kernel> <bytes_per_launch>'), or None when not measured."""
    f = ROOT / "profiles" / "hbm_traffic.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        parts = line.split()
        if len(parts) == 4 and parts[0] == "TRAFFIC" and parts[1] == key and parts[2] == kernel:
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


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pos, q, L, p, reps=1):
    """The oracle (FP64 CPU ESP: spreading, FFTs, influence, gather and a cell-list pair sum, its
    parallel loops on OpenMP over the host cores) evaluating the same system with the same mesh,
    P, r_c and kappa as the headline configuration."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from esp_hoomd_amd import solver as S

    N = len(pos)
    h = L / p.nx
    sigma = np.pi / (h * p.kappa)
    D, cw = S.default_window(p.order, sigma)
    w = S.build_window(p.order, D, cw)
    scr, dscr, rlo, dri = S.build_split_table(0, p.kappa, p.rcut, 4096)
    o = O.OracleESP((L, L, L), (p.nx, p.ny, p.nz), p.order, w, scr, dscr, rlo, dri, p.rcut, k_cut=1.2 * p.kappa)
    o.compute(pos[:64], q[:64])  # load the library
    t0 = time.perf_counter()
    for _ in range(reps):
        o.compute(pos, q)
    dt = (time.perf_counter() - t0) / reps
    threads = O.num_threads()
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {
        "value": N / dt,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "nproc": affinity,
        "cpu_model": _cpu_model(),
        "sample": f"the oracle (C, FP64, OpenMP on {threads} threads; cell-list pair sum, serial spreading) "
                  f"evaluating the whole N={N} system once per step with the headline parameters (mesh {p.nx}^3, "
                  f"P={p.order}, r_c={p.rcut:.4g}, kappa={p.kappa:.4g}); mean of {reps} evaluation(s), {dt:.2f} s each",
    }


# ---------------------------------------------------------------------------------------
def _n_pairs(state):
    import ctypes as C
    from esp_hoomd_amd import _lib
    n = np.empty(state.N, np.uint32)
    _lib.check(_lib.load().esp_memcpy_dtoh(n.ctypes.data_as(C.c_void_p), C.c_void_p(state.nlist_ptrs[0]), n.nbytes))
    return int(n.astype(np.int64).sum())


def run_single(params, pos, q, L, steps, warmup, ref, profile=True, skin=0.4):
    """Time `steps` evaluations of one resident system; error vs ref.  With `profile`, the timed
    region carries HIP events around the pair kernel only (on its own stream, so the step is
    not slowed by markers), and a separate pass of min(steps, 20) evaluations with events at
    every stage boundary gives the per-stage breakdown."""
    from esp_hoomd_amd import _lib
    from esp_hoomd_amd.accuracy import rel_rms_error
    from esp_hoomd_amd.device import Event, Stream, synchronize
    from esp_hoomd_amd.md import NeighborList, State
    from esp_hoomd_amd.solver import Solver

    st = State(pos, q, [L, L, L])
    nl = NeighborList(buffer=skin)
    nl.update(st, params.rcut, force_rebuild=True)
    solver = Solver(params)
    stream = Stream()
    for _ in range(warmup):
        solver.enqueue_state(st, stream=stream)
    stream.synchronize()
    err = rel_rms_error(st.force.numpy(), ref)
    solver.set_profiling(2 if profile else 0)
    solver.reset_stage_times()
    e0, e1 = Event(), Event()
    synchronize()
    e0.record(stream)
    for k in range(steps):
        solver.enqueue_state(st, timestep=k, stream=stream)
    e1.record(stream)
    stream.synchronize()
    synchronize()
    ms = e0.elapsed_ms(e1) / steps
    timed = solver.stage_times() if profile else {}
    avg = {}
    if profile:
        solver.set_profiling(1)
        solver.reset_stage_times()
        for k in range(min(steps, 20)):
            solver.enqueue_state(st, timestep=k, stream=stream)
        stream.synchronize()
        avg = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in solver.stage_times().items()}
    solver.set_profiling(0)
    pair_timed = timed.get("pair", (0.0, 0))
    return {"ms_per_step": ms, "particle_steps_per_s": len(pos) / (ms * 1e-3), "error": err,
            "stages_ms": avg, "pair_ms_timed": pair_timed[0] / pair_timed[1] if pair_timed[1] else 0.0,
            "n_pairs": _n_pairs(st), "tile": int(solver.info().tile)}


def md_leg(params, pos, vel, q, L, steps, warmup, skin=0.4):
    """NVE velocity-Verlet MD of the resident system with this ESP configuration plus WCA
    (sigma = epsilon = 1), dt = 0.002: per step the first half kick and drift, the neighbour
    list's displacement check (one host synchronisation) and its rebuilds when a particle has
    moved half the skin, the ESP forces, WCA over the inner list and the second half kick; every
    100 steps the particles are re-sorted into cell order (and the list rebuilt).
    Timed by wall clock over `steps` steps after `warmup`."""
    from esp_hoomd_amd.device import synchronize
    from esp_hoomd_amd.md import ESPForce, NeighborList, Simulation, State, WCA
    from esp_hoomd_amd.solver import Solver

    st = State(pos, q, [L, L, L], velocities=vel)
    # particles put back in cell order every 100 steps (HOOMD's SFCPack keeps its particles
    # sorted the same way; without it they diffuse away from their lattice order and the pair,
    # binning and spreading passes lose their memory locality step by step)
    sim = Simulation(st, dt=0.002, nlist=NeighborList(skin), sort_every=100)
    sim.add_force(ESPForce(Solver(params)))
    sim.add_force(WCA(1.0, 1.0))
    sim.run(warmup)
    ke0, pe0 = sim.energies()
    b0 = sim.nlist.builds
    synchronize()
    t0 = time.perf_counter()
    sim.run(steps)
    dt = time.perf_counter() - t0
    ke, pe = sim.energies()
    return {"steps": steps, "ms_per_step": dt / steps * 1e3, "particle_steps_per_s": len(pos) * steps / dt,
            "list_builds": sim.nlist.builds - b0, "sorts": sim.sorts, "energy_change_per_KE0": abs(ke + pe - ke0 - pe0) / ke0,
            "temperature": sim.temperature()}


# LDS peak: 32 banks x 4 bytes per clock per CU (the CDNA LDS width), 256 CUs at 2.4 GHz
LDS_PEAK_TBS = 128 * 256 * 2.4e9 / 1e12


def measured_lds(key, kernel):
    """(lds_util, conflict) of `kernel` for workload `key` from the rocprofv3 counter summary
    profiles/esp_step_N1M_lds_valu_counters.txt (lines 'LDSUTIL <key> <kernel> <util> <conflict>')."""
    f = ROOT / "profiles" / "esp_step_N1M_lds_valu_counters.txt"
    if not f.exists():
        return None, None
    for line in f.read_text().splitlines():
        parts = line.split()
        if len(parts) == 5 and parts[0] == "LDSUTIL" and parts[1] == key and parts[2] == kernel:
            return float(parts[3]), float(parts[4])
    return None, None


def lds_roofline(res, params, N, n_charged, precision):
    """LDS side of the spreading and gathering kernels, which are bound by LDS, not HBM: the
    algorithmic LDS bytes of one launch -- spread: clearing, accumulating into (one atomic per
    stencil tap: n_charged P^3) and reading out the tile boxes; gather: filling the boxes from
    the potential mesh and reading the P^3 taps of every particle -- over the stage's average
    duration, against the LDS peak; beside it the counter-measured LDS utilisation and bank
    conflicts of the same kernels (rocprofv3, committed summary)."""
    P, T = params.order, res.get("tile", 8) or 8
    el = 8 if precision == "fp64" else 4
    B = T + P - 1
    tiles = math.ceil(params.nx / T) * math.ceil(params.ny / T) * math.ceil(params.nz / T)
    box = tiles * B ** 3 * el
    bytes_ = {"spread": n_charged * P ** 3 * el + 2 * box, "gather": n_charged * P ** 3 * el + box}
    key = workload_key(N, params, precision)
    out = {"peak": LDS_PEAK_TBS, "unit": "TB/s", "peak_basis": "128 B/clock/CU x 256 CUs x 2.4 GHz"}
    for stage in ("spread", "gather"):
        t = res["stages_ms"].get(stage, 0.0)
        achieved = bytes_[stage] / (t * 1e-3) / 1e12 if t > 0 else 0.0
        util, conflict = measured_lds(key, stage_kernel(stage, precision))
        out[stage] = {"kernel": stage_kernel(stage, precision), "algorithmic_bytes": bytes_[stage], "avg_launch_ms": t,
                      "achieved": achieved, "frac": achieved / LDS_PEAK_TBS, "measured_lds_util": util,
                      "measured_bank_conflict": conflict}
    return out


def roofline(res, params, N, n_charged, precision):
    """Dominant kernel by the per-stage pass; its launch time from the timed region's events
    when it is the pair kernel (the one the timed region marks)."""
    avg = dict(res["stages_ms"])
    dom = max(avg, key=avg.get)
    if dom == "pair" and res.get("pair_ms_timed", 0.0) > 0:
        avg["pair"] = res["pair_ms_timed"]
    R = 8 if precision == "fp64" else 4
    M = params.nx * params.ny * params.nz
    Mh = (params.nx // 2 + 1) * params.ny * params.nz
    b = kernel_bytes(dom, N, n_charged, res["n_pairs"], M, Mh, R)
    achieved = b / (avg[dom] * 1e-3) / 1e9 if avg[dom] > 0 else 0.0
    kern = stage_kernel(dom, precision)
    return {
        "kernel": f"{dom} ({kern})",
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": measured_traffic(workload_key(N, params, precision), kern),
        "algorithmic_bytes": b,
        "avg_launch_ms": avg[dom],
        "lds": lds_roofline(res, params, N, n_charged, precision),
    }


def batched_leg(n_sys, steps, warmup, seed=6):
    """512 independent electrolytes (or n_sys), tuned per system, one batched launch per stage."""
    from esp_hoomd_amd import _lib, accuracy, tune, workloads as W
    from esp_hoomd_amd.batch import Batch
    from esp_hoomd_amd.device import Event, Stream, synchronize
    from esp_hoomd_amd.md import NeighborList, State
    from esp_hoomd_amd.solver import Params, Solver

    rng = np.random.default_rng(seed)
    Ns = W.batch_sizes(n_sys, seed)
    rng.uniform(size=n_sys)  # the sizes' draws (keeps the error sample below as before)
    systems = W.batch_systems(n_sys, seed)
    # the model is fitted on random electrolytes; rescale it to these lattice systems with
    # measured errors on three of them (smallest, median, largest), then tune every system
    order = np.argsort(Ns)
    pilots = [order[0], order[len(order) // 2], order[-1]]
    scale = 1.0
    for i in pilots:
        pos, q, L = systems[i]
        ch = tune.autotune(pos, q, (L, L, L), TARGET, _lib.ESP_PREC_MIXED, orders=(6, 8, 10), mesh_range=(16, 64),
                           cost=tune.CostModel(rebuild_every=0.0))
        scale = max(scale, ch.method_scale)
    t0 = time.perf_counter()
    pairs, states, choices = [], [], []
    for pos, q, L in systems:
        ch = tune.tune(len(pos), (L, L, L), TARGET, _lib.ESP_PREC_MIXED, orders=(6, 8, 10), mesh_range=(16, 64),
                       method_scale=scale, safety=1.1, batched=True)[0]
        st = State(pos, q, [L, L, L])
        nl = NeighborList(buffer=0.4)
        nl.update(st, ch.params.rcut, force_rebuild=True)
        st._nl = nl  # keep the list alive with its state
        pairs.append((Solver(ch.params), st))
        choices.append(ch)
    setup_s = time.perf_counter() - t0
    batch = Batch(use_graph=True)
    stream = Stream()
    for _ in range(warmup):
        batch.enqueue(pairs, stream=stream)
    stream.synchronize()
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(steps):
        batch.enqueue(pairs, stream=stream)
    e1.record(stream)
    stream.synchronize()
    ms = e0.elapsed_ms(e1) / steps

    def sequential(precision):
        """The same systems one after another through the plain single-system path (same
        ESP parameters, the library's automatic tile for single systems)."""
        solos = [Solver(Params(**{**ch.params.__dict__, "tile": 0, "precision": precision})) for ch in choices]
        # one untimed call each: initialisation (tables, influence, buffers) is not timed
        for s, (_, st) in zip(solos, pairs):
            s.enqueue_state(st, stream=stream)
        stream.synchronize()
        e0.record(stream)
        for s, (_, st) in zip(solos, pairs):
            s.enqueue_state(st, stream=stream)
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1)

    seq_mixed_ms = sequential(_lib.ESP_PREC_MIXED)
    seq_fp64_ms = sequential(_lib.ESP_PREC_FP64)
    # measured errors on a sample of systems (batched forces against the tight reference)
    batch.enqueue(pairs, stream=stream)
    stream.synchronize()
    sample = list(rng.choice(n_sys, size=min(8, n_sys), replace=False))
    errs = []
    for i in sample:
        pos, q, L = systems[i]
        ref = accuracy.reference_forces(pos, q, (L, L, L))
        errs.append(accuracy.rel_rms_error(pairs[i][1].force.numpy(), ref))
    total_N = int(Ns.sum())
    meshes = sorted({c.params.nx for c in choices})
    orders = {int(P): int(sum(1 for c in choices if c.params.order == P)) for P in (6, 8, 10)}
    return {
        "systems": n_sys,
        "total_particles": total_N,
        "ms_per_batch_step": ms,
        "solves_per_s": n_sys / (ms * 1e-3),
        "particle_steps_per_s": total_N / (ms * 1e-3),
        "sequential_single_mixed_ms": seq_mixed_ms,
        "sequential_single_fp64_ms": seq_fp64_ms,
        "batching_speedup_vs_single_mixed": seq_mixed_ms / ms,
        "speedup_vs_single_fp64": seq_fp64_ms / ms,
        "error_sample_max": float(max(errs)),
        "error_sample_median": float(np.median(errs)),
        "meshes": [meshes[0], meshes[-1]],
        "orders": orders,
        "model_scale": scale,
        "setup_s": setup_s,
        "graph": batch.stats(),
    }


MEAN_REQUEST_WORK = (2e5 - 1e3) / math.log(200) * (1000 - 10) / math.log(100)  # E[N] E[steps] of the stream


def _serve_run(scales, seconds, rate, seed, budget_gb, verify, steady_from=None):
    from esp_hoomd_amd import _lib
    from esp_hoomd_amd.serve import Server, poisson_requests
    reqs = poisson_requests(int(rate * seconds * 1.5) + 8, rate=rate, seed=seed)
    srv = Server(hbm_budget=int(budget_gb * (1 << 30)), max_active=128, precision=_lib.ESP_PREC_MIXED, target=TARGET,
                 keep_final_positions=verify)
    srv.method_scales = dict(scales)
    for r in reqs:
        srv.submit(r)
    st = srv.run(time_limit=seconds)
    st["offered_rate_per_s"] = rate
    if steady_from is not None:
        st["steady"] = srv.rate_after(steady_from, seconds)
    if verify:
        # every finished request's force error at its final positions, against the tight reference
        t0 = time.perf_counter()
        errs = np.asarray(srv.verify_results()) if srv.done else np.zeros(1)
        st["force_error"] = {"verified_results": len(srv.done), "p50": float(np.median(errs)),
                             "p99": float(np.percentile(errs, 99)), "max": float(errs.max()),
                             "over_target": int((errs > TARGET).sum()), "verify_s": time.perf_counter() - t0}
    return st


def served_leg(seconds, rate=None, overload_rate=100.0, capacity_seconds=12.0, ramp_seconds=3.0, seed=7, budget_gb=64):
    """Two phases on the same request generator: the server's capacity (a stream offered faster
    than it can serve for `capacity_seconds`: the particle-steps/s it sustains once full, i.e.
    over the ticks after the first `ramp_seconds`, in which it fills up from empty), then the
    served stream at `rate` -- by default 70 % of the measured capacity -- for `seconds`, with
    latencies, queue delays and every finished result's force error."""
    from esp_hoomd_amd import _lib
    from esp_hoomd_amd.serve import Server
    t0 = time.perf_counter()
    scales = Server(precision=_lib.ESP_PREC_MIXED, target=TARGET).calibrate()
    calib_s = time.perf_counter() - t0
    cap = _serve_run(scales, capacity_seconds, overload_rate, seed + 1, budget_gb, verify=False, steady_from=ramp_seconds)
    steady = cap["steady"]
    capacity = {"particle_steps_per_s": steady["particle_steps_per_s"], "solves_per_s": steady["solves_per_s"],
                "mean_active": steady["mean_active"], "ticks": steady["ticks"],
                "seconds": capacity_seconds, "ramp_excluded_s": ramp_seconds,
                "whole_run_particle_steps_per_s": cap["particle_steps_per_s"],
                "requests_per_s": cap["requests_per_s"], "offered_rate_per_s": overload_rate,
                "queue_delay_s_p50": cap["queue_delay_s_p50"]}
    if rate is None:
        rate = 0.7 * cap["particle_steps_per_s"] / MEAN_REQUEST_WORK
    st = _serve_run(scales, seconds, rate, seed, budget_gb, verify=True)
    st["capacity"] = capacity
    st["hbm_budget_bytes"] = int(budget_gb * (1 << 30))
    st["calibration"] = {"method_scale_by_density": scales, "seconds": calib_s}
    return st


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--N", type=int, default=1_000_000)
    ap.add_argument("--legs", default="literal,fp64,mixed,md,shuffled,batched,served",
                    help="comma list of literal, fp64, mixed, md, shuffled, batched, served")
    ap.add_argument("--equilibrate", type=int, default=1000,
                    help="Langevin steps (T = 1, gamma = 1, seed 4) before anything is measured")
    ap.add_argument("--md-steps", type=int, default=1000, help="timed NVE steps of the md leg (the 1M-particle configuration's 1,000)")
    ap.add_argument("--batch-systems", type=int, default=512)
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="offered requests/s of the served leg (default: 70 %% of the server's measured capacity)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="MIB",
                    help="only run the record-copy kernel over MIB MiB, --steps times (counter calibration)")
    ap.add_argument("--only", choices=["literal", "fp64", "mixed"], default=None,
                    help="run one single-system leg with fixed parameters and no tuning (for profilers)")
    ap.add_argument("--params", default=None, help="mesh,P,rc,kappa for --only (default: the literal ones)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py: only single-GPU runs are supported (got --gpus {args.gpus}); "
              "the library runs one system or batch per GPU and has no multi-GPU sharding", file=sys.stderr)
        return 2
    if args.calibrate_copy > 0:
        return copy_calibration(args.calibrate_copy, args.steps)

    from esp_hoomd_amd import _lib, accuracy, tune, workloads as W
    from esp_hoomd_amd.md import State
    from esp_hoomd_amd.solver import Params

    legs = set(args.legs.split(","))
    t_all = time.perf_counter()
    pos, q, L = W.lattice_electrolyte(args.N, 0.1)
    box = (L, L, L)
    n_charged = int(np.count_nonzero(q))
    literal = Params(nx=128, ny=128, nz=128, order=10, kappa=np.pi / (1.3 * L / 128), rcut=3.5,
                     precision=_lib.ESP_PREC_FP64)
    vel = None
    eq_s = 0.0
    if args.equilibrate > 0:
        st = State(pos, q, list(box))
        t0 = time.perf_counter()
        W.equilibrate(st, literal, steps=args.equilibrate)
        eq_s = time.perf_counter() - t0
        progress(f"equilibrated {args.equilibrate} steps in {eq_s:.1f} s")
        pos, vel = st.positions().copy(), st.velocities().copy()
        del st

    if args.only:
        prec = _lib.ESP_PREC_FP64 if args.only in ("literal", "fp64") else _lib.ESP_PREC_MIXED
        if args.params:
            n, P, rc, kappa = args.params.split(",")
            p = Params(nx=int(n), ny=int(n), nz=int(n), order=int(P), kappa=float(kappa), rcut=float(rc), precision=prec)
        else:
            p = Params(nx=128, ny=128, nz=128, order=10, kappa=literal.kappa, rcut=3.5, precision=prec)
        res = run_single(p, pos, q, L, args.steps, args.warmup, np.zeros((args.N, 3)))
        pname = "fp64" if prec == _lib.ESP_PREC_FP64 else "mixed"
        print(json.dumps({"leg": args.only, "params": [p.nx, p.order, p.rcut, p.kappa], "ms_per_step": res["ms_per_step"],
                          "stages_ms": res["stages_ms"], "roofline": roofline(res, p, args.N, n_charged, pname)}))
        return 0

    ref = accuracy.reference_forces(pos, q, box)
    out_legs = {}
    if "literal" in legs:
        r = run_single(literal, pos, q, L, args.steps, args.warmup, ref)
        out_legs["literal"] = {"params": {"mesh": 128, "order": 10, "rcut": 3.5, "kappa": literal.kappa,
                                          "c": literal.kappa * 3.5}, **{k: r[k] for k in ("ms_per_step",
                                          "particle_steps_per_s", "error", "stages_ms")}}
        progress(f"literal leg {r['ms_per_step']:.3f} ms/step")
    tuned = {}
    for name, prec in (("fp64", _lib.ESP_PREC_FP64), ("mixed", _lib.ESP_PREC_MIXED)):
        if name not in legs:
            continue
        ch = tune.autotune(pos, q, box, TARGET, prec, measure_time=True, top=8, reference=ref)
        r = run_single(ch.params, pos, q, L, args.steps, args.warmup, ref)
        tuned[name] = (ch, r)
        out_legs[name] = {"params": ch.describe(), **{k: r[k] for k in ("ms_per_step", "particle_steps_per_s", "error",
                                                                       "stages_ms")}}
        progress(f"{name} leg {r['ms_per_step']:.3f} ms/step, error {r['error']:.3g}, {ch.describe()}")
    head = "mixed" if "mixed" in tuned else ("fp64" if "fp64" in tuned else None)
    if head is not None and "md" in legs:
        out_legs["md"] = {"precision": head, **md_leg(tuned[head][0].params, pos, vel, q, L, args.md_steps, 20)}
        progress("md leg done")
    if head is not None and "shuffled" in legs:
        # the same system with its particles in a random order (no spatial locality of the
        # neighbour records; HOOMD itself keeps particles sorted along a space-filling curve)
        perm = np.random.default_rng(11).permutation(len(pos))
        r = run_single(tuned[head][0].params, pos[perm], q[perm], L, args.steps, args.warmup, ref[perm], profile=False)
        out_legs["shuffled"] = {"precision": head, **{k: r[k] for k in ("ms_per_step", "particle_steps_per_s", "error")}}
        progress(f"shuffled leg {r['ms_per_step']:.3f} ms/step")
    if "batched" in legs:
        out_legs["batched"] = batched_leg(args.batch_systems, max(args.steps // 5, 5), 3)
        progress("batched leg done")
    if "served" in legs:
        # the server's capacity first, then the stream offered at 70 % of it (or --serve-rate)
        out_legs["served"] = served_leg(args.serve_seconds, rate=args.serve_rate if args.serve_rate > 0 else None)
        progress("served leg done")

    if head is not None:
        ch, r = tuned[head]
        value, ms, params, dtype = r["particle_steps_per_s"], r["ms_per_step"], ch.params, head
    else:
        r = None
        value, ms, params, dtype = None, None, literal, "fp64"
    state_desc = (f"equilibrated {args.equilibrate} Langevin steps (T=1, gamma=1, dt=0.002) from a jittered simple-cubic "
                  f"lattice" if args.equilibrate > 0 else "jittered simple-cubic lattice")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 influence/reductions)" if dtype == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"1 system: N={args.N} primitive-model electrolyte (q=+-1, density 0.1), {state_desc}, "
                        f"L={L:.3f}; a step is one full ESP force evaluation (pair correction + mesh + energy/virial) "
                        f"of resident inputs; {head or 'literal'} precision, ESP parameters autotuned to a measured "
                        f"relative RMS force error <= {TARGET:g} against the tight FP64 reference: "
                        f"mesh {params.nx}^3, P={params.order}, r_c={params.rcut:.4g}, kappa={params.kappa:.4g}",
            "force_error": r["error"] if r else None,
            "target_error": TARGET,
            "equilibration_s": eq_s,
        },
        "legs": out_legs,
    }
    if r is not None:
        out["roofline"] = roofline(r, params, args.N, n_charged, head)
        out["stages_ms"] = {k: round(v, 5) for k, v in r["stages_ms"].items()}
    if "fp64" in tuned and "mixed" in tuned:
        f64 = tuned["fp64"][1]["particle_steps_per_s"]
        out["speedup_vs_fp64_single"] = {
            "mixed_single": tuned["mixed"][1]["particle_steps_per_s"] / f64,
            "mixed_batched": out_legs.get("batched", {}).get("particle_steps_per_s", 0.0) / f64,
            "mixed_served": out_legs.get("served", {}).get("capacity", {}).get("particle_steps_per_s", 0.0) / f64,
            "mixed_served_at_offered_load": out_legs.get("served", {}).get("particle_steps_per_s", 0.0) / f64,
        }
    if not args.no_cpu_baseline and head is not None:
        out["cpu_baseline"] = cpu_baseline(pos, q, L, params)
    else:
        out["cpu_baseline"] = None
    out["bench_wall_s"] = time.perf_counter() - t_all
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
