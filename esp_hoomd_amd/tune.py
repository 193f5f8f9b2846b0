"""Autotuner: ESP parameters (mesh, stencil order P, cutoff r_c, splitting kappa) at a force-error
target, for one system or for each system of a batch.

Accuracy model.  For the prolate split the relative RMS force error of a neutral electrolyte
is, to good approximation, sqrt(E_split(c)^2 + E_window(P, sigma)^2 [+ E_fp32^2]) where
c = kappa * r_c is the prolate bandwidth and sigma = pi / (h kappa) the oversampling of the
band limit by the mesh (h = mesh spacing):

* E_split(c) = 10^(-0.416 c - 0.551), fitted to the oracle's error against a direct Ewald sum
  for random +-1 electrolytes at number density 0.1 (c = 6 ... 24, best P and sigma);
* E_window(P, sigma) is the table below (the excess over E_split at fixed c, maximum over
  c = 8 ... 20 of the same measurements), interpolated log-linearly in sigma and P;
* E_fp32 = 1.5e-6 in mixed precision (FP32 spread/FFT/gather/pair against FP64 at equal
  parameters: 2e-7 on 1e3-1e4-particle systems, 1.3e-6 to 2.3e-6 on the 1M-particle lattice);
* the default reciprocal-space cutoff |k| <= 1.2 kappa raises the method error by up to
  ~1.3x (measured on the same systems), so the first two terms carry a factor CUT_FACTOR.

Cost model.  Per-step time on one MI355X, fitted to mesh-only, pair-only and full steps of
the 1M-particle workload (P = 6-12, meshes 128^3-216^3, r_c 4-8): the pair pass alone costs
`pair_us` plus `pair_ps` per full-list neighbour entry; the mesh chain costs `overhead_us`,
`bin_ns` per particle for binning, `stencil_ps * P^3` per particle for spread + gather, and per
mesh cell `cell_ps` for the FFT passes plus `halo_ps` times the tile-box volume ratio
((T + P - 1) / T)^3 for the halos of the spreading tiles.  The pair pass runs beside the
mesh chain, which hides min(hide_pair * pair, hide_mesh * mesh) of it.  The mesh-chain fit
is within 7% (mixed) and 9% (FP64) RMS, the full step within about 5% near the optimum.  A
simulation also rebuilds its neighbour list (out to r_c + skin) every `rebuild_every` steps
on average: the cell-list build costs `nlist_us` plus `nlist_ps` per list entry (k_nlist at
1e6 particles: 3.0 ms for 1.78e8 entries), and its share per step is added to the step (the
1e6-particle NVE run at T = 1, dt = 0.002, skin 0.4 rebuilds every ~19 steps; 0 leaves the
list out, for force evaluations on a fixed list).  The tuner picks, over FFT-friendly meshes
and the allowed orders, the smallest r_c that meets the target (minimising sigma * c at fixed
h) and then the cheapest candidate.  `tune_measured` times the best few candidates on the GPU
(the force evaluation and one list build) instead of trusting the model, and `verify`
measures the chosen configuration's error on a proxy system against the tight FP64
reference.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .accuracy import fft_size_at_least
from .solver import Params

SPLIT_FIT = (-0.416, -0.551)  # log10 E_split = a c + b
FP32_ERROR = 1.5e-6
CUT_FACTOR = 1.2
SIGMA_GRID = (1.15, 1.30, 1.50, 1.75, 2.00)
WINDOW_ERROR = {
    4: (2.8e-2, 2.0e-2, 1.34e-2, 1.05e-2, 9.0e-3),
    6: (9.2e-4, 5.3e-4, 2.8e-4, 1.54e-4, 9.4e-5),
    8: (3.2e-5, 1.25e-5, 4.3e-6, 1.9e-6, 1.0e-6),
    10: (9.7e-7, 2.7e-7, 7.7e-8, 2.4e-8, 9.9e-9),
    12: (2.9e-8, 6.6e-9, 1.3e-9, 3.0e-10, 9.2e-11),
    14: (9.3e-10, 1.5e-10, 2.5e-11, 4.0e-12, 9.0e-13),
}


def split_error(c: float) -> float:
    a, b = SPLIT_FIT
    return 10.0 ** (a * c + b)


def window_error(P: float, sigma: float) -> float:
    """Window (mesh) error for order P at oversampling sigma, log-linear interpolation."""
    orders = sorted(WINDOW_ERROR)
    P = float(min(max(P, orders[0]), orders[-1]))
    lo = max(o for o in orders if o <= P)
    hi = min(o for o in orders if o >= P)

    def at(o):
        ls = np.log(np.asarray(WINDOW_ERROR[o]))
        s = np.asarray(SIGMA_GRID)
        if sigma <= s[0]:
            slope = (ls[1] - ls[0]) / (s[1] - s[0])
            return ls[0] + slope * (sigma - s[0])
        if sigma >= s[-1]:
            slope = (ls[-1] - ls[-2]) / (s[-1] - s[-2])
            return ls[-1] + slope * (sigma - s[-1])
        return float(np.interp(sigma, s, ls))

    if lo == hi:
        return float(math.exp(at(lo)))
    w = (P - lo) / (hi - lo)
    return float(math.exp((1 - w) * at(lo) + w * at(hi)))


def predicted_error(c: float, sigma: float, P: int, precision: int = _lib.ESP_PREC_MIXED, scale: float = 1.0) -> float:
    """Model error; `scale` multiplies the method part (split + window) for a system whose
    structure differs from the model's random electrolyte."""
    e2 = (scale * CUT_FACTOR) ** 2 * (split_error(c) ** 2 + window_error(P, sigma) ** 2)
    if precision == _lib.ESP_PREC_MIXED:
        e2 += FP32_ERROR ** 2
    return math.sqrt(e2)


@dataclass
class CostModel:
    """Per-step cost coefficients on one MI355X (see the module docstring); pairs are
    (FP64, mixed)."""
    pair_ps: Tuple[float, float] = (4.7, 2.47)        # per neighbour entry, pair pass alone (1e6 particles:
                                                      # FP32 0.39 ms at 1.31e8 entries, 0.46 ms at 1.58e8;
                                                      # FP64 0.535 ms at 9.5e7)
    pair_us: Tuple[float, float] = (90.0, 70.0)       # fixed cost of the pair pass
    bin_ns: Tuple[float, float] = (0.15, 0.11)        # per particle (bin + scan + scatter)
    stencil_ps: Tuple[float, float] = (0.53, 0.26)    # per particle per P^3 (spread + gather)
    cell_ps: Tuple[float, float] = (11.8, 31.7)       # per mesh cell (FFT passes)
    halo_ps: Tuple[float, float] = (12.7, 7.2)        # per mesh cell per ((T + P - 1) / T)^3, T = 8
    hide_pair: Tuple[float, float] = (0.25, 0.8)      # the mesh chain hides at most this share
    hide_mesh: Tuple[float, float] = (0.15, 0.4)      # of the pair pass, and at most this share of itself
    overhead_us: Tuple[float, float] = (120.0, 84.0)  # launch and reduction latency of the mesh chain
    overlap: bool = True                              # False: branches add up (no hiding)
    nlist_ps: float = 16.9                            # neighbour-list build per list entry (FP64 positions)
    nlist_us: float = 60.0                            # fixed cost of a build (cell count, scan, fill, order)
    rebuild_every: float = 19.0                       # steps per list build; 0 = no list upkeep

    def nlist_ms(self, N: int, neighbours: float) -> float:
        """One neighbour-list build out to r_c + skin (`neighbours` entries per particle), in ms."""
        return N * neighbours * self.nlist_ps * 1e-9 + self.nlist_us * 1e-3

    def branch_ms(self, N: int, n_cells: int, P: int, neighbours: float, precision: int) -> Tuple[float, float]:
        """(mesh chain, pair pass) each run alone, in ms."""
        k = 0 if precision == _lib.ESP_PREC_FP64 else 1
        halo = ((8 + P - 1) / 8.0) ** 3
        mesh = (N * P ** 3 * self.stencil_ps[k] + n_cells * (self.cell_ps[k] + self.halo_ps[k] * halo)
                + N * self.bin_ns[k] * 1e3) * 1e-9 + self.overhead_us[k] * 1e-3
        pair = N * neighbours * self.pair_ps[k] * 1e-9 + self.pair_us[k] * 1e-3
        return mesh, pair

    def force_ms(self, N: int, n_cells: int, P: int, neighbours: float, precision: int) -> float:
        """One force evaluation (pair pass beside the mesh chain), in ms."""
        k = 0 if precision == _lib.ESP_PREC_FP64 else 1
        mesh, pair = self.branch_ms(N, n_cells, P, neighbours, precision)
        hidden = min(self.hide_pair[k] * pair, self.hide_mesh[k] * mesh) if self.overlap else 0.0
        return mesh + pair - hidden

    def step_ms(self, N: int, n_cells: int, P: int, neighbours: float, precision: int) -> float:
        """A simulation step: the force evaluation plus the amortised neighbour-list build."""
        ms = self.force_ms(N, n_cells, P, neighbours, precision)
        if self.rebuild_every > 0:
            ms += self.nlist_ms(N, neighbours) / self.rebuild_every
        return ms


#: Systems solved inside a batch (many small systems, pair and mesh items of all systems sharing
#: the GPU) are tuned with an additive model that gives the pair pass no overlap credit: on the
#: 512-system batch of 500-50,000 particles its choices measured 13.0 ms per batch step against
#: 15.4 ms for the single-system fit above.  It prices the batched solve alone (no list upkeep).
BATCHED_COST = CostModel(pair_ps=(10.0, 6.5), pair_us=(0.0, 0.0), bin_ns=(0.18, 0.15), stencil_ps=(1.17, 0.62),
                         cell_ps=(90.0, 70.0), halo_ps=(0.0, 0.0), overhead_us=(25.0, 25.0), overlap=False,
                         rebuild_every=0.0)


@dataclass
class Choice:
    params: Params
    c: float
    sigma: float
    predicted_error: float
    predicted_ms: float
    neighbours: float
    measured_ms: Optional[float] = None         # force evaluation, timed on the GPU
    measured_error: Optional[float] = None
    method_scale: float = 1.0
    measured_list_ms: Optional[float] = None    # one neighbour-list build, timed on the GPU
    rebuild_every: float = 0.0                  # steps per list build the ranking assumed

    def ranked_ms(self) -> float:
        """Measured per-step cost used for ranking: force evaluation plus the amortised list
        build."""
        ms = self.measured_ms if self.measured_ms is not None else self.predicted_ms
        if self.measured_list_ms is not None and self.rebuild_every > 0:
            ms += self.measured_list_ms / self.rebuild_every
        return ms

    def describe(self) -> dict:
        p = self.params
        return {"mesh": [int(p.nx), int(p.ny), int(p.nz)], "order": int(p.order), "rcut": round(float(p.rcut), 4),
                "kappa": round(float(p.kappa), 5),
                "c": round(self.c, 3), "sigma": round(self.sigma, 3), "predicted_error": self.predicted_error,
                "predicted_ms": self.predicted_ms, "measured_ms": self.measured_ms, "measured_error": self.measured_error,
                "measured_list_ms": self.measured_list_ms, "rebuild_every": self.rebuild_every}


def fft_sizes(lo: int, hi: int) -> List[int]:
    out, n = [], max(lo, 2)
    while n <= hi:
        n = fft_size_at_least(n, lo)
        if n <= hi:
            out.append(n)
        n += 1
    return out


SIGMA_MIN = SIGMA_GRID[0]  # the accuracy model is not extrapolated below the measured oversampling


def best_split(P: int, target: float, precision: int, c_max: float = 32.0, scale: float = 1.0):
    """(c, sigma) meeting the target at order P with the smallest sigma * c, i.e. the smallest
    r_c = sigma c h / pi at any mesh spacing h; None if the target is out of reach."""
    best = None
    fp32 = FP32_ERROR ** 2 if precision == _lib.ESP_PREC_MIXED else 0.0
    for c in np.arange(4.0, c_max + 1e-9, 0.125):
        budget2 = (target ** 2 - fp32) / (scale * CUT_FACTOR) ** 2 - split_error(c) ** 2
        if budget2 <= 0:
            continue
        budget = math.sqrt(budget2)
        lo, hi = SIGMA_MIN, 4.0
        if window_error(P, hi) > budget:
            continue
        if window_error(P, lo) <= budget:
            hi = lo
        else:
            for _ in range(40):
                mid = 0.5 * (lo + hi)
                if window_error(P, mid) <= budget:
                    hi = mid
                else:
                    lo = mid
        if best is None or hi * c < best[1] * best[0]:
            best = (float(c), hi)
    return best


def batch_tile(P: int, precision: int) -> int:
    """Tile edge for a system solved inside a batch: 12, else 8, whichever first has a
    thread-per-particle LDS box ((T + P - 1)^3 cells plus the window table) within 64 KiB
    (FP64 cells take 8 bytes, so FP64 at P = 10 gets 8).  Batches get their parallelism from
    many systems, so tiles bigger than the single-system 8 (more particles per workgroup, less
    halo) win there.  On MI355X, for 512 systems of 500-50,000 particles in mixed precision:
    9.6 ms per step at 12, 10.0 at 16, 10.0 at 10, 12.0 at 8."""
    el = 8 if precision == _lib.ESP_PREC_FP64 else 4
    for T in (12, 8):
        B = T + P - 1
        if ((B ** 3 + 1) * el + 15) // 16 * 16 + P * 16 * el <= 64 * 1024:
            return T
    return 4


def tune(N: int, box: Sequence[float], target: float = 1e-5, precision: int = _lib.ESP_PREC_MIXED,
         orders: Sequence[int] = (6, 8, 10, 12), mesh_range: Tuple[int, int] = (16, 512), skin: float = 0.4,
         cost: Optional[CostModel] = None, safety: float = 1.25, top: int = 1, method_scale: float = 1.0,
         batched: bool = False) -> List[Choice]:
    """Candidate configurations meeting `target / safety` by the accuracy model, cheapest first
    by the cost model (the first `top` are returned).  `batched` picks the tile edge for
    solving inside a batch (batch_tile); otherwise the library's automatic tile is used."""
    cost = cost or (BATCHED_COST if batched else CostModel())
    L = np.asarray(box[:3], float)
    rho = N / float(np.prod(L))
    goal = target / safety
    splits = {P: best_split(P, goal, precision, scale=method_scale) for P in orders}
    cands: List[Choice] = []
    for n0 in fft_sizes(*mesh_range):
        h = float(L.max()) / n0
        mesh = [fft_size_at_least(math.ceil(Li / h - 1e-9), 4) for Li in L]
        hmax = max(Li / ni for Li, ni in zip(L, mesh))
        for P in orders:
            if P > min(mesh) // 2:
                continue
            if splits[P] is None:
                continue
            c, sigma = splits[P]
            rc = sigma * c * hmax / math.pi
            if rc + skin > 0.5 * float(L.min()):
                continue
            nbar = rho * 4.0 / 3.0 * math.pi * (rc + skin) ** 3
            ms = float(cost.step_ms(N, int(np.prod(mesh)), P, nbar, precision))
            rc = float(rc)
            p = Params(nx=mesh[0], ny=mesh[1], nz=mesh[2], order=P, kappa=c / rc, rcut=rc, precision=precision,
                       tile=batch_tile(P, precision) if batched else 0)
            cands.append(Choice(p, c, sigma, predicted_error(c, sigma, P, precision, method_scale), ms, nbar,
                                rebuild_every=cost.rebuild_every))
    if not cands:
        raise ValueError(f"no ESP configuration reaches {target:g} for N={N}, box={tuple(L)} within the mesh range")
    cands.sort(key=lambda ch: ch.predicted_ms)
    return cands[:top]


def measure_step_ms(choice: Choice, positions: np.ndarray, charges: np.ndarray, box: Sequence[float],
                    steps: int = 10, skin: float = 0.4) -> float:
    """Average time of one ESP force evaluation for this choice on the GPU (HIP events); also
    times one neighbour-list build into choice.measured_list_ms."""
    from .device import Event, Stream, synchronize
    from .md import NeighborList, State
    from .solver import Solver
    st = State(positions, charges, list(box[:3]))
    nl = NeighborList(buffer=skin)
    nl.update(st, choice.params.rcut, force_rebuild=True)  # sizes the list
    synchronize()
    t0 = time.perf_counter()
    nl.update(st, choice.params.rcut, force_rebuild=True)
    synchronize()
    choice.measured_list_ms = (time.perf_counter() - t0) * 1e3
    solver = Solver(choice.params)
    stream = Stream()
    for _ in range(2):
        solver.enqueue_state(st, stream=stream)
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(steps):
        solver.enqueue_state(st, stream=stream)
    e1.record(stream)
    stream.synchronize()
    return e0.elapsed_ms(e1) / steps


def tune_measured(positions: np.ndarray, charges: np.ndarray, box: Sequence[float], target: float = 1e-5,
                  precision: int = _lib.ESP_PREC_MIXED, top: int = 4, **kw) -> List[Choice]:
    """Model candidates, then the best `top` timed on the GPU with this system; fastest first."""
    cands = tune(len(positions), box, target, precision, top=top, **kw)
    for ch in cands:
        ch.measured_ms = measure_step_ms(ch, positions, charges, box)
    cands.sort(key=lambda ch: ch.ranked_ms())
    return cands


def autotune(positions: np.ndarray, charges: np.ndarray, box: Sequence[float], target: float = 1e-5,
             precision: int = _lib.ESP_PREC_MIXED, measure_time: bool = False, top: int = 4, max_rounds: int = 5,
             reference: Optional[np.ndarray] = None, **kw) -> Choice:
    """Tune for this very system: model candidates (timed on the GPU when `measure_time`), the
    best one's force error measured against the tight FP64 reference, and the accuracy goal
    tightened by the measured/predicted ratio until the measured error meets `target`.

    The relative error depends on the system's structure (close contacts raise the RMS force
    and lower the relative error), which the density-0.1 random-electrolyte model cannot know;
    the measurement closes that gap."""
    from .accuracy import forces_with, reference_forces, rel_rms_error
    ref = reference if reference is not None else reference_forces(positions, charges, box)
    safety = float(kw.pop("safety", 1.25))
    scale = 1.0
    N = len(positions)
    last = None
    for _ in range(max_rounds):
        cands = tune(N, box, target, precision, top=top if measure_time else 1, safety=safety, method_scale=scale, **kw)
        if measure_time:
            for ch in cands:
                ch.measured_ms = measure_step_ms(ch, positions, charges, box)
            cands.sort(key=lambda ch: ch.ranked_ms())
        best = cands[0]
        best.measured_error = rel_rms_error(forces_with(best.params, positions, charges, box), ref)
        best.method_scale = scale
        last = best
        if best.measured_error <= target:
            return best
        # rescale the method part of the model to this system (the FP32 part is not rescaled)
        fp32 = FP32_ERROR if precision == _lib.ESP_PREC_MIXED else 0.0
        method_meas = math.sqrt(max(best.measured_error ** 2 - fp32 ** 2, (0.5 * best.measured_error) ** 2))
        method_pred = math.sqrt(max(best.predicted_error ** 2 - fp32 ** 2, 1e-300))
        scale *= max(method_meas / method_pred, 1.05)
        safety = 1.1
    raise ValueError(f"autotune: {target:g} not met after {max_rounds} rounds (last {last.describe() if last else None})")


def proxy_system(choice: Choice, density: float, charge_values: Sequence[float] = (1.0, -1.0),
                 charge_weights: Optional[Sequence[float]] = None, min_particles: int = 4000, seed: int = 5):
    """A random neutral electrolyte at `density` in a cubic box whose mesh spacing equals the
    choice's, large enough for its cutoff (the force error is intensive)."""
    p = choice.params
    h = p.rcut * math.pi / (choice.sigma * choice.c)
    n = fft_size_at_least(max(2 * p.order, math.ceil(2.5 * (p.rcut + 0.4) / h),
                              math.ceil((min_particles / density) ** (1 / 3) / h)))
    L = n * h
    N = int(round(density * L ** 3 / 2)) * 2
    rng = np.random.default_rng(seed)
    vals = np.asarray(charge_values, float)
    w = np.ones(len(vals)) / len(vals) if charge_weights is None else np.asarray(charge_weights, float) / np.sum(charge_weights)
    q = rng.choice(vals, size=N, p=w)
    # neutralise by flipping signs of random particles until the sum is zero (integer charges)
    for _ in range(4 * N):
        s = q.sum()
        if abs(s) < 1e-12:
            break
        cand = np.flatnonzero(np.sign(q) == np.sign(s))
        i = rng.choice(cand)
        if abs(2 * q[i]) <= abs(s) + 1e-12 or abs(q[i]) == 1.0:
            q[i] = -q[i]
    pos = rng.uniform(-L / 2, L / 2, (N, 3))
    params = Params(nx=n, ny=n, nz=n, order=p.order, kappa=p.kappa, rcut=p.rcut, precision=p.precision)
    return pos, q, (L, L, L), params


def verify(choice: Choice, density: float, **kw) -> float:
    """Measured relative RMS force error of the choice on a proxy system (GPU, tight reference)."""
    from .accuracy import measure_error
    pos, q, box, params = proxy_system(choice, density, **kw)
    choice.measured_error = measure_error(params, pos, q, box)
    return choice.measured_error


__all__ = ["CostModel", "Choice", "tune", "tune_measured", "verify", "proxy_system", "predicted_error", "split_error",
           "window_error", "fft_sizes"]
