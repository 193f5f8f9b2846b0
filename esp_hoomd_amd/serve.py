"""Serving ESP solves: continuous batching of MD requests on one GPU.

A request is an independent periodic electrolyte (N particles, number density, charges) that
asks for `steps` MD steps (velocity Verlet, WCA + ESP electrostatics).  The server

* admits arrived requests while the HBM footprint of everything resident (particle state,
  neighbour list, mesh, tables and scratch) stays under a budget and a slot is free —
  requests that do not fit wait, in arrival order;
* keeps every admitted request's particle, mesh and neighbour-list state resident on the
  device until it finishes;
* advances all resident requests by one MD step per tick, with the ESP forces of all of
  them evaluated by one batched launch per stage (replayed from a captured hipGraph while
  the set of resident requests is unchanged), so requests that arrive while others run join
  the next tick and their work overlaps on the device;
* records per-step latency as each request sees it (wall time from its previous step, or its
  admission, to the end of the tick that advanced it, host work included), queueing delay,
  completed requests and particle-steps.

Per tick the host issues a fixed number of launches whatever the number of resident
requests: batched first half-kick and drift, one displacement check for all neighbour lists
(only lists that need it are rebuilt), the batched ESP forces (hipGraph replay), batched WCA
and the batched second half-kick.

ESP parameters per request come from the autotuner's model at the force-error target (cached
per size/density bucket), with the model's method scale measured on probe requests
(`calibrate`); `verify_results` measures every finished request's force error against the
tight FP64 reference.  The step is deterministic in mixed precision: a request
computes the same trajectory alone or in any mix (the tests check this).
"""
from __future__ import annotations

import ctypes as C
import heapq
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .batch import Batch
from .device import Stream, synchronize
from .md import BatchMD, NeighborList, State, WCA, maxwell_boltzmann
from .solver import Params, Solver
from . import tune as _tune
from .workloads import lattice_positions, neutral_charges

WCA_SIGMA = 1.0
SKIN = 0.4
#: margin of the per-request tuning below the force-error target: a result's error varies with
#: its structure along the trajectory around the calibrated model (one in ~850 served results
#: exceeded 1e-5 at a margin of 1.25)
SAFETY = 1.4


@dataclass
class Request:
    rid: int
    N: int
    density: float
    steps: int
    arrival: float = 0.0          # seconds after the server starts
    seed: int = 0
    q2_frac: float = 0.1
    kT: float = 1.0
    # filled by the server
    params: Optional[Params] = None
    footprint: int = 0
    admitted_at: Optional[float] = None
    finished_at: Optional[float] = None
    steps_done: int = 0
    state: Optional[State] = field(default=None, repr=False)
    solver: Optional[Solver] = field(default=None, repr=False)
    nlist: Optional[NeighborList] = field(default=None, repr=False)
    first_forces: Optional[np.ndarray] = field(default=None, repr=False)
    final_positions: Optional[np.ndarray] = field(default=None, repr=False)
    final_forces: Optional[np.ndarray] = field(default=None, repr=False)
    force_error: Optional[float] = None  # relative RMS electrostatic force error (verify_results)
    prepared: Optional[tuple] = field(default=None, repr=False)
    admitted_wall: float = 0.0
    last_step_end: float = 0.0

    def setup(self) -> Tuple[np.ndarray, np.ndarray, float]:
        N = self.N - (self.N % 2)
        pos, L = lattice_positions(N, self.density, self.seed)
        return pos, neutral_charges(N, self.q2_frac, self.seed), L


def footprint_bytes(N: int, params: Params, neighbours: float) -> int:
    """Device bytes a request holds while resident: particle arrays (position, charge,
    velocity, force, virial: 152 B), neighbour list (4 B per entry at 1.3x the mean count,
    plus counts and heads), the WCA inner list (16 entries, count and head: 76 B), sorting
    scratch (~48 B), and the solver's meshes and tables."""
    M = params.nx * params.ny * params.nz
    Mh = (params.nx // 2 + 1) * params.ny * params.nz
    R = 8 if params.precision == _lib.ESP_PREC_FP64 else 4
    per_particle = 152 + 48 + 12 + 76 + 4 * 1.3 * neighbours
    mesh = M * 8 + 2 * Mh * R + M * R + 2 * Mh * 8
    return int(N * per_particle + mesh + (1 << 20))


class Server:
    """Continuous-batching ESP/MD server (see the module docstring)."""

    def __init__(self, hbm_budget: int = 32 << 30, max_active: int = 64, precision: int = _lib.ESP_PREC_MIXED,
                 target: float = 1e-5, dt: float = 0.002, use_graph: bool = True, keep_results: bool = False,
                 tuner_orders: Sequence[int] = (6, 8, 10), mesh_range: Tuple[int, int] = (16, 128),
                 keep_final_positions: bool = False):
        self.hbm_budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.precision = precision
        self.target = target
        self.dt = dt
        self.keep_results = keep_results
        self.keep_final_positions = keep_final_positions or keep_results
        # accuracy-model method scale per probe density (calibrate()); 1.0 until calibrated
        self.method_scales: Dict[float, float] = {}
        self.tuner_orders = tuple(tuner_orders)
        self.mesh_range = mesh_range
        self.batch = Batch(use_graph=use_graph)       # every resident request, every tick
        self.batch_new = Batch(use_graph=False)       # initial forces of newly admitted requests
        self.md = BatchMD()                           # integration, WCA and list upkeep of the resident set
        # the resident set's steps run on their own (non-blocking) stream; admissions (uploads,
        # first neighbour list, first forces) run on the null stream meanwhile, so the host
        # prepares new requests while the GPU advances the others
        self.stream = Stream()
        self._inflight = None  # (requests, launch time, particles) of the step on the stream
        self._member_ids: List[int] = []
        self._esp_args = None
        self.step_latency: List[float] = []
        self.wca = WCA(WCA_SIGMA, 1.0)
        self.pending: List[Tuple[float, int, Request]] = []
        self.active: List[Request] = []
        self.done: List[Request] = []
        self.resident_bytes = 0
        self.peak_bytes = 0
        self.tick_ms: List[float] = []
        self.tick_active: List[int] = []
        self.tick_particles: List[int] = []
        self.waited_for_memory = 0
        # wall seconds per phase of the loop (admission, first forces of new requests, member
        # bookkeeping, list upkeep, ESP launch, the rest of the step up to its synchronisation)
        self.admit_s: Dict[str, float] = {k: 0.0 for k in ("state", "nlist", "solver", "first_forces")}
        self.phase_s: Dict[str, float] = {k: 0.0 for k in ("admit", "wait", "retire", "members", "nlists", "esp",
                                                          "step_rest")}
        self._tune_cache: Dict[Tuple[int, int], Params] = {}
        self._lib = _lib.load()

    # ---- request lifecycle ----------------------------------------------------------
    def submit(self, req: Request) -> None:
        """Queue a request.  Its particle data (the client's input) is generated here, so a
        stream submitted ahead of run() is served without client-side work in the loop."""
        pos, q, L = req.setup()
        params, nbar = self._params_for(len(pos), L)
        req.prepared = (pos, q, L, params, footprint_bytes(len(pos), params, nbar))
        heapq.heappush(self.pending, (req.arrival, req.rid, req))

    # ---- accuracy ---------------------------------------------------------------------
    def calibrate(self, densities: Sequence[float] = (0.05, 0.1, 0.2), N: int = 20_000, q2_frac: float = 0.1,
                  seed: int = 99) -> Dict[float, float]:
        """Measure the accuracy model's method scale on probe requests (the request generator's
        systems at each density) with the autotuner, which compares forces against the tight
        FP64 reference and rescales the model until the target is met.  The relative error of
        these lattice-started systems depends mostly on density (their RMS force falls with the
        lattice spacing), so a request uses the scale of the nearest probe density at or below
        its own (the larger scale of the two neighbours when between)."""
        for rho in densities:
            pos, L = lattice_positions(int(N), float(rho), seed)
            q = neutral_charges(int(N), q2_frac, seed)
            ch = _tune.autotune(pos, q, (L, L, L), self.target, self.precision, orders=self.tuner_orders,
                                mesh_range=self.mesh_range, batched=True)
            self.method_scales[float(rho)] = float(ch.method_scale)
        self._tune_cache.clear()
        return dict(self.method_scales)

    def _scale_for(self, density: float) -> float:
        if not self.method_scales:
            return 1.0
        rhos = sorted(self.method_scales)
        below = [r for r in rhos if r <= density]
        above = [r for r in rhos if r >= density]
        cands = []
        if below:
            cands.append(self.method_scales[below[-1]])
        if above:
            cands.append(self.method_scales[above[0]])
        return max(cands)

    def verify_results(self, reqs: Optional[Sequence[Request]] = None) -> List[float]:
        """Relative RMS electrostatic force error of each finished request at its final
        positions: its ESP configuration against the tight FP64 reference (on the GPU, outside
        the serving loop).  Needs keep_final_positions (or keep_results).  Sets
        request.force_error and returns the errors."""
        from .accuracy import forces_with, reference_forces, rel_rms_error
        errs = []
        for r in (self.done if reqs is None else reqs):
            if r.final_positions is None:
                raise RuntimeError("Server.verify_results: final positions were not kept")
            _, q, L = r.setup()
            box = (L, L, L)
            r.force_error = rel_rms_error(forces_with(r.params, r.final_positions, q, box),
                                          reference_forces(r.final_positions, q, box))
            errs.append(r.force_error)
        return errs

    def _params_for(self, N: int, L: float) -> Tuple[Params, float]:
        density = N / L ** 3
        scale = self._scale_for(density)
        key = (int(round(math.log(N) * 8)), int(round(math.log(density) * 16)), scale)
        if key not in self._tune_cache:
            ch = _tune.tune(N, (L, L, L), self.target, self.precision, orders=self.tuner_orders,
                            mesh_range=self.mesh_range, batched=True, method_scale=scale, safety=SAFETY)[0]
            self._tune_cache[key] = (ch.params, ch.c, ch.sigma)
        p, c, sigma = self._tune_cache[key]
        # same c and sigma, mesh spacing and cutoff rescaled to this box
        n = _tune.fft_size_at_least(max(2 * p.order, math.ceil(L / (math.pi * p.rcut / (sigma * c)) - 1e-9)))
        h = L / n
        rc = sigma * c * h / math.pi
        if rc + SKIN > 0.5 * L:
            ch = _tune.tune(N, (L, L, L), self.target, self.precision, orders=self.tuner_orders, mesh_range=self.mesh_range,
                            batched=True, method_scale=scale, safety=SAFETY)[0]
            return ch.params, ch.neighbours
        params = Params(nx=n, ny=n, nz=n, order=p.order, kappa=c / rc, rcut=rc, precision=self.precision, tile=p.tile)
        return params, density * 4.0 / 3.0 * math.pi * (rc + SKIN) ** 3

    def _admit(self, now: float) -> bool:
        changed = False
        new: List[Request] = []
        while self.pending and self.pending[0][0] <= now and len(self.active) < self.max_active:
            req = self.pending[0][2]
            if req.prepared is None:
                pos, q, L = req.setup()
                params, nbar = self._params_for(len(pos), L)
                req.prepared = (pos, q, L, params, footprint_bytes(len(pos), params, nbar))
            pos, q, L, params, need = req.prepared
            if self.resident_bytes + need > self.hbm_budget and self.active:
                self.waited_for_memory += 1
                break
            heapq.heappop(self.pending)
            req.params, req.footprint, req.prepared = params, need, None
            ta = time.perf_counter()
            # velocities drawn on the host; force and virial are overwritten by the first forces
            st = State(pos, q, [L, L, L], velocities=maxwell_boltzmann(np.ones(len(pos)), req.kT, seed=req.seed + 1),
                       zero_outputs=False, pooled=True)
            req.state = st
            tb = time.perf_counter()
            req.nlist = NeighborList(buffer=SKIN, inner=self.wca.r_cut)
            req.nlist.update(st, max(params.rcut, self.wca.r_cut), force_rebuild=True)
            tc = time.perf_counter()
            req.solver = Solver(params)
            self.admit_s["state"] += tb - ta
            self.admit_s["nlist"] += tc - tb
            self.admit_s["solver"] += time.perf_counter() - tc
            req.admitted_at = now
            req.admitted_wall = time.perf_counter()
            self.active.append(req)
            self.resident_bytes += need
            self.peak_bytes = max(self.peak_bytes, self.resident_bytes)
            new.append(req)
            changed = True
        if new:
            # forces at the initial positions (the first half kick needs them), on the null
            # stream beside the resident set's step; a separate batch, so the resident set's
            # captured graph stays valid
            tf = time.perf_counter()
            self._forces(new, self.batch_new)
            if self.keep_results:
                for r in new:
                    r.first_forces = r.state.force.numpy().copy()
            _lib.check(self._lib.esp_stream_synchronize(None))
            self.admit_s["first_forces"] += time.perf_counter() - tf
        return changed

    def _forces(self, reqs: List[Request], batch: Batch) -> None:
        """Net force of the given requests: batched ESP (overwrites), then WCA (adds)."""
        batch.enqueue([(r.solver, r.state) for r in reqs])
        for r in reqs:
            self.wca.add_to(r.state)

    def _retire(self, now: float) -> None:
        keep = []
        for r in self.active:
            if r.steps_done >= r.steps:
                r.finished_at = now
                if self.keep_final_positions:
                    r.final_positions = r.state.positions().copy()
                if self.keep_results:
                    r.final_forces = r.state.force.numpy().copy()
                self.resident_bytes -= r.footprint
                r.state = r.solver = r.nlist = None
                self.done.append(r)
            else:
                keep.append(r)
        self.active = keep

    def _sync_members(self) -> None:
        """Re-describe the resident set to the batched MD and ESP launches after it changed."""
        ids = [r.rid for r in self.active]
        if ids == self._member_ids:
            return
        self._member_ids = ids
        self.md.set([(r.state, r.nlist) for r in self.active], seeds=[r.seed for r in self.active])
        self._esp_args = self.batch.prepare([(r.solver, r.state) for r in self.active])

    def _complete(self) -> None:
        """Wait for the step in flight and account it: each request's step latency runs from its
        previous step's completion (or its admission) to now."""
        if self._inflight is None:
            return
        reqs, t0, particles = self._inflight
        tw = time.perf_counter()
        self.stream.synchronize()
        end = time.perf_counter()
        self.phase_s["wait"] += end - tw
        for r in reqs:
            r.steps_done += 1
            self.step_latency.append((end - (r.last_step_end if r.last_step_end else r.admitted_wall)) * 1e3)
            r.last_step_end = end
        self.tick_ms.append((end - t0) * 1e3)
        self.tick_active.append(len(reqs))
        self.tick_particles.append(particles)
        self._inflight = None

    def _launch(self) -> None:
        """Enqueue one MD step of every resident request on the server's stream: batched first
        half kick + drift, one displacement check (rebuilding only the lists that need it),
        batched ESP forces (captured graph), batched WCA, batched second half kick.  Returns
        once the step is enqueued; _complete waits for it."""
        ph = self.phase_s
        t0 = time.perf_counter()
        self._sync_members()
        t1 = time.perf_counter()
        st = self.stream
        self.md.integrate_first(self.dt, stream=st)
        for i in self.md.update_nlists(stream=st):
            Batch.refresh(self._esp_args, i)
        t2 = time.perf_counter()
        self.batch.enqueue_prepared(self._esp_args, stream=st)
        t3 = time.perf_counter()
        self.md.wca(WCA_SIGMA, 1.0, stream=st)
        self.md.integrate_second(self.dt, step=len(self.tick_ms) + 1, stream=st)
        t4 = time.perf_counter()
        ph["members"] += t1 - t0
        ph["nlists"] += t2 - t1
        ph["esp"] += t3 - t2
        ph["step_rest"] += t4 - t3
        self._inflight = (list(self.active), t0, sum(r.state.N for r in self.active))

    def tick(self, now: float) -> None:
        """One MD step of every resident request, waited for."""
        self._launch()
        self._complete()

    def run(self, time_limit: float = 1e9) -> dict:
        """Serve until every submitted request is done (or time_limit seconds pass)."""
        start = time.perf_counter()
        while self.pending or self.active or self._inflight is not None:
            now = time.perf_counter() - start
            if now > time_limit:
                break
            # admissions overlap the step in flight
            ta = time.perf_counter()
            self._admit(now)
            self.phase_s["admit"] += time.perf_counter() - ta
            self._complete()
            tr = time.perf_counter()
            self._retire(tr - start)
            self.phase_s["retire"] += time.perf_counter() - tr
            if not self.active:
                # idle until the next arrival
                wait = self.pending[0][0] - (time.perf_counter() - start) if self.pending else 0.0
                if wait > 0:
                    time.sleep(min(wait, 0.01))
                continue
            self._launch()
        self._complete()
        wall = time.perf_counter() - start
        return self.stats(wall)

    def stats(self, wall: float) -> dict:
        # per-step latency as a request sees it: the wall time from its previous step (or its
        # admission) to the end of the tick that advanced it, host work between ticks included
        lat = np.asarray(self.step_latency) if self.step_latency else np.zeros(1)
        steps = int(sum(self.tick_active))
        psteps = int(sum(self.tick_particles))
        qdelay = [r.admitted_at - r.arrival for r in self.done if r.admitted_at is not None]
        return {
            "wall_s": wall,
            "requests_done": len(self.done),
            "requests_per_s": len(self.done) / wall if wall > 0 else 0.0,
            "solves": steps,
            "solves_per_s": steps / wall if wall > 0 else 0.0,
            "particle_steps_per_s": psteps / wall if wall > 0 else 0.0,
            "step_latency_ms_p50": float(np.percentile(lat, 50)),
            "step_latency_ms_p99": float(np.percentile(lat, 99)),
            "queue_delay_s_p50": float(np.percentile(qdelay, 50)) if qdelay else 0.0,
            "queue_delay_s_p99": float(np.percentile(qdelay, 99)) if qdelay else 0.0,
            "ticks": len(self.tick_ms),
            "mean_active": float(np.mean(self.tick_active)) if self.tick_active else 0.0,
            "peak_resident_bytes": self.peak_bytes,
            "admissions_deferred_for_memory": self.waited_for_memory,
            "batch": self.batch.stats(),
            "phase_s": {k: round(v, 3) for k, v in self.phase_s.items()},
            "admit_s": {k: round(v, 3) for k, v in self.admit_s.items()},
        }


def poisson_requests(n: int, rate: float, seed: int = 7, N_range=(1_000, 200_000), density_range=(0.05, 0.2),
                     steps_range=(10, 1000), q2_frac: float = 0.1) -> List[Request]:
    """n requests with Poisson arrivals at `rate` per second; N and steps log-uniform,
    density uniform."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    N = np.exp(rng.uniform(np.log(N_range[0]), np.log(N_range[1]), n)).astype(int) // 2 * 2
    dens = rng.uniform(*density_range, n)
    steps = np.exp(rng.uniform(np.log(steps_range[0]), np.log(steps_range[1]), n)).astype(int)
    return [Request(rid=i, N=int(N[i]), density=float(dens[i]), steps=int(steps[i]), arrival=float(t[i]),
                    seed=int(seed * 100003 + i), q2_frac=q2_frac) for i in range(n)]


__all__ = ["Request", "Server", "poisson_requests", "footprint_bytes", "lattice_positions", "neutral_charges"]
