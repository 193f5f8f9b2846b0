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
requests: batched first half-kick and drift, the rebuilds of the neighbour lists that need one,
the batched ESP forces (hipGraph replay), batched WCA, the batched second half-kick and a
predicted displacement check.  The tick has no wait inside it: which lists a tick rebuilds is
known before it starts, from the previous tick's check of the position its drift will produce
(the same arithmetic as the drift, so the decisions are the ones a check after the drift would
give).  A rebuilt list that overflowed its rows is found when the tick completes; that request's
step is then repeated alone from a copy of its state taken before the tick (same bits).

An admitted request's uploads are enqueued behind the tick in flight from pinned staging, and
it joins the next tick held: that tick builds its first list and computes its forces at the
initial positions with everyone else's, without integrating it.  Retired requests hand their
solver and list objects to new ones, so admissions reuse device arrays instead of allocating.
Between two ticks the host waits once, for the tick to complete.

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
from .device import PinnedPool, Stream, synchronize
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
    error: Optional[str] = None          # why the request failed (set up, admission); it takes no steps
    prepared: Optional[tuple] = field(default=None, repr=False)
    snapshot: Optional[tuple] = field(default=None, repr=False)  # state before a tick that rebuilds its list
    first_tick: bool = False                     # admitted: the next tick computes its first forces
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
                 keep_final_positions: bool = False, predict_lists: bool = True, reuse_contexts: bool = True):
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
        self.md = BatchMD()                           # integration, WCA and list upkeep of the resident set
        self._pinned = PinnedPool()                   # staging of the admitted requests' uploads
        self._orphans: List[State] = []               # states of failed admissions, freed after a wait
        self._held_uploaded = False                   # the batched MD holds some member this tick
        # solvers and lists of retired requests, handed to new ones (device memory reused; at
        # most max_active held beside the resident footprint)
        self.reuse_contexts = bool(reuse_contexts)
        self._ctx_pool: List[tuple] = []
        self.contexts_reused = 0
        # per request id: whether its next step must rebuild its list (predicted displacement)
        self._need: Dict[int, bool] = {}
        self.redone_steps = 0
        # False: check the displacements after each drift and wait for it inside the tick (the
        # decisions are the same; kept to show that)
        self.predict_lists = bool(predict_lists)
        # the resident set's steps and the admitted requests' uploads run on the server's own
        # (non-blocking) stream; the host prepares new requests while the GPU advances the others
        self.stream = Stream()
        self._inflight = None  # (requests, launch time, step number, held flags) of the tick on the stream
        self._member_ids: List[int] = []
        self._esp_args = None
        self.step_latency: List[float] = []
        self.wca = WCA(WCA_SIGMA, 1.0)
        self.pending: List[Tuple[float, int, Request]] = []
        self.active: List[Request] = []
        self.done: List[Request] = []
        self.failed: List[Request] = []
        self.resident_bytes = 0
        self.peak_bytes = 0
        self.tick_ms: List[float] = []
        self.tick_active: List[int] = []
        self.tick_particles: List[int] = []
        self.tick_end: List[float] = []  # completion time of each tick, seconds since the first run() began
        self._run_start = None
        self.waited_for_memory = 0
        # wall seconds per phase of the loop (admission, member
        # bookkeeping, list upkeep, ESP launch, the rest of the step up to its synchronisation)
        self.admit_s: Dict[str, float] = {k: 0.0 for k in ("state", "nlist", "solver", "prepare")}
        self.phase_s: Dict[str, float] = {k: 0.0 for k in ("admit", "wait", "retire", "members", "nlists", "esp",
                                                          "step_rest")}
        self._tune_cache: Dict[Tuple[int, int], Params] = {}
        self._lib = _lib.load()

    # ---- request lifecycle ----------------------------------------------------------
    def submit(self, req: Request) -> None:
        """Queue a request.  Its particle data and initial velocities (the client's input) are
        generated here, so a stream submitted ahead of run() is served without client-side work
        in the loop."""
        try:
            req.prepared = self._prepare(req)
        except (ValueError, ArithmeticError, _lib.EspError) as e:
            self._fail(req, f"set-up: {e}")
            return
        if req.prepared[4] > self.hbm_budget:
            self._fail(req, f"footprint {req.prepared[4]} bytes exceeds the HBM budget of {self.hbm_budget} bytes")
            req.prepared = None
            return
        heapq.heappush(self.pending, (req.arrival, req.rid, req))

    def _fail(self, req: Request, msg: str) -> None:
        """A request that cannot be served: recorded with its reason; the others are unaffected."""
        req.error = msg
        self.failed.append(req)

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

    def _prepare(self, req: Request) -> tuple:
        """The request's particles, parameters, footprint and initial velocities (host work, done
        at submission when the request is known ahead)."""
        if not (req.N >= 2 and req.density > 0 and math.isfinite(req.density) and req.steps >= 1):
            raise ValueError(f"request {req.rid}: needs N >= 2, a positive density and steps >= 1")
        pos, q, L = req.setup()
        params, nbar = self._params_for(len(pos), L)
        vel = maxwell_boltzmann(np.ones(len(pos)), req.kT, seed=req.seed + 1)
        return (pos, q, L, params, footprint_bytes(len(pos), params, nbar), vel)

    def _admit(self, now: float) -> bool:
        """Admit arrived requests that fit.  Their uploads are enqueued on the server's stream
        behind the step in flight (from pinned staging; nothing here waits for the GPU) and they
        join the next tick held: that tick builds their first neighbour list and computes their
        forces at the initial positions with everyone else's, without integrating them, and
        predicts their first drift; their MD steps start with the tick after."""
        changed = False
        while self.pending and self.pending[0][0] <= now and len(self.active) < self.max_active:
            req = self.pending[0][2]
            if req.prepared is None:
                req.prepared = self._prepare(req)
            pos, q, L, params, need, vel = req.prepared
            if self.resident_bytes + need > self.hbm_budget and self.active:
                self.waited_for_memory += 1
                break
            heapq.heappop(self.pending)
            req.params, req.footprint, req.prepared = params, need, None
            ta = time.perf_counter()
            try:
                # force and virial are overwritten by the first forces
                req.state = State(pos, q, [L, L, L], velocities=vel, zero_outputs=False, pooled=True, stream=self.stream,
                                  staging=self._pinned)
                tb = time.perf_counter()
                # a retired request's solver and list, when one is free: their device arrays are
                # reused (set_params / reset rebuild everything that depends on the parameters)
                pooled = self._take_context(len(pos), params)
                if pooled is not None:
                    req.solver, req.nlist = pooled
                    req.nlist.reset(max(params.rcut, self.wca.r_cut))
                    self.contexts_reused += 1
                else:
                    req.nlist = NeighborList(buffer=SKIN, inner=self.wca.r_cut)
                    req.nlist._ensure(max(params.rcut, self.wca.r_cut))
                tc = time.perf_counter()
                if pooled is not None:
                    req.solver.set_params(params)
                else:
                    req.solver = Solver(params)
                td = time.perf_counter()
                # tables, window and influence function now, while the step in flight runs,
                # not in the next tick's batched call (about 1.3 ms of host time per request)
                req.solver.prepare(req.state, stream=self.stream)
            except (ValueError, _lib.EspError) as e:
                # this request alone fails; its uploads (if any) drain with the stream
                self._fail(req, f"admission: {e}")
                if req.state is not None:
                    self._orphans.append(req.state)
                req.state = req.solver = req.nlist = None
                continue
            self.admit_s["state"] += tb - ta
            self.admit_s["nlist"] += tc - tb
            self.admit_s["solver"] += td - tc
            self.admit_s["prepare"] += time.perf_counter() - td
            req.admitted_at = now
            req.admitted_wall = time.perf_counter()
            req.first_tick = True
            self._need[req.rid] = True  # no list yet
            self.active.append(req)
            self.resident_bytes += need
            self.peak_bytes = max(self.peak_bytes, self.resident_bytes)
            changed = True
        return changed

    @staticmethod
    def _cells(params: Params) -> int:
        return int(params.nx) * int(params.ny) * int(params.nz)

    def _take_context(self, N: int, params: Params):
        """The pooled (solver, list) with the smallest arrays that already hold a system of N
        particles on this mesh, or None.  A context too small for the request is left in the
        pool: growing its arrays would free the old ones, and hipFree waits for the whole GPU
        (the step in flight), which admissions are meant to overlap."""
        cells = self._cells(params)
        best = None
        for k, (_, _, n_cap, c_cap) in enumerate(self._ctx_pool):
            if n_cap >= N and c_cap >= cells and (best is None or (c_cap, n_cap) < self._ctx_pool[best][2:]):
                best = k
        if best is None:
            return None
        solver, nlist, _, _ = self._ctx_pool.pop(best)
        return solver, nlist

    def _pool_context(self, r: Request) -> None:
        # keep the largest arrays seen by this context (they never shrink)
        n_cap = max(r.state.N, getattr(r.solver, "_pool_n", 0))
        c_cap = max(self._cells(r.params), getattr(r.solver, "_pool_cells", 0))
        r.solver._pool_n, r.solver._pool_cells = n_cap, c_cap
        self._ctx_pool.append((r.solver, r.nlist, n_cap, c_cap))
        if len(self._ctx_pool) > 2 * self.max_active:
            # drop the smallest (its arrays are freed now, rarely)
            self._ctx_pool.remove(min(self._ctx_pool, key=lambda e: (e[3], e[2])))

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
                if self.reuse_contexts:
                    self._pool_context(r)
                r.state = r.solver = r.nlist = r.snapshot = None
                self._need.pop(r.rid, None)
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
        self.md.set([(r.state, r.nlist) for r in self.active], seeds=[r.seed for r in self.active], stream=self.stream)
        self._esp_args = self.batch.prepare([(r.solver, r.state) for r in self.active])

    def _complete(self) -> None:
        """Wait for the step in flight and account it: each request's step latency runs from its
        previous step's completion (or its admission) to now.  Reads the tick's predicted
        displacement checks and repeats the step of any request whose rebuilt list overflowed.
        Requests held in the tick (admitted: their first forces) took no MD step in it."""
        if self._inflight is None:
            return
        reqs, t0, step_no, held = self._inflight
        tw = time.perf_counter()
        self.stream.synchronize()
        self.phase_s["wait"] += time.perf_counter() - tw
        if self.predict_lists:
            overflowed = self.md.finish_rebuilds()
            for r, need in zip(reqs, self.md.read_predicted()):
                self._need[r.rid] = need
            for i in overflowed:
                self._redo_step(reqs[i], step_no, held[i])
        end = time.perf_counter()
        particles = 0
        stepped = 0
        for r, h in zip(reqs, held):
            if h:
                r.first_tick = False
                if self.keep_results:
                    r.first_forces = r.state.force.numpy().copy()
                r.state.release_staging()
                continue
            r.steps_done += 1
            stepped += 1
            particles += r.state.N
            self.step_latency.append((end - (r.last_step_end if r.last_step_end else r.admitted_wall)) * 1e3)
            r.last_step_end = end
        self.tick_ms.append((end - t0) * 1e3)
        self.tick_active.append(stepped)
        self.tick_particles.append(particles)
        self.tick_end.append(end - (self._run_start if self._run_start is not None else t0))
        self._inflight = None
        for st in self._orphans:
            st.release_staging()
        self._orphans = []

    def _snapshot(self, r: Request) -> None:
        """Copy a request's positions, velocities and forces (the state a step starts from) on
        the server's stream, for repeating its step if the list it rebuilds overflows."""
        from .device import DeviceArray
        st = r.state
        if r.snapshot is None:
            r.snapshot = tuple(DeviceArray(a.shape, a.dtype, pooled=True) for a in (st.postype, st.vel, st.force))
        for dst, src in zip(r.snapshot, (st.postype, st.vel, st.force)):
            _lib.check(self._lib.esp_memcpy_dtod(dst._ptr, src._ptr, src.nbytes, self.stream.handle))

    def _redo_step(self, r: Request, step_no: int, held: bool) -> None:
        """Repeat request r's last tick alone: restore the state it started from, drift (unless
        it was held), build its list with the raised capacity (waiting for it), then forces,
        WCA, second half kick (unless held) and the displacement prediction -- the launches of a
        tick for one member, so the result is the same bits the tick would have given."""
        st, h = r.state, self.stream.handle
        for src, dst in zip(r.snapshot, (st.postype, st.vel, st.force)):
            _lib.check(self._lib.esp_memcpy_dtod(dst._ptr, src._ptr, src.nbytes, h))
        solo = BatchMD()
        solo.set([(st, r.nlist)], seeds=[r.seed])
        if not held:
            solo.integrate_first(self.dt, stream=self.stream)
        r.nlist.update(st, max(r.params.rcut, self.wca.r_cut), force_rebuild=True, stream=self.stream)
        solo.set([(st, r.nlist)], seeds=[r.seed])  # the list arrays may have moved
        Batch(use_graph=False).enqueue([(r.solver, st)], stream=self.stream)
        solo.wca(WCA_SIGMA, 1.0, stream=self.stream)
        if not held:
            solo.integrate_second(self.dt, step=step_no, stream=self.stream)
        solo.predict(self.dt, stream=self.stream)
        self.stream.synchronize()
        self._need[r.rid] = solo.read_predicted()[0]
        self._member_ids = []  # re-describe the resident set (list pointers may have changed)
        self.redone_steps += 1

    def _launch(self) -> None:
        """Enqueue one MD step of every resident request on the server's stream: batched first
        half kick + drift, the list rebuilds the previous tick's predicted check asked for,
        batched ESP forces (captured graph), batched WCA, batched second half kick and the
        predicted check for the next tick.  Nothing in it waits for the GPU; _complete waits
        for it."""
        ph = self.phase_s
        t0 = time.perf_counter()
        self._sync_members()
        t1 = time.perf_counter()
        st = self.stream
        held = [r.first_tick for r in self.active]
        if any(held) or self._held_uploaded:
            self.md.set_hold(held, stream=st)
            self._held_uploaded = any(held)
        if self.predict_lists:
            need = [self._need.get(r.rid, True) for r in self.active]
            for r, nd in zip(self.active, need):
                if nd:
                    self._snapshot(r)
            self.md.integrate_first(self.dt, stream=st)
            moved = set(self.md.rebuild_async(need, stream=st))
        else:
            self.md.integrate_first(self.dt, stream=st)
            moved = set(self.md.update_nlists(stream=st))
        for i, r in enumerate(self.active):
            if held[i] and i not in moved:
                r.nlist._publish(r.state)  # a reused list keeps its arrays: point the new State at them
                moved.add(i)
        for i in moved:
            Batch.refresh(self._esp_args, i)
        t2 = time.perf_counter()
        self.batch.enqueue_prepared(self._esp_args, stream=st)
        t3 = time.perf_counter()
        step_no = len(self.tick_ms) + 1
        self.md.wca(WCA_SIGMA, 1.0, stream=st)
        self.md.integrate_second(self.dt, step=step_no, stream=st)
        if self.predict_lists:
            self.md.predict(self.dt, stream=st)
        t4 = time.perf_counter()
        ph["members"] += t1 - t0
        ph["nlists"] += t2 - t1
        ph["esp"] += t3 - t2
        ph["step_rest"] += t4 - t3
        self._inflight = (list(self.active), t0, step_no, held)

    def tick(self, now: float) -> None:
        """One MD step of every resident request, waited for."""
        self._launch()
        self._complete()

    def run(self, time_limit: float = 1e9) -> dict:
        """Serve until every submitted request is done (or time_limit seconds pass)."""
        start = time.perf_counter()
        if self._run_start is None:
            self._run_start = start
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

    def rate_after(self, t0: float, t1: float) -> dict:
        """Sustained rates over the ticks that completed between t0 and t1 seconds after the
        first run() began (a server filling up from empty is excluded by a t0 past its ramp)."""
        sel = [k for k, t in enumerate(self.tick_end) if t0 < t <= t1]
        span = (self.tick_end[sel[-1]] - self.tick_end[sel[0] - 1]) if sel and sel[0] > 0 else 0.0
        if not sel or span <= 0:
            return {"particle_steps_per_s": 0.0, "solves_per_s": 0.0, "mean_active": 0.0, "ticks": 0}
        return {"particle_steps_per_s": sum(self.tick_particles[k] for k in sel) / span,
                "solves_per_s": sum(self.tick_active[k] for k in sel) / span,
                "mean_active": float(np.mean([self.tick_active[k] for k in sel])), "ticks": len(sel)}

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
            "requests_failed": len(self.failed),
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
            "steps_repeated_for_list_overflow": self.redone_steps,
            "contexts_reused": self.contexts_reused,
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
