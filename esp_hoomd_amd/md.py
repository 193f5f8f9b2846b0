"""Device-resident molecular dynamics around the ESP force (the roles HOOMD-blue plays for the
reference plugin): particle state, a GPU cell-list neighbour list, the WCA repulsion and
velocity-Verlet (NVE) / Langevin integration.

Every per-step operation runs on the GPU through the C ABI; the host only sequences calls.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceArray, synchronize
from .nlist import build_exclusion_list


class State:
    """Particle data in the HOOMD layout: postype (x, y, z, type), charge, velocity (vx, vy,
    vz, mass), net force (fx, fy, fz, energy) and per-particle virial (6 x pitch)."""

    def __init__(self, positions, charges, box, velocities=None, masses=1.0, types=None, exclusions=None,
                 zero_outputs: bool = True, pooled: bool = False, stream=None, staging=None):
        pos = np.asarray(positions, dtype=np.float64)
        self.N = len(pos)
        box = list(box)
        self.box = np.array(box[:3] + [0.0, 0.0, 0.0] if len(box) == 3 else box, dtype=np.float64)
        pt = np.zeros((self.N, 4))
        pt[:, :3] = pos
        if types is not None:
            pt[:, 3] = types
        v = np.zeros((self.N, 4))
        if velocities is not None:
            v[:, :3] = velocities
        v[:, 3] = masses
        q = np.asarray(charges, dtype=np.float64)
        # pooled: the arrays come from and return to the device block cache (device.DeviceArray)
        # stream + staging (a device.PinnedPool): uploads are enqueued on the stream from pinned
        # blocks, held in self.staging until release_staging() (after the stream passed them)
        self.staging = []
        if stream is not None and staging is not None:
            self.postype = DeviceArray(pt.shape, np.float64, pooled=pooled)
            self.charge = DeviceArray(q.shape, np.float64, pooled=pooled)
            self.vel = DeviceArray(v.shape, np.float64, pooled=pooled)
            self._staging_pool = staging
            for d, a in ((self.postype, pt), (self.charge, q), (self.vel, v)):
                self.staging.append(d.upload_async(a, stream, staging))
        else:
            self.postype = DeviceArray.from_numpy(pt, pooled=pooled)
            self.charge = DeviceArray.from_numpy(q, pooled=pooled)
            self.vel = DeviceArray.from_numpy(v, pooled=pooled)
        # zero_outputs=False leaves force and virial unset (for callers whose first force
        # evaluation overwrites them)
        self.force = DeviceArray((self.N, 4), np.float64, zero=zero_outputs, pooled=pooled)
        self.virial_pitch = max(self.N, 1)
        self.virial = DeviceArray((6, self.virial_pitch), np.float64, zero=zero_outputs, pooled=pooled)
        self.ext = DeviceArray((8,), np.float64, zero=True, pooled=pooled)
        self.exclusions = list(exclusions) if exclusions else None
        self.n_ex = self.ex_list = None
        self.ex_pitch = 0
        if self.exclusions:
            n_ex, ex_list, pitch = build_exclusion_list(self.N, self.exclusions)
            self.n_ex = DeviceArray.from_numpy(n_ex)
            self.ex_list = DeviceArray.from_numpy(ex_list)
            self.ex_pitch = pitch
        self.nlist_ptrs = None  # (n_neigh, nlist, head) device pointers of the current list
        self.inner_ptrs = None  # (n_neigh, nlist, head, r_inner) of the current inner list, if any

    def release_staging(self) -> None:
        """Return the pinned upload blocks (call once the upload stream has passed them)."""
        for blk in self.staging:
            self._staging_pool.release(blk)
        self.staging = []

    @property
    def volume(self) -> float:
        return float(np.prod(self.box[:3]))

    def c_system(self) -> _lib.EspSystem:
        s = _lib.EspSystem()
        s.N = self.N
        s.postype = self.postype.ptr
        s.charge = self.charge.ptr
        for i in range(6):
            s.box[i] = float(self.box[i])
        if self.nlist_ptrs is not None:
            s.n_neigh, s.nlist, s.head_list = self.nlist_ptrs
        if self.n_ex is not None:
            s.n_ex, s.ex_list, s.ex_pitch = self.n_ex.ptr, self.ex_list.ptr, self.ex_pitch
        return s

    def c_out(self, flags=0) -> _lib.EspOut:
        o = _lib.EspOut()
        o.force, o.virial, o.virial_pitch = self.force.ptr, self.virial.ptr, self.virial_pitch
        o.ext_energy_virial = self.ext.ptr
        o.flags = flags
        return o

    def positions(self) -> np.ndarray:
        """Positions in the arrays' current order (see permute)."""
        return self.postype.numpy()[:, :3]

    def in_input_order(self, a: np.ndarray) -> np.ndarray:
        """Rows of a per-particle host array in the current order, put back in the order the
        particles were given in (the identity until permute is called)."""
        if getattr(self, "tag", None) is None:
            return a
        out = np.empty_like(a)
        out[self.tag.numpy().astype(np.int64)] = a
        return out

    def permute(self, order_ptr: int, stream=None) -> None:
        """Reorder the particles: row i of positions, velocities, net forces and charges becomes
        row order[i] (a device permutation of [0, N), e.g. NeighborList.cell_order()), so that
        particles close in space are close in memory again (the role of HOOMD's SFCPack sorter).
        `tag` follows, mapping rows back to the input order.  Neighbour lists must be rebuilt
        afterwards; per-particle virials are not carried (recomputed with the next forces)."""
        if self.exclusions:
            raise ValueError("State.permute: exclusion lists index particles and are not remapped")
        lib = _lib.load()
        if getattr(self, "tag", None) is None:
            self.tag = DeviceArray.from_numpy(np.arange(self.N, dtype=np.uint32))
        spare = getattr(self, "_spare", None) or {}
        h = stream.handle if stream is not None else None
        for name in ("postype", "vel", "force", "charge", "tag"):
            src = getattr(self, name)
            dst = spare.get(name)
            if dst is None or dst.shape != src.shape or dst.dtype != src.dtype:
                dst = DeviceArray(src.shape, src.dtype)
            row = src.nbytes // max(self.N, 1)
            _lib.check(lib.esp_gather_rows(self.N, row, C.c_void_p(order_ptr), src._ptr, dst._ptr, h))
            spare[name] = src
            setattr(self, name, dst)
        self._spare = spare

    def velocities(self) -> np.ndarray:
        return self.vel.numpy()[:, :3]

    def set_positions(self, pos) -> None:
        pt = self.postype.numpy()
        pt[:, :3] = pos
        self.postype.upload(pt)

    def set_velocities(self, v) -> None:
        a = self.vel.numpy()
        a[:, :3] = v
        self.vel.upload(a)

    def thermalize(self, kT: float, seed: int = 1) -> None:
        """Maxwell-Boltzmann velocities at kT with zero net momentum."""
        a = self.vel.numpy()
        a[:, :3] = maxwell_boltzmann(a[:, 3], kT, seed)
        self.vel.upload(a)


class NeighborList:
    """GPU cell-list neighbour list (full list, HOOMD layout) with a skin `buffer`; rebuilt when
    a particle has moved more than buffer / 2 since the last build.  With `inner` > 0 each build
    also keeps the pairs within inner + buffer in a second list, which short-ranged forces
    (WCA) with a cutoff <= inner walk instead of the full list."""

    def __init__(self, buffer: float = 0.4, inner: float = 0.0):
        if buffer < 0 or inner < 0:
            raise ValueError("NeighborList: buffer and inner must be >= 0")
        self.buffer = float(buffer)
        self.inner = float(inner)
        self.r_cut = 0.0
        self._nl = None

    def _ensure(self, r_cut: float):
        lib = _lib.load()
        if self._nl is None or r_cut > self.r_cut:
            if self._nl is not None:
                lib.esp_nlist_destroy(self._nl)
            nl = C.c_void_p()
            _lib.check(lib.esp_nlist_create(C.byref(nl), r_cut, self.buffer))
            self._nl = nl
            self.r_cut = r_cut
            if self.inner > 0:
                _lib.check(lib.esp_nlist_set_inner(self._nl, self.inner))

    def set_inner(self, inner: float) -> None:
        """Change the inner radius (0: none); the next update rebuilds."""
        if inner < 0:
            raise ValueError("NeighborList: inner must be >= 0")
        if float(inner) != self.inner:
            self.inner = float(inner)
            if self._nl is not None:
                _lib.check(_lib.load().esp_nlist_set_inner(self._nl, self.inner))

    def _publish(self, state: "State") -> None:
        """Point the state at the current list (and inner list) arrays."""
        lib = _lib.load()
        a, b, c = C.c_void_p(), C.c_void_p(), C.c_void_p()
        nmax = C.c_uint()
        _lib.check(lib.esp_nlist_arrays(self._nl, C.byref(a), C.byref(b), C.byref(c), C.byref(nmax)))
        state.nlist_ptrs = (a.value, b.value, c.value)
        self.nmax = nmax.value
        state.inner_ptrs = None
        self.nmax_inner = 0
        if 0 < self.inner < self.r_cut:
            _lib.check(lib.esp_nlist_inner_arrays(self._nl, C.byref(a), C.byref(b), C.byref(c), C.byref(nmax)))
            state.inner_ptrs = (a.value, b.value, c.value, self.inner)
            self.nmax_inner = nmax.value

    def reset(self, r_cut: float) -> None:
        """Make the list one for pair cutoff r_cut (rebuilt at the next update), keeping its
        arrays for reuse."""
        if self._nl is None:
            self._ensure(r_cut)
            return
        _lib.check(_lib.load().esp_nlist_set_cutoff(self._nl, float(r_cut)))
        self.r_cut = float(r_cut)

    def set_capacity(self, nmax: int, nmax_inner: int = 0) -> None:
        """Row capacities of the next builds (0: sized from the density); an overflowing build
        raises them."""
        if self._nl is None:
            raise RuntimeError("NeighborList.set_capacity: no list yet (call update first)")
        _lib.check(_lib.load().esp_nlist_set_capacity(self._nl, int(nmax), int(nmax_inner)))

    def update(self, state: State, r_cut: float, force_rebuild: bool = False, stream=None) -> bool:
        """Rebuild the list if needed (or always, with force_rebuild) on `stream` (None: the null
        stream); returns whether it was rebuilt.  Synchronises with the stream."""
        lib = _lib.load()
        self._ensure(r_cut)
        s = state.c_system()
        rebuilt = C.c_int(0)
        h = stream.handle if stream is not None else None
        if force_rebuild:
            rc = lib.esp_nlist_build(self._nl, C.byref(s), h)
            rebuilt.value = 1
        else:
            rc = lib.esp_nlist_update(self._nl, C.byref(s), C.byref(rebuilt), h)
        if rc != _lib.ESP_OK:
            raise _lib.EspError(rc, lib.esp_nlist_last_error(self._nl).decode())
        self._publish(state)
        return bool(rebuilt.value)

    def cell_order(self) -> int:
        """Device pointer to the particle indices in cell order of the last build."""
        p, n = C.c_void_p(), C.c_uint()
        rc = _lib.load().esp_nlist_cell_order(self._nl, C.byref(p), C.byref(n))
        if rc != _lib.ESP_OK:
            raise _lib.EspError(rc, _lib.load().esp_nlist_last_error(self._nl).decode())
        return p.value

    @property
    def builds(self) -> int:
        n = C.c_ulonglong()
        _lib.check(_lib.load().esp_nlist_stats(self._nl, C.byref(n)))
        return int(n.value)

    def host_arrays(self, state: State, inner: bool = False):
        """(n_neigh, [neighbour indices per particle]) of the list (or the inner list) copied to
        the host (for tests)."""
        import ctypes
        ptrs, width = (state.inner_ptrs, self.nmax_inner) if inner else (state.nlist_ptrs, self.nmax)
        if ptrs is None:
            raise RuntimeError("NeighborList.host_arrays: no such list")
        n = np.empty(state.N, np.uint32)
        _lib.check(_lib.load().esp_memcpy_dtoh(n.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptrs[0]), n.nbytes))
        flat = np.empty(state.N * width, np.uint32)
        _lib.check(_lib.load().esp_memcpy_dtoh(flat.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptrs[1]), flat.nbytes))
        flat = flat.reshape(state.N, width)
        return n, [flat[i, : n[i]] for i in range(state.N)]

    def __del__(self):
        try:
            if self._nl is not None:
                _lib.load().esp_nlist_destroy(self._nl)
        except Exception:
            pass


class WCA:
    """Weeks-Chandler-Andersen repulsion: LJ truncated at 2^(1/6) sigma and shifted by epsilon."""

    def __init__(self, sigma: float = 1.0, epsilon: float = 1.0):
        self.sigma = float(sigma)
        self.epsilon = float(epsilon)

    @property
    def r_cut(self) -> float:
        return 2.0 ** (1.0 / 6.0) * self.sigma

    def add_to(self, state: State, stream=None) -> None:
        s = state.c_system()
        if state.inner_ptrs is not None and self.r_cut <= state.inner_ptrs[3]:
            s.n_neigh, s.nlist, s.head_list = state.inner_ptrs[:3]
        o = state.c_out()
        _lib.check(_lib.load().esp_wca_forces(C.byref(s), self.sigma, self.epsilon, C.byref(o),
                                              stream.handle if stream else None))


class ESPForce:
    """The whole ESP force of one Solver (pair correction, mesh part, energy and virial) as a
    Simulation force: one esp_compute_forces call per step that overwrites the net force."""

    def __init__(self, solver):
        self.solver = solver
        self.r_cut = float(solver.params.rcut)

    def _esp_flags(self) -> int:
        return _lib.ESP_COMPUTE_ALL

    def _esp_solver(self):
        return self.solver


class NVE:
    """Velocity-Verlet integration."""

    gamma = 0.0
    kT = 0.0
    seed = 0


class Langevin:
    """Velocity Verlet with a Langevin bath (drag gamma, temperature kT, counter-based RNG)."""

    def __init__(self, kT: float, gamma: float = 1.0, seed: int = 4):
        if kT < 0 or gamma <= 0:
            raise ValueError("Langevin: kT must be >= 0 and gamma > 0")
        self.kT, self.gamma, self.seed = float(kT), float(gamma), int(seed)


class Simulation:
    """Minimal MD driver: forces (ESP Spectral/Local, WCA) on a shared neighbour list and one
    integration method, all device resident."""

    def __init__(self, state: State, dt: float, method=None, nlist: NeighborList | None = None, sort_every: int = 0):
        """sort_every > 0: every that many steps the particles are put back in cell order
        (State.permute with the list's cell order) and the list is rebuilt, as HOOMD's SFCPack
        updater keeps its particles sorted; neighbours in space stay neighbours in memory, which
        the pair, binning and spreading passes rely on."""
        if dt <= 0:
            raise ValueError("Simulation: dt must be > 0")
        if sort_every and state.exclusions:
            raise ValueError("Simulation: sort_every needs a state without exclusions")
        self.sort_every = int(sort_every)
        self._force_rebuild = False
        self.sorts = 0
        self.state = state
        self.dt = float(dt)
        self.method = method if method is not None else NVE()
        self.nlist = nlist if nlist is not None else NeighborList(0.4)
        self.forces = []
        self.timestep = 0
        self._forces_valid = False

    def add_force(self, f) -> None:
        self.forces.append(f)
        if hasattr(f, "_attach"):
            f._attach(self)
        self._forces_valid = False

    def _r_cut(self) -> float:
        rc = 0.0
        for f in self.forces:
            rc = max(rc, getattr(f, "r_cut", 0.0) or 0.0)
        return rc

    def compute_forces(self) -> None:
        st = self.state
        if self.forces:
            # WCA walks an inner list when a longer-ranged force sets the list's cutoff
            wca = [f.r_cut for f in self.forces if isinstance(f, WCA)]
            self.nlist.set_inner(max(wca) if wca and max(wca) < self._r_cut() else 0.0)
            self.nlist.update(st, self._r_cut(), force_rebuild=self._force_rebuild)
            self._force_rebuild = False
        # the ESP objects (Spectral mesh part, Local pair part) share one solver: one call with
        # the union of their flags overwrites the net force; other forces then add to it
        flags = 0
        solver = None
        for f in self.forces:
            if hasattr(f, "_esp_flags"):
                flags |= f._esp_flags()
                solver = solver or f._esp_solver()
        if solver is not None:
            solver.enqueue_state(st, self.timestep, flags)
        else:
            st.force.zero()
            st.virial.zero()
            st.ext.zero()
        for f in self.forces:
            if isinstance(f, WCA):
                f.add_to(st)
        self._forces_valid = True

    def run(self, steps: int) -> None:
        lib = _lib.load()
        st = self.state
        if not self._forces_valid:
            self.compute_forces()
        box = np.ascontiguousarray(st.box[:3])
        for _ in range(int(steps)):
            if self.sort_every and self.timestep % self.sort_every == 0 and self.nlist._nl is not None and self.nlist.builds > 0:
                st.permute(self.nlist.cell_order())
                self._force_rebuild = True
                self.sorts += 1
            _lib.check(lib.esp_integrate_first(st.N, st.postype.ptr, st.vel.ptr, st.force.ptr, self.dt, _lib.dptr(box), None))
            self.timestep += 1
            self.compute_forces()
            m = self.method
            _lib.check(lib.esp_integrate_second(st.N, st.vel.ptr, st.force.ptr, self.dt, m.gamma, m.kT, m.seed,
                                                self.timestep, None))
        synchronize()

    def energies(self):
        """(kinetic, potential) with the ESP external (mesh, self, background) energy included."""
        ke, pe = C.c_double(), C.c_double()
        _lib.check(_lib.load().esp_kinetic_potential(self.state.N, self.state.vel.ptr, self.state.force.ptr,
                                                      C.byref(ke), C.byref(pe)))
        ext = self.state.ext.numpy()[0] if any(hasattr(f, "_esp_flags") for f in self.forces) else 0.0
        return ke.value, pe.value + ext

    def temperature(self) -> float:
        ke, _ = self.energies()
        return 2.0 * ke / (3.0 * (self.state.N - 1)) if self.state.N > 1 else 0.0


def maxwell_boltzmann(masses, kT: float, seed: int = 1) -> np.ndarray:
    """(N, 3) Maxwell-Boltzmann velocities at kT (seeded) with zero net momentum."""
    m = np.broadcast_to(np.asarray(masses, np.float64), (len(np.atleast_1d(masses)),))
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((len(m), 3)) * np.sqrt(kT / m)[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    return v


def jittered_lattice(N: int, rho: float, seed: int = 2, jitter: float = 0.1):
    """Seeded jittered simple-cubic lattice of N = m^3 sites at density rho (box centred on 0)."""
    m = round(N ** (1.0 / 3.0))
    if m ** 3 != N:
        raise ValueError("N must be a perfect cube")
    L = (N / rho) ** (1.0 / 3.0)
    a = L / m
    g = (np.arange(m) + 0.5) * a - L / 2
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    pos += np.random.default_rng(seed).uniform(-jitter * a, jitter * a, pos.shape)
    return pos, L


def random_packing(N: int, L: float, min_dist: float, seed: int = 0):
    """Uniform positions in a cubic box with every pair at least min_dist apart (rejection)."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-L / 2, L / 2, (N, 3))
    for _ in range(1000):
        t = cKDTree(pos + L / 2, boxsize=L)
        pairs = t.query_pairs(min_dist, output_type="ndarray")
        if len(pairs) == 0:
            return pos
        bad = np.unique(pairs[:, 1])
        pos[bad] = rng.uniform(-L / 2, L / 2, (len(bad), 3))
    raise RuntimeError("random_packing: could not place particles")


class BatchMD:
    """Velocity Verlet, WCA and neighbour-list upkeep for many States, one launch each."""

    def __init__(self):
        b = C.c_void_p()
        _lib.check(_lib.load().esp_md_batch_create(C.byref(b)))
        self._b = b
        self.members: list = []

    def _check(self, rc):
        if rc != _lib.ESP_OK:
            raise _lib.EspError(rc, _lib.load().esp_md_batch_last_error(self._b).decode())

    def set(self, members, seeds=None, stream=None) -> None:
        """members: list of (State, NeighborList or None).  With a stream the descriptor uploads
        are enqueued on it without waiting (synchronise with it before the next set)."""
        n = len(members)
        systems = (_lib.EspSystem * max(n, 1))()
        vel = (C.c_void_p * max(n, 1))()
        outs = (_lib.EspOut * max(n, 1))()
        nls = (C.c_void_p * max(n, 1))()
        sd = (C.c_uint64 * max(n, 1))()
        for i, (st, nl) in enumerate(members):
            systems[i] = st.c_system()
            vel[i] = st.vel.ptr
            outs[i] = st.c_out()
            nls[i] = nl._nl.value if (nl is not None and nl._nl is not None) else None
            sd[i] = int(seeds[i]) if seeds is not None else 0
        if stream is not None:
            self._check(_lib.load().esp_md_batch_set_async(self._b, n, systems, vel, outs, nls, sd, stream.handle))
        else:
            self._check(_lib.load().esp_md_batch_set(self._b, n, systems, vel, outs, nls, sd))
        self.members = list(members)

    def set_hold(self, hold, stream=None) -> None:
        """Members flagged in `hold` are not integrated by the next integrate calls."""
        n = len(self.members)
        h = (C.c_ubyte * max(n, 1))(*[1 if x else 0 for x in hold])
        self._check(_lib.load().esp_md_batch_set_hold(self._b, h, stream.handle if stream else None))

    def integrate_first(self, dt: float, stream=None) -> None:
        self._check(_lib.load().esp_md_batch_integrate_first(self._b, dt, stream.handle if stream else None))

    def integrate_second(self, dt: float, step: int, gamma: float = 0.0, kT: float = 0.0, stream=None) -> None:
        self._check(_lib.load().esp_md_batch_integrate_second(self._b, dt, gamma, kT, int(step), stream.handle if stream else None))

    def wca(self, sigma: float = 1.0, epsilon: float = 1.0, stream=None) -> None:
        self._check(_lib.load().esp_md_batch_wca(self._b, sigma, epsilon, stream.handle if stream else None))

    def update_nlists(self, stream=None) -> list:
        """Rebuild the lists of members that moved more than buffer / 2; returns the indices
        rebuilt (their States' list pointers are refreshed)."""
        n = len(self.members)
        rebuilt = C.c_uint(0)
        which = (C.c_ubyte * max(n, 1))()
        self._check(_lib.load().esp_md_batch_update_nlists(self._b, C.byref(rebuilt), which, stream.handle if stream else None))
        out = [i for i in range(n) if which[i]]
        for i in out:
            st, nl = self.members[i]
            nl._publish(st)
        return out

    # ---- list upkeep without a wait inside the step -------------------------------------
    def predict(self, dt: float, stream=None) -> None:
        """Enqueue the predicted displacement check for the next first half step at dt (its
        result reaches the host with the stream; read it after synchronising)."""
        self._check(_lib.load().esp_md_batch_predict(self._b, float(dt), stream.handle if stream else None))

    def read_predicted(self) -> list:
        """Members whose list the next step must rebuild (after synchronising with the step that
        ran predict())."""
        n = len(self.members)
        need = (C.c_ubyte * max(n, 1))()
        self._check(_lib.load().esp_md_batch_read_predicted(self._b, need))
        return [bool(need[i]) for i in range(n)]

    def rebuild_async(self, which, stream=None) -> list:
        """Enqueue the rebuilds of the members flagged in `which` (no wait); returns the members
        whose list arrays moved (their States are re-pointed here)."""
        n = len(self.members)
        w = (C.c_ubyte * max(n, 1))(*[1 if x else 0 for x in which])
        moved = (C.c_ubyte * max(n, 1))()
        self._check(_lib.load().esp_md_batch_rebuild_async(self._b, w, moved, stream.handle if stream else None))
        out = [i for i in range(n) if moved[i]]
        for i in out:
            st, nl = self.members[i]
            nl._publish(st)
        return out

    def finish_rebuilds(self) -> list:
        """After synchronising with the step that ran rebuild_async: the members whose rebuilt
        list overflowed its rows (their step must be repeated)."""
        n = len(self.members)
        over = (C.c_ubyte * max(n, 1))()
        cnt = C.c_uint(0)
        self._check(_lib.load().esp_md_batch_finish_rebuilds(self._b, over, C.byref(cnt)))
        return [i for i in range(n) if over[i]]

    def __del__(self):
        try:
            if self._b.value:
                _lib.load().esp_md_batch_destroy(self._b)
        except Exception:
            pass


__all__ = ["State", "NeighborList", "WCA", "ESPForce", "NVE", "Langevin", "Simulation", "BatchMD", "jittered_lattice",
           "random_packing", "maxwell_boltzmann"]
