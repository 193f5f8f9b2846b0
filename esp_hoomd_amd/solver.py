"""Python handle over one ESP context of the C ABI, with device-resident system state.

``Solver`` owns an ``esp_ctx``; ``DeviceSystem`` owns the device copies of positions,
charges, neighbour and exclusion lists and the output arrays.  Both are thin: every
computation runs in the HIP library.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .device import DeviceArray
from .nlist import build_exclusion_list, build_full_nlist


@dataclass
class Params:
    """Arguments of ESPForceCompute::setParams plus the MI355X options."""

    nx: int = 32
    ny: int = 32
    nz: int = 32
    order: int = 6
    kappa: float = 1.0
    rcut: float = 3.0
    alpha: float = 0.0
    n_table: int = 4096
    precision: int = _lib.ESP_PREC_FP64
    split: int = _lib.ESP_SPLIT_PSWF
    window_c: float = 0.0
    window_degree: int = 0
    pair_lanes: int = 0
    k_cut: float = 0.0
    tile: int = 0

    def to_c(self) -> _lib.EspParams:
        p = _lib.EspParams()
        p.nx, p.ny, p.nz = int(self.nx), int(self.ny), int(self.nz)
        p.order = int(self.order)
        p.kappa, p.rcut, p.alpha = float(self.kappa), float(self.rcut), float(self.alpha)
        p.n_table = int(self.n_table)
        p.precision, p.split = int(self.precision), int(self.split)
        p.window_c, p.window_degree = float(self.window_c), int(self.window_degree)
        p.pair_lanes = int(self.pair_lanes)
        p.k_cut = float(self.k_cut)
        p.tile = int(self.tile)
        return p


def potential_pair_forces(system: "DeviceSystem", param: _lib.EspPairParam, r_cut: float, energy_shift: bool = False):
    """HOOMD's generic pair loop over EvaluatorPairPSWF (esp_potential_pair_forces); host copies
    of force (N, 4) and virial (6, N)."""
    s = system.c_system()
    o = system.c_out(_lib.ESP_COMPUTE_PAIR)
    _lib.check(_lib.load().esp_potential_pair_forces(C.byref(s), C.byref(param), float(r_cut), int(energy_shift), C.byref(o), None))
    _lib.check(_lib.load().esp_device_synchronize())
    return {"force": system.force.numpy(), "virial": system.virial.numpy()[:, : system.N]}


class DeviceSystem:
    """Device-resident particle state in the HOOMD layout."""

    def __init__(self, pos, charge, box, r_list=None, exclusions=None, group=None, types=None):
        pos = np.asarray(pos, dtype=np.float64)
        self.N = len(pos)
        self.box = np.array(list(box[:3]) + [0.0, 0.0, 0.0] if len(box) == 3 else list(box), dtype=np.float64)
        pt = np.zeros((self.N, 4), dtype=np.float64)
        pt[:, :3] = pos
        if types is not None:
            pt[:, 3] = types
        self.postype = DeviceArray.from_numpy(pt)
        self.charge = DeviceArray.from_numpy(np.asarray(charge, dtype=np.float64))
        self.force = DeviceArray((self.N, 4), np.float64, zero=True)
        self.virial_pitch = max(self.N, 1)
        self.virial = DeviceArray((6, self.virial_pitch), np.float64, zero=True)
        self.ext = DeviceArray((8,), np.float64, zero=True)
        self.group = None
        self.group_size = 0
        if group is not None:
            g = np.asarray(group, dtype=np.uint32)
            self.group = DeviceArray.from_numpy(g)
            self.group_size = len(g)
        self.n_neigh = self.nlist = self.head = None
        self.n_ex = self.ex_list = None
        self.ex_pitch = 0
        self.exclusions = list(exclusions) if exclusions is not None else None
        if exclusions is not None and len(self.exclusions):
            n_ex, ex_list, pitch = build_exclusion_list(self.N, self.exclusions)
            self.n_ex = DeviceArray.from_numpy(n_ex)
            self.ex_list = DeviceArray.from_numpy(ex_list)
            self.ex_pitch = pitch
        if r_list is not None:
            self.set_nlist(*build_full_nlist(pos, self.box, r_list, self.exclusions))

    def set_nlist(self, n_neigh, nlist, head):
        self.n_neigh = DeviceArray.from_numpy(np.asarray(n_neigh, np.uint32))
        self.nlist = DeviceArray.from_numpy(np.asarray(nlist, np.uint32) if len(nlist) else np.zeros(1, np.uint32))
        self.head = DeviceArray.from_numpy(np.asarray(head, np.uint64))

    def set_positions(self, pos) -> None:
        pt = self.postype.numpy()
        pt[:, :3] = pos
        self.postype.upload(pt)

    def c_system(self) -> _lib.EspSystem:
        s = _lib.EspSystem()
        s.N = self.N
        s.postype = self.postype.ptr
        s.charge = self.charge.ptr
        s.group_size = self.group_size
        s.group_idx = self.group.ptr if self.group is not None else None
        for i in range(6):
            s.box[i] = float(self.box[i])
        if self.nlist is not None:
            s.n_neigh, s.nlist, s.head_list = self.n_neigh.ptr, self.nlist.ptr, self.head.ptr
        if self.n_ex is not None:
            s.n_ex, s.ex_list, s.ex_pitch = self.n_ex.ptr, self.ex_list.ptr, self.ex_pitch
        return s

    def c_out(self, flags: int = 0) -> _lib.EspOut:
        o = _lib.EspOut()
        o.force = self.force.ptr
        o.virial = self.virial.ptr
        o.virial_pitch = self.virial_pitch
        o.ext_energy_virial = self.ext.ptr
        o.flags = flags
        return o


class Solver:
    """One ESP context (ESPForceCompute on the GPU)."""

    def __init__(self, params: Params):
        lib = _lib.load()
        self.params = params
        ctx = C.c_void_p()
        cp = params.to_c()
        rc = lib.esp_create(C.byref(ctx), C.byref(cp))
        if rc != _lib.ESP_OK:
            raise _lib.EspError(rc, lib.esp_last_error(None).decode())
        self._ctx = ctx

    @property
    def ctx(self):
        return self._ctx

    def set_params(self, params: Params) -> None:
        cp = params.to_c()
        _lib.check(_lib.load().esp_set_params(self._ctx, C.byref(cp)), self._ctx)
        self.params = params

    def invalidate(self) -> None:
        _lib.check(_lib.load().esp_invalidate(self._ctx), self._ctx)

    def enqueue(self, system: DeviceSystem, timestep: int = 0, flags: int = 0, stream=None) -> None:
        s = system.c_system()
        o = system.c_out(flags)
        _lib.check(_lib.load().esp_compute_forces(self._ctx, timestep, C.byref(s), C.byref(o),
                                                  stream.handle if stream is not None else None), self._ctx)

    def enqueue_state(self, state, timestep: int = 0, flags: int = 0, stream=None) -> None:
        """Enqueue on an md.State (device arrays including its current neighbour list)."""
        s = state.c_system()
        o = state.c_out(flags)
        _lib.check(_lib.load().esp_compute_forces(self._ctx, timestep, C.byref(s), C.byref(o),
                                                  stream.handle if stream is not None else None), self._ctx)

    def prepare(self, system, stream=None) -> None:
        """Do the set-up the first computation on `system` (a DeviceSystem or an md.State)
        would do -- tables, window, influence function (enqueued on `stream`), scratch --
        without computing, so that it can happen off a step's critical path."""
        s = system.c_system()
        o = system.c_out(0)
        _lib.check(_lib.load().esp_prepare(self._ctx, C.byref(s), C.byref(o), stream.handle if stream is not None else None),
                   self._ctx)

    def compute(self, system: DeviceSystem, timestep: int = 0, flags: int = 0):
        """Run one evaluation and return host copies: force (N,4), virial (6,N), ext (7,)."""
        self.enqueue(system, timestep, flags)
        _lib.check(_lib.load().esp_device_synchronize())
        return {
            "force": system.force.numpy(),
            "virial": system.virial.numpy()[:, : system.N],
            "ext": system.ext.numpy()[:7],
        }

    def pair_forces(self, system: DeviceSystem):
        s = system.c_system()
        o = system.c_out(_lib.ESP_COMPUTE_PAIR)
        _lib.check(_lib.load().esp_pair_forces(self._ctx, C.byref(s), C.byref(o), None), self._ctx)
        _lib.check(_lib.load().esp_device_synchronize())
        return {"force": system.force.numpy(), "virial": system.virial.numpy()[:, : system.N]}

    def set_graph(self, on: bool = True) -> None:
        """Replay the whole step as a captured hipGraph while the call's arguments stay the same
        (see esp_set_graph)."""
        _lib.check(_lib.load().esp_set_graph(self._ctx, 1 if on else 0), self._ctx)

    def graph_stats(self):
        """(captures, replays) of the step graph."""
        c, r = C.c_ulonglong(), C.c_ulonglong()
        _lib.check(_lib.load().esp_get_graph_stats(self._ctx, C.byref(c), C.byref(r)), self._ctx)
        return int(c.value), int(r.value)

    def set_profiling(self, on=True) -> None:
        """True / 1: HIP-event times of every stage; 2: of the pair stage only; False / 0: off."""
        level = 1 if on is True else (0 if on is False else int(on))
        _lib.check(_lib.load().esp_set_profiling(self._ctx, level), self._ctx)

    def reset_stage_times(self) -> None:
        _lib.check(_lib.load().esp_reset_stage_times(self._ctx), self._ctx)

    def stage_times(self):
        """{stage name: (total ms, calls)} since the last reset (synchronises)."""
        lib = _lib.load()
        ms = np.zeros(_lib.ESP_N_STAGES)
        calls = (C.c_uint * _lib.ESP_N_STAGES)()
        _lib.check(lib.esp_get_stage_times(self._ctx, _lib.dptr(ms), calls), self._ctx)
        return {lib.esp_stage_name(i).decode(): (float(ms[i]), int(calls[i])) for i in range(_lib.ESP_N_STAGES)}

    def info(self) -> _lib.EspInfo:
        inf = _lib.EspInfo()
        _lib.check(_lib.load().esp_get_info(self._ctx, C.byref(inf)), self._ctx)
        return inf

    def charge_sums(self):
        q, q2 = C.c_double(), C.c_double()
        _lib.check(_lib.load().esp_get_charge_sums(self._ctx, C.byref(q), C.byref(q2)), self._ctx)
        return q.value, q2.value

    def split_table(self):
        lib = _lib.load()
        n = C.c_uint()
        _lib.check(lib.esp_get_split_table(self._ctx, C.byref(n), None, None, None, None), self._ctx)
        n = n.value
        scr = np.zeros(5 * n)
        dscr = np.zeros(5 * n)
        rlo = np.zeros(n)
        dri = np.zeros(n)
        nn = C.c_uint()
        _lib.check(lib.esp_get_split_table(self._ctx, C.byref(nn), _lib.dptr(scr), _lib.dptr(dscr), _lib.dptr(rlo),
                                           _lib.dptr(dri)), self._ctx)
        return scr.reshape(n, 5), dscr.reshape(n, 5), rlo, dri

    def fixed_point_scale(self):
        """(scale, lowered) of the last step's fixed-point mesh accumulation (both precisions)."""
        sc, low = C.c_float(), C.c_int()
        _lib.check(_lib.load().esp_get_fixed_point_scale(self._ctx, C.byref(sc), C.byref(low)), self._ctx)
        return float(sc.value), bool(low.value)

    def pair_param(self, V_cut: float = 0.0) -> _lib.EspPairParam:
        """EvaluatorPairPSWF::param_type for this context's split (the table addresses stay valid
        until the parameters or table change)."""
        h, d, n = C.c_void_p(), C.c_void_p(), C.c_uint()
        _lib.check(_lib.load().esp_get_pair_table(self._ctx, C.byref(h), C.byref(d), C.byref(n)), self._ctx)
        p = _lib.EspPairParam()
        p.kappa, p.alpha, p.rcut, p.V_cut = self.params.kappa, self.params.alpha, self.params.rcut, V_cut
        p.table_host, p.table_device, p.n_segs = h.value, d.value, n.value
        return p

    def set_split_table(self, screen, dscreen, r_lo, dr_inv) -> None:
        screen = np.ascontiguousarray(screen, np.float64).reshape(-1)
        dscreen = np.ascontiguousarray(dscreen, np.float64).reshape(-1)
        r_lo = np.ascontiguousarray(r_lo, np.float64)
        dr_inv = np.ascontiguousarray(dr_inv, np.float64)
        _lib.check(_lib.load().esp_set_split_table(self._ctx, len(r_lo), _lib.dptr(screen), _lib.dptr(dscreen),
                                                   _lib.dptr(r_lo), _lib.dptr(dr_inv)), self._ctx)

    def window(self):
        lib = _lib.load()
        P, D = C.c_uint(), C.c_uint()
        _lib.check(lib.esp_get_window(self._ctx, C.byref(P), C.byref(D), None), self._ctx)
        coef = np.zeros(P.value * (D.value + 1))
        _lib.check(lib.esp_get_window(self._ctx, C.byref(P), C.byref(D), _lib.dptr(coef)), self._ctx)
        return P.value, D.value, coef.reshape(P.value, D.value + 1)

    def table_entry(self, i: int) -> _lib.EspTableEntry:
        e = _lib.EspTableEntry()
        _lib.check(_lib.load().esp_get_table_entry(self._ctx, int(i), C.byref(e)), self._ctx)
        return e

    def close(self) -> None:
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            _lib.load().esp_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- host-only mathematics (no GPU) ------------------------------------------------------
def build_split_table(split: int, kappa: float, rcut: float, n_segs: int = 4096):
    lib = _lib.load()
    scr = np.zeros(5 * n_segs)
    dscr = np.zeros(5 * n_segs)
    rlo = np.zeros(n_segs)
    dri = np.zeros(n_segs)
    _lib.check(lib.esp_build_split_table(split, kappa, rcut, n_segs, _lib.dptr(scr), _lib.dptr(dscr), _lib.dptr(rlo),
                                         _lib.dptr(dri)))
    return scr.reshape(n_segs, 5), dscr.reshape(n_segs, 5), rlo, dri


def split_eval(split: int, kappa: float, rcut: float, s):
    s = np.ascontiguousarray(s, np.float64)
    S = np.zeros_like(s)
    dS = np.zeros_like(s)
    _lib.check(_lib.load().esp_split_eval(split, kappa, rcut, len(s), _lib.dptr(s), _lib.dptr(S), _lib.dptr(dS)))
    return S, dS


def split_fourier(split: int, kappa: float, rcut: float, k):
    k = np.ascontiguousarray(k, np.float64)
    F = np.zeros_like(k)
    dF = np.zeros_like(k)
    _lib.check(_lib.load().esp_split_fourier(split, kappa, rcut, len(k), _lib.dptr(k), _lib.dptr(F), _lib.dptr(dF)))
    return F, dF


def split_constants(split: int, kappa: float, rcut: float):
    a, b = C.c_double(), C.c_double()
    _lib.check(_lib.load().esp_split_constants(split, kappa, rcut, C.byref(a), C.byref(b)))
    return a.value, b.value


def prolate_eval(c: float, x):
    x = np.ascontiguousarray(x, np.float64)
    psi = np.zeros_like(x)
    integ = np.zeros_like(x)
    _lib.check(_lib.load().esp_prolate_eval(c, len(x), _lib.dptr(x), _lib.dptr(psi), _lib.dptr(integ)))
    return psi, integ


def build_window(P: int, D: int, window_c: float):
    coef = np.zeros(P * (D + 1))
    _lib.check(_lib.load().esp_build_window(P, D, window_c, _lib.dptr(coef)))
    return coef.reshape(P, D + 1)


def window_fourier(coef: np.ndarray, theta):
    coef = np.ascontiguousarray(coef, np.float64)
    P, D1 = coef.shape
    theta = np.ascontiguousarray(theta, np.float64)
    out = np.zeros_like(theta)
    _lib.check(_lib.load().esp_window_fourier(P, D1 - 1, _lib.dptr(coef.reshape(-1)), len(theta), _lib.dptr(theta),
                                              _lib.dptr(out)))
    return out


def default_window(P: int, sigma: float):
    lib = _lib.load()
    return lib.esp_default_window_degree(P), lib.esp_default_window_c(P, sigma)
