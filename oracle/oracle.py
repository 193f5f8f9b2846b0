"""ctypes wrapper of the CPU oracle ``oracle/liboracle.so`` -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the CPU-baseline leg of bench.py load this module.
The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
_LIB = _DIR / "liboracle.so"
_lib = None

_DP = C.POINTER(C.c_double)


class EspIn(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("postype", _DP), ("q", _DP), ("L", C.c_double * 3), ("n", C.c_int * 3),
        ("P", C.c_int), ("D", C.c_int), ("wcoef", _DP), ("n_segs", C.c_int), ("screen", _DP), ("dscreen", _DP),
        ("r_lo", _DP), ("dr_inv", _DP), ("rc", C.c_double), ("split_kind", C.c_int), ("kappa", C.c_double),
        ("n_ex", C.POINTER(C.c_uint)), ("ex_list", C.POINTER(C.c_uint)), ("ex_pitch", C.c_uint),
        ("k_cut", C.c_double), ("alpha", C.c_double), ("tilt", C.c_double * 3),
    ]


class EspOut(C.Structure):
    _fields_ = [
        ("force", _DP), ("virial", _DP), ("virial_pitch", C.c_uint), ("ext_energy", C.c_double),
        ("ext_virial", C.c_double * 6), ("mesh", _DP), ("phi", _DP), ("green", _DP),
    ]


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB.exists():
        subprocess.run(["make", "-C", str(_DIR), "-s"], check=True)
    lib = C.CDLL(str(_LIB))
    lib.oracle_fft3d.argtypes = [C.c_int, C.c_int, C.c_int, _DP, C.c_int]
    lib.oracle_fft3d.restype = C.c_int
    lib.oracle_split_eval.argtypes = [_DP, _DP, _DP, _DP, C.c_uint, C.c_double, C.c_int, _DP, _DP, _DP, _DP, _DP]
    lib.oracle_split_eval.restype = None
    lib.oracle_ewald.argtypes = [C.c_int, _DP, _DP, _DP, C.c_double, C.c_double, C.c_double, C.c_int,
                                 C.POINTER(C.c_int), _DP, _DP, _DP]
    lib.oracle_ewald.restype = C.c_int
    lib.oracle_esp.argtypes = [C.POINTER(EspIn), C.POINTER(EspOut)]
    lib.oracle_esp.restype = C.c_int
    lib.oracle_window_fourier2.argtypes = [_DP, C.c_int, C.c_int, C.c_double]
    lib.oracle_window_fourier2.restype = C.c_double
    lib.oracle_window_eval.argtypes = [_DP, C.c_int, C.c_int, C.c_double, _DP, _DP]
    lib.oracle_window_eval.restype = None
    lib.oracle_split_F.argtypes = [C.POINTER(EspIn), C.c_double, _DP, _DP]
    lib.oracle_split_F.restype = None
    lib.oracle_ewald_box.argtypes = [C.c_int, _DP, _DP, _DP, C.c_double, C.c_double, C.c_double, _DP, _DP, _DP]
    lib.oracle_ewald_box.restype = C.c_int
    lib.oracle_num_threads.argtypes = []
    lib.oracle_num_threads.restype = C.c_int
    _lib = lib
    return lib


def num_threads() -> int:
    """OpenMP threads the oracle's parallel loops use."""
    return int(load().oracle_num_threads())


def _p(a):
    return a.ctypes.data_as(_DP)


def fft3d(a: np.ndarray, sign: int = -1) -> np.ndarray:
    """3-D FFT of a complex array shaped (nz, ny, nx) (x fastest)."""
    b = np.ascontiguousarray(a, dtype=np.complex128).copy()
    nz, ny, nx = b.shape
    load().oracle_fft3d(nx, ny, nz, b.view(np.float64).ctypes.data_as(_DP), sign)
    return b


def split_eval(screen, dscreen, r_lo, dr_inv, rc, r):
    screen = np.ascontiguousarray(screen, np.float64).reshape(-1)
    dscreen = np.ascontiguousarray(dscreen, np.float64).reshape(-1)
    r_lo = np.ascontiguousarray(r_lo, np.float64)
    dr_inv = np.ascontiguousarray(dr_inv, np.float64)
    r = np.ascontiguousarray(r, np.float64)
    n = len(r)
    L, nd, S, dS = (np.zeros(n) for _ in range(4))
    load().oracle_split_eval(_p(screen), _p(dscreen), _p(r_lo), _p(dr_inv), len(r_lo), rc, n, _p(r), _p(L), _p(nd),
                             _p(S), _p(dS))
    return L, nd, S, dS


def ewald(pos, q, box, alpha, rcut, kcut, targets=None, energy=True):
    """FP64 Ewald sum.  Returns (forces (n_t,3), energy or None, virial (6,) or None)."""
    pos = np.asarray(pos, np.float64)
    N = len(pos)
    pt = np.zeros((N, 4))
    pt[:, :3] = pos
    q = np.ascontiguousarray(q, np.float64)
    L = np.ascontiguousarray(box[:3], np.float64)
    if targets is None:
        tptr = None
        nt = N
    else:
        targets = np.ascontiguousarray(targets, np.int32)
        tptr = targets.ctypes.data_as(C.POINTER(C.c_int))
        nt = len(targets)
    f = np.zeros((nt, 3))
    e = np.zeros(1)
    w = np.zeros(6)
    load().oracle_ewald(N, _p(pt), _p(q), _p(L), alpha, rcut, kcut, nt, tptr, _p(f),
                        _p(e) if energy else None, _p(w) if energy else None)
    return f, (e[0] if energy else None), (w if energy else None)


def box_half_width(box):
    """Half the smallest perpendicular width of a HOOMD box (Lx, Ly, Lz, xy, xz, yz)."""
    Lx, Ly, Lz, xy, xz, yz = (list(box) + [0.0, 0.0, 0.0])[:6]
    n1 = np.sqrt(1.0 + xy * xy + (xy * yz - xz) ** 2) / Lx
    n2 = np.sqrt(1.0 + yz * yz) / Ly
    return 0.5 / max(n1, n2, 1.0 / Lz)


def ewald_box(pos, q, box, alpha, rcut, kcut):
    """FP64 Ewald sum in a general (triclinic) HOOMD box (Lx, Ly, Lz, xy, xz, yz): nearest-image
    real space (rcut below half the smallest width), k = 2 pi B^-T m.  Returns (forces, energy,
    virial)."""
    pos = np.asarray(pos, np.float64)
    N = len(pos)
    pt = np.zeros((N, 4))
    pt[:, :3] = pos
    q = np.ascontiguousarray(q, np.float64)
    b = np.ascontiguousarray((list(box) + [0.0, 0.0, 0.0])[:6], np.float64)
    f = np.zeros((N, 3))
    e = np.zeros(1)
    w = np.zeros(6)
    if load().oracle_ewald_box(N, _p(pt), _p(q), _p(b), alpha, rcut, kcut, _p(f), _p(e), _p(w)) != 0:
        raise ValueError("rcut must be below half the smallest perpendicular width of the box")
    return f, e[0], w


def ewald_box_auto(pos, q, box, tol=1e-15, rcut=None):
    """ewald_box with parameters chosen for ~tol relative accuracy."""
    if rcut is None:
        rcut = 0.99 * box_half_width(box)
    s = np.sqrt(-np.log(tol))
    alpha = s / rcut
    return ewald_box(pos, q, box, alpha, rcut, 2.0 * alpha * s)


def ewald_auto(pos, q, box, tol=1e-15, targets=None, energy=True, rcut=None):
    """Ewald with parameters chosen for ~tol relative accuracy."""
    L = np.asarray(box[:3], np.float64)
    if rcut is None:
        rcut = 0.49 * float(L.min())
    s = np.sqrt(-np.log(tol))
    alpha = s / rcut
    kcut = 2.0 * alpha * s
    return ewald(pos, q, box, alpha, rcut, kcut, targets=targets, energy=energy)


class OracleESP:
    """Holds the tables of one ESP configuration and evaluates it on the CPU."""

    def __init__(self, box, mesh, P, wcoef, screen, dscreen, r_lo, dr_inv, rc, split_kind=0, kappa=0.0, k_cut=0.0, alpha=0.0):
        self.L = np.asarray(box[:3], np.float64)
        self.tilt = np.asarray((list(box) + [0.0, 0.0, 0.0])[3:6], np.float64)
        self.mesh = tuple(int(m) for m in mesh)
        self.P = int(P)
        self.wcoef = np.ascontiguousarray(wcoef, np.float64)
        self.D = self.wcoef.shape[1] - 1
        self.screen = np.ascontiguousarray(screen, np.float64).reshape(-1)
        self.dscreen = np.ascontiguousarray(dscreen, np.float64).reshape(-1)
        self.r_lo = np.ascontiguousarray(r_lo, np.float64)
        self.dr_inv = np.ascontiguousarray(dr_inv, np.float64)
        self.rc = float(rc)
        self.split_kind = int(split_kind)
        self.kappa = float(kappa)
        self.k_cut = float(k_cut)
        self.alpha = float(alpha)

    def _in(self, pt, q, n_ex=None, ex_list=None, ex_pitch=0):
        s = EspIn()
        s.N = len(q)
        s.postype = _p(pt)
        s.q = _p(q)
        for a in range(3):
            s.L[a] = self.L[a]
            s.n[a] = self.mesh[a]
            s.tilt[a] = self.tilt[a]
        s.P, s.D = self.P, self.D
        s.wcoef = _p(self.wcoef)
        s.n_segs = len(self.r_lo)
        s.screen, s.dscreen, s.r_lo, s.dr_inv = _p(self.screen), _p(self.dscreen), _p(self.r_lo), _p(self.dr_inv)
        s.rc = self.rc
        s.split_kind = self.split_kind
        s.kappa = self.kappa
        s.k_cut = self.k_cut
        s.alpha = self.alpha
        if n_ex is not None:
            s.n_ex = n_ex.ctypes.data_as(C.POINTER(C.c_uint))
            s.ex_list = ex_list.ctypes.data_as(C.POINTER(C.c_uint))
            s.ex_pitch = ex_pitch
        return s

    def compute(self, pos, q, exclusions=None, want_mesh=False):
        pos = np.asarray(pos, np.float64)
        N = len(pos)
        pt = np.zeros((N, 4))
        pt[:, :3] = pos
        q = np.ascontiguousarray(q, np.float64)
        n_ex = ex_list = None
        pitch = 0
        if exclusions is not None and len(exclusions):
            per = [[] for _ in range(N)]
            for a, b in exclusions:
                per[a].append(b)
                per[b].append(a)
            pitch = max(len(x) for x in per)
            n_ex = np.array([len(x) for x in per], np.uint32)
            ex_list = np.zeros((N, pitch), np.uint32)
            for a, lst in enumerate(per):
                ex_list[a, : len(lst)] = lst
            ex_list = ex_list.reshape(-1)
        sin = self._in(pt, q, n_ex, ex_list, pitch)
        force = np.zeros((N, 4))
        vir = np.zeros((6, N))
        out = EspOut()
        out.force = _p(force)
        out.virial = _p(vir)
        out.virial_pitch = N
        M = self.mesh[0] * self.mesh[1] * self.mesh[2]
        mesh = phi = green = None
        if want_mesh:
            mesh, phi, green = np.zeros(M), np.zeros(M), np.zeros(M)
            out.mesh, out.phi, out.green = _p(mesh), _p(phi), _p(green)
        if load().oracle_esp(C.byref(sin), C.byref(out)) != 0:
            raise ValueError("in a tilted box rc must be below half the smallest perpendicular width")
        res = {"force": force, "virial": vir, "ext_energy": out.ext_energy, "ext_virial": np.array(out.ext_virial[:])}
        if want_mesh:
            shp = (self.mesh[2], self.mesh[1], self.mesh[0])
            res.update(mesh=mesh.reshape(shp), phi=phi.reshape(shp), green=green.reshape(shp))
        res["energy"] = force[:, 3].sum() + out.ext_energy
        res["total_virial"] = vir.sum(axis=1) + res["ext_virial"]
        return res

    def split_F(self, k):
        pt = np.zeros((1, 4))
        q = np.zeros(1)
        s = self._in(pt, q)
        F, dF = C.c_double(), C.c_double()
        load().oracle_split_F(C.byref(s), float(k), C.byref(F), C.byref(dF))
        return F.value, dF.value


def window_fourier2(wcoef, theta):
    wcoef = np.ascontiguousarray(wcoef, np.float64)
    P, D1 = wcoef.shape
    return load().oracle_window_fourier2(_p(wcoef), P, D1 - 1, float(theta))
