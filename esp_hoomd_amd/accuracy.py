"""Force-accuracy measurement on the GPU.

The reported error of every result is the relative RMS of the per-particle electrostatic
force against a reference, sqrt(sum |F - F_ref|^2 / sum |F_ref|^2).  The reference is this
library's own FP64 ESP at tight settings (prolate bandwidth c = 20, stencil order 12,
oversampling 1.5), whose own error against a direct Ewald sum is about 2e-9 — three orders
of magnitude below the 1e-5 target — and is itself checked against the Ewald sum in the
test suite.  Everything runs on the device; nothing here calls a CPU solver.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .device import synchronize
from .md import NeighborList, State
from .solver import Params, Solver

REF_C = 20.0        # prolate bandwidth of the reference split
REF_SIGMA = 1.5     # mesh oversampling of the reference
REF_ORDER = 12      # reference stencil order
REF_RC_MAX = 8.0    # largest reference cutoff (it shrinks with small boxes)


def fft_size_at_least(n: int, minimum: int = 8) -> int:
    """Smallest 2^a 3^b 5^c >= max(n, minimum)."""
    m = max(int(n), int(minimum))
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def reference_params(box: Sequence[float]) -> Params:
    """Tight FP64 ESP parameters for a periodic box (Lx, Ly, Lz)."""
    L = np.asarray(box[:3], float)
    rc = min(REF_RC_MAX, 0.45 * float(L.min()))
    kappa = REF_C / rc
    h = math.pi / (REF_SIGMA * kappa)
    n = [fft_size_at_least(math.ceil(Li / h), 2 * REF_ORDER) for Li in L]
    return Params(nx=n[0], ny=n[1], nz=n[2], order=REF_ORDER, kappa=kappa, rcut=rc, precision=_lib.ESP_PREC_FP64)


def reference_forces(positions: np.ndarray, charges: np.ndarray, box: Sequence[float]) -> np.ndarray:
    """Electrostatic forces (N, 3) of the tight FP64 reference, computed on the GPU."""
    st = State(positions, charges, list(box[:3]))
    params = reference_params(box)
    nl = NeighborList(buffer=0.0)
    nl.update(st, params.rcut, force_rebuild=True)
    solver = Solver(params)
    solver.enqueue_state(st)
    synchronize()
    return st.force.numpy()[:, :3].copy()


def rel_rms_error(F: np.ndarray, F_ref: np.ndarray, subset: Optional[np.ndarray] = None) -> float:
    """sqrt(sum |F - F_ref|^2 / sum |F_ref|^2) over all particles or a subset."""
    F = np.asarray(F)[:, :3]
    F_ref = np.asarray(F_ref)[:, :3]
    if subset is not None:
        F = F[subset]
        F_ref = F_ref[subset]
    den = float((F_ref * F_ref).sum())
    return float(np.sqrt(((F - F_ref) ** 2).sum() / den)) if den > 0 else 0.0


def forces_with(params: Params, positions: np.ndarray, charges: np.ndarray, box: Sequence[float],
                skin: float = 0.4) -> np.ndarray:
    """Electrostatic forces (N, 3) of one ESP configuration on the GPU (GPU neighbour list)."""
    st = State(positions, charges, list(box[:3]))
    nl = NeighborList(buffer=skin)
    nl.update(st, params.rcut, force_rebuild=True)
    solver = Solver(params)
    solver.enqueue_state(st)
    synchronize()
    return st.force.numpy()[:, :3].copy()


def measure_error(params: Params, positions: np.ndarray, charges: np.ndarray, box: Sequence[float],
                  reference: Optional[np.ndarray] = None) -> float:
    """Relative RMS force error of `params` on this system against the tight reference."""
    if reference is None:
        reference = reference_forces(positions, charges, box)
    return rel_rms_error(forces_with(params, positions, charges, box), reference)


__all__ = ["fft_size_at_least", "reference_params", "reference_forces", "rel_rms_error", "forces_with",
           "measure_error", "REF_C", "REF_SIGMA", "REF_ORDER"]
