"""Python surface of the ESP plugin: ``Spectral`` (long-range mesh part, ESPForceCompute),
``Local`` (short-range PSWF pair part, PotentialPair<EvaluatorPairPSWF>) and
``make_esp_forces``.

Mirrors the reference's ``esp.py`` (src/esp.py:62-712): the same constructor arguments,
defaults, validation rules, properties, ``set_params`` / ``invalidate`` behaviour, error types
and the ``Coulomb`` / ``Pair`` aliases.  The objects attach to this package's device-resident
``md.Simulation`` (HOOMD-blue is not required); all computation runs in the HIP library.

Documented differences from the reference: ``MAX_ORDER`` is 16 (the reference's Python layer
caps at 8, esp.py:87, while its C++ allows 12, ESPForceCompute.h:63); ``kappa`` is the band
limit of the long-range part in 1/length, i.e. the prolate bandwidth is c = kappa * r_cut.
``alpha`` (Debye screening) multiplies the long-range influence function by exp(-alpha |k|), as
the reference's GPU influence build does (ESPForceComputeGPU.cu:199); the pair part, self and
background terms do not depend on it, and energy and virial carry the factor consistently.
"""
from __future__ import annotations

import warnings
from typing import Optional, Sequence, Tuple

from . import _lib
from .solver import Params, Solver

#: Maximum supported stencil order P.
MAX_ORDER: int = 16
#: Default introspection-table size (esp.py:94).
DEFAULT_TABLE_SEGMENTS: int = 512
#: Minimum introspection-table size (esp.py:97).
MIN_TABLE_SEGMENTS: int = 16


def _check_positive_int(val, name):
    # esp.py:103-105
    if not isinstance(val, int) or isinstance(val, bool) or val <= 0:
        raise ValueError(f"ESP: '{name}' must be a positive integer, got {val!r}.")


def _check_positive_float(val, name):
    # esp.py:108-110
    if not (isinstance(val, (int, float)) and not isinstance(val, bool) and val > 0):
        raise ValueError(f"ESP: '{name}' must be a positive number, got {val!r}.")


def _check_nonneg_float(val, name):
    # esp.py:113-115
    if not (isinstance(val, (int, float)) and not isinstance(val, bool) and val >= 0):
        raise ValueError(f"ESP: '{name}' must be >= 0, got {val!r}.")


def _validate_resolution(res: Sequence[int]) -> Tuple[int, int, int]:
    # esp.py:118-125
    res = tuple(int(v) for v in res)
    if len(res) != 3:
        raise ValueError(f"ESP: 'resolution' must be a 3-tuple (Nx, Ny, Nz), got length {len(res)}.")
    for i, (v, label) in enumerate(zip(res, ("Nx", "Ny", "Nz"))):
        _check_positive_int(v, f"resolution[{i}] ({label})")
    return res  # type: ignore[return-value]


def _estimate_kappa(r_cut: float, resolution) -> float:
    # esp.py:128-130
    return 3.2 / r_cut


class Spectral:
    """Long-range Coulomb force by the ESP mesh method (ESPForceCompute on the GPU).

    Parameters follow esp.py:208-217: ``nlist``, ``default_r_cut``, ``resolution``, ``order``,
    ``kappa`` (heuristic 3.2 / r_cut with a warning when None), ``alpha``, ``n_table`` and
    ``nlist_buffer`` (default 0.1 r_cut), plus ``precision`` ("fp64" or "mixed").
    """

    def __init__(self, nlist, default_r_cut: float, resolution: Sequence[int] = (32, 32, 32), order: int = 6,
                 kappa: Optional[float] = None, alpha: float = 0.0, n_table: int = DEFAULT_TABLE_SEGMENTS,
                 nlist_buffer: Optional[float] = None, precision: str = "fp64"):
        _check_positive_float(default_r_cut, "default_r_cut")
        res = _validate_resolution(resolution)
        if not (isinstance(order, int) and 1 <= order <= MAX_ORDER):
            raise ValueError(f"ESP: 'order' must be an integer in [1, {MAX_ORDER}], got {order!r}.")
        if kappa is None:
            kappa = _estimate_kappa(default_r_cut, res)
            warnings.warn(
                f"ESP: 'kappa' not specified; using heuristic kappa = {kappa:.4f} "
                f"(= 3.2 / r_cut). For production runs, tune kappa explicitly.",
                stacklevel=2,
            )
        else:
            _check_positive_float(kappa, "kappa")
        _check_nonneg_float(alpha, "alpha")
        if not (isinstance(n_table, int) and n_table >= MIN_TABLE_SEGMENTS):
            raise ValueError(f"ESP: 'n_table' must be an integer >= {MIN_TABLE_SEGMENTS}, got {n_table!r}.")
        if precision not in ("fp64", "mixed"):
            raise ValueError("ESP: 'precision' must be 'fp64' or 'mixed'.")
        self._nlist = nlist
        self._r_cut = float(default_r_cut)
        self._resolution = res
        self._order = int(order)
        self._kappa = float(kappa)
        self._alpha = float(alpha)
        self._n_table = int(n_table)
        self._precision = precision
        self._nlist_buffer = 0.1 * self._r_cut if nlist_buffer is None else float(nlist_buffer)
        self._cpp_obj: Optional[Solver] = None
        self._pair: Optional["Local"] = None
        self._simulation = None

    # -- properties (esp.py:266-313) --------------------------------------------------------
    @property
    def kappa(self) -> float:
        return self._cpp_obj.params.kappa if self._cpp_obj is not None else self._kappa

    @property
    def resolution(self) -> Tuple[int, int, int]:
        if self._cpp_obj is not None:
            p = self._cpp_obj.params
            return (p.nx, p.ny, p.nz)
        return self._resolution

    @property
    def order(self) -> int:
        return self._cpp_obj.params.order if self._cpp_obj is not None else self._order

    @property
    def r_cut(self) -> float:
        return self._cpp_obj.params.rcut if self._cpp_obj is not None else self._r_cut

    @property
    def alpha(self) -> float:
        return self._cpp_obj.params.alpha if self._cpp_obj is not None else self._alpha

    @property
    def q_sum(self) -> float:
        if self._cpp_obj is None:
            raise RuntimeError("ESP: q_sum not available before simulation attachment.")
        return self._cpp_obj.charge_sums()[0]

    @property
    def q2_sum(self) -> float:
        if self._cpp_obj is None:
            raise RuntimeError("ESP: q2_sum not available before simulation attachment.")
        return self._cpp_obj.charge_sums()[1]

    @property
    def table(self):
        """ESPTableEntry list (introspection only, ESPForceCompute.cc:336-372)."""
        if self._cpp_obj is None:
            raise RuntimeError("ESP: table not available before simulation attachment.")
        return [self._cpp_obj.table_entry(i) for i in range(self._n_table)]

    # -- runtime parameter update (esp.py:319-394) ----------------------------------------------
    def set_params(self, *, resolution=None, order=None, kappa=None, r_cut=None, alpha=None, n_table=None) -> None:
        if self._cpp_obj is None:
            raise RuntimeError(
                "ESP: 'set_params()' called before the Simulation was attached. "
                "Pass initial values to the Spectral constructor instead.")
        if resolution is not None:
            self._resolution = _validate_resolution(resolution)
        if order is not None:
            if not (isinstance(order, int) and 1 <= order <= MAX_ORDER):
                raise ValueError(f"ESP: invalid order {order!r}.")
            self._order = int(order)
        if kappa is not None:
            _check_positive_float(kappa, "kappa")
            self._kappa = float(kappa)
        if r_cut is not None:
            _check_positive_float(r_cut, "r_cut")
            self._r_cut = float(r_cut)
        if alpha is not None:
            _check_nonneg_float(alpha, "alpha")
            self._alpha = float(alpha)
        if n_table is not None:
            if not (isinstance(n_table, int) and n_table >= MIN_TABLE_SEGMENTS):
                raise ValueError(f"ESP: invalid n_table {n_table!r}.")
            self._n_table = int(n_table)
        self._cpp_obj.set_params(self._params())
        if self._pair is not None:
            self._pair._sync_params_from_spectral()

    def invalidate(self) -> None:
        if self._cpp_obj is not None:
            self._cpp_obj.invalidate()

    # -- lifecycle ------------------------------------------------------------------------------
    def _params(self) -> Params:
        nx, ny, nz = self._resolution
        return Params(nx=nx, ny=ny, nz=nz, order=self._order, kappa=self._kappa, rcut=self._r_cut, alpha=self._alpha,
                      n_table=self._n_table,
                      precision=_lib.ESP_PREC_FP64 if self._precision == "fp64" else _lib.ESP_PREC_MIXED)

    def _attach(self, simulation) -> None:
        # Spectral._attach_hook (esp.py:409-430)
        self._cpp_obj = Solver(self._params())
        self._simulation = simulation
        if self._pair is not None and self._pair._simulation is simulation:
            self._pair._sync_params_from_spectral()

    def _detach(self) -> None:
        self._cpp_obj = None
        self._simulation = None

    def _esp_flags(self) -> int:
        return _lib.ESP_COMPUTE_MESH | _lib.ESP_COMPUTE_ENERGY_VIRIAL

    def _esp_solver(self) -> Solver:
        return self._cpp_obj

    def __repr__(self) -> str:
        return (f"esp.Spectral(resolution={self._resolution}, order={self._order}, kappa={self._kappa:.4g}, "
                f"r_cut={self._r_cut}, alpha={self._alpha})")


class _ParamDict(dict):
    """Per type-pair parameter dictionary; values are validated by Local._validate_params."""

    def __init__(self, owner):
        super().__init__()
        self._owner = owner

    def __setitem__(self, key, value):
        super().__setitem__(key, self._owner._validate_params(dict(value)))


class Local:
    """Short-range real-space complement of :class:`Spectral` (PotentialPair<EvaluatorPairPSWF>).

    Arguments follow esp.py:505-537: ``nlist``, ``default_r_cut``, ``spectral`` (or the deprecated
    ``coulomb``), ``default_r_on`` and ``mode`` ("none" only).
    """

    _accepted_modes = ("none",)

    def __init__(self, nlist, default_r_cut: float, spectral: Optional[Spectral] = None,
                 coulomb: Optional[Spectral] = None, default_r_on: float = 0.0, mode: str = "none"):
        if spectral is None and coulomb is not None:
            warnings.warn("The 'coulomb' parameter is deprecated; use 'spectral' instead.", DeprecationWarning,
                          stacklevel=2)
            spectral = coulomb
        if spectral is None:
            raise ValueError("'spectral' (or deprecated 'coulomb') must be provided.")
        if not isinstance(spectral, Spectral):
            raise TypeError(f"ESP: 'spectral' must be an esp.Spectral instance, got {type(spectral).__name__!r}.")
        if mode not in ("none",):
            raise ValueError("ESP: Pair mode must be 'none' for ESP (L(r) already vanishes at r_c).")
        _check_positive_float(default_r_cut, "default_r_cut")
        self._nlist = nlist
        self._default_r_cut = float(default_r_cut)
        self._default_r_on = float(default_r_on)
        self._mode = mode
        self._spectral = spectral
        spectral._pair = self
        self.params = _ParamDict(self)
        self.r_cut_by_pair = {}
        self._cpp_obj = None
        self._simulation = None

    @staticmethod
    def _validate_params(params: dict) -> dict:
        # esp.py:543-558: ESP pair parameters carry no user-settable fields
        if params and set(params.keys()) - {"_reserved", "kappa", "rcut", "alpha", "V_cut"}:
            warnings.warn(
                "ESP: Local.params accepts an empty dict {}; all ESP parameters are read from the "
                f"associated Spectral object. Ignoring keys: {set(params.keys())!r}.",
                stacklevel=3,
            )
            return {}
        return params

    @property
    def r_cut(self) -> float:
        return self._spectral.r_cut

    def _attach(self, simulation) -> None:
        # Local._attach_hook (esp.py:564-567)
        self._simulation = simulation
        self._cpp_obj = True
        self._sync_params_from_spectral()

    def _sync_params_from_spectral(self) -> None:
        # esp.py:569-610: push {kappa, rcut, alpha, V_cut=0} to every registered type pair
        if self._cpp_obj is None:
            return
        if self._spectral._cpp_obj is None:
            raise RuntimeError(
                "ESP: Local._sync_params_from_spectral() called but the associated Spectral object is "
                "not attached. Attach Spectral before Local.")
        sp = self._spectral
        for key in list(self.params.keys()):
            dict.__setitem__(self.params, key, dict(kappa=sp.kappa, rcut=sp.r_cut, alpha=sp.alpha, V_cut=0.0))

    def _detach(self) -> None:
        self._cpp_obj = None
        self._simulation = None
        self._spectral._pair = None

    def _esp_flags(self) -> int:
        return _lib.ESP_COMPUTE_PAIR

    def _esp_solver(self) -> Solver:
        if self._spectral._cpp_obj is None:
            raise RuntimeError("ESP: Local needs its Spectral attached to the same simulation.")
        return self._spectral._cpp_obj

    def __repr__(self) -> str:
        return (f"esp.Local(r_cut={self._spectral._r_cut}, kappa={self._spectral._kappa:.4g}, "
                f"alpha={self._spectral._alpha})")


def make_esp_forces(nlist, r_cut: float, resolution: Sequence[int] = (32, 32, 32), order: int = 6,
                    kappa: Optional[float] = None, alpha: float = 0.0, n_table: int = DEFAULT_TABLE_SEGMENTS,
                    precision: str = "fp64") -> Tuple[Spectral, Local]:
    """Coupled (Spectral, Local) pair, as esp.py:632-693."""
    spectral = Spectral(nlist=nlist, default_r_cut=r_cut, resolution=resolution, order=order, kappa=kappa,
                        alpha=alpha, n_table=n_table, precision=precision)
    local = Local(nlist=nlist, default_r_cut=r_cut, spectral=spectral)
    return spectral, local


Coulomb = Spectral
Pair = Local

__all__ = ["Spectral", "Local", "make_esp_forces", "Coulomb", "Pair", "MAX_ORDER", "DEFAULT_TABLE_SEGMENTS",
           "MIN_TABLE_SEGMENTS"]
