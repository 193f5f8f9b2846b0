"""HOOMD-blue front end of the MI355X ESP plugin (the reference's src/esp.py on hoomd.md):
``Spectral`` (hoomd.md.force.Force, the long-range part, esp.py:137) and ``Local``
(hoomd.md.pair.Pair over PotentialPairPSWF[GPU], esp.py:456), with ``make_esp_forces``.

Arguments, defaults and validation are those of esp_hoomd_amd.esp (the ctypes mirror of the
same interface).  By default ``Spectral`` evaluates the whole ESP force -- mesh, pair
correction and exclusion correction -- in one library call over the neighbour list, so no
``Local`` is attached; ``make_esp_forces(..., fused=False)`` instead returns a mesh-only
``Spectral`` and a ``Local`` whose pair parameters (cutoff, kappa and the split table's
addresses) follow the Spectral's.  Needs a HOOMD-blue GPU device (HIP build) and the compiled
``_esp`` module of this directory (CMakeLists.txt).
"""
from __future__ import annotations

import hoomd
from hoomd.data.parameterdicts import TypeParameterDict
from hoomd.data.typeparam import TypeParameter
from hoomd.md import _md

from esp_hoomd_amd.esp import (DEFAULT_TABLE_SEGMENTS, MAX_ORDER, _check_positive_float, _estimate_kappa,
                               _validate_resolution)

from . import _esp

_PRECISIONS = {"fp64": 0, "mixed": 1}


class Spectral(hoomd.md.force.Force):
    """Long-range Coulomb force by the ESP mesh method on the GPU (esp.py:137-455)."""

    def __init__(self, nlist, default_r_cut, resolution=(32, 32, 32), order=6, kappa=None, alpha=0.0,
                 n_table=DEFAULT_TABLE_SEGMENTS, precision="fp64", fused=True):
        super().__init__()
        _check_positive_float(default_r_cut, "default_r_cut")
        self._resolution = _validate_resolution(resolution)
        if not (isinstance(order, int) and 1 <= order <= MAX_ORDER):
            raise ValueError(f"ESP: 'order' must be an integer in [1, {MAX_ORDER}], got {order!r}.")
        if precision not in _PRECISIONS:
            raise ValueError(f"ESP: 'precision' must be one of {sorted(_PRECISIONS)}, got {precision!r}.")
        self._nlist = nlist
        self._r_cut = float(default_r_cut)
        self._order = int(order)
        self._kappa = float(kappa) if kappa is not None else _estimate_kappa(self._r_cut, self._resolution)
        self._alpha = float(alpha)
        self._n_table = int(n_table)
        self._precision = precision
        self._fused = bool(fused)
        self._local = None

    @property
    def nlist(self):
        return self._nlist

    def _attach_hook(self):
        if not isinstance(self._simulation.device, hoomd.device.GPU):
            raise RuntimeError("ESP: the MI355X library needs a hoomd.device.GPU")
        if self._nlist._attached and self._simulation != self._nlist._simulation:
            self._nlist._detach()
        self._nlist._attach(self._simulation)
        # the library walks HOOMD's full neighbour list
        self._nlist._cpp_obj.setStorageMode(_md.NeighborList.storageMode.full)
        group = self._simulation.state._get_group(hoomd.filter.All())
        self._cpp_obj = _esp.ESPForceComputeMI355X(self._simulation.state._cpp_sys_def, self._nlist._cpp_obj, group)
        self._push_params()
        if self._local is not None and self._local._attached:
            self._local._sync_from_spectral()

    def _detach_hook(self):
        self._nlist._detach()

    def _push_params(self):
        nx, ny, nz = self._resolution
        self._cpp_obj.setParams(nx, ny, nz, self._order, self._kappa, self._r_cut, self._alpha, 4096,
                                _PRECISIONS[self._precision])
        self._cpp_obj.setFused(self._fused)

    def _require_attached(self, what):
        if not self._attached:
            raise RuntimeError(f"ESP: '{what}' is only available after the force is attached to a simulation.")

    @property
    def kappa(self):
        return self._kappa

    @property
    def resolution(self):
        return self._resolution

    @property
    def order(self):
        return self._order

    @property
    def r_cut(self):
        return self._r_cut

    @property
    def alpha(self):
        return self._alpha

    @property
    def q_sum(self):
        self._require_attached("q_sum")
        return self._cpp_obj.getQSum()

    @property
    def q2_sum(self):
        self._require_attached("q2_sum")
        return self._cpp_obj.getQ2Sum()

    @property
    def table(self):
        self._require_attached("table")
        return [self._cpp_obj.getTableEntry(i) for i in range(self._cpp_obj.getTableSize())]

    def set_params(self, *, resolution=None, order=None, kappa=None, r_cut=None, alpha=None, n_table=None):
        """Change parameters of an attached force (esp.py:300-352)."""
        self._require_attached("set_params")
        if resolution is not None:
            self._resolution = _validate_resolution(resolution)
        if order is not None:
            self._order = int(order)
        if kappa is not None:
            self._kappa = float(kappa)
        if r_cut is not None:
            _check_positive_float(r_cut, "r_cut")
            self._r_cut = float(r_cut)
        if alpha is not None:
            self._alpha = float(alpha)
        if n_table is not None:
            self._n_table = int(n_table)
        self._push_params()
        if self._local is not None and self._local._attached:
            self._local._sync_from_spectral()

    def invalidate(self):
        self._require_attached("invalidate")
        self._cpp_obj.invalidate()


class Local(hoomd.md.pair.Pair):
    """Short-range PSWF pair correction through HOOMD's PotentialPair (esp.py:456-600); its
    parameters follow the Spectral's (cutoff, kappa and split table)."""

    _ext_module = _esp
    _cpp_class_name = "PotentialPairPSWF"
    _accepted_modes = ("none",)

    def __init__(self, nlist, default_r_cut, spectral, default_r_on=0.0, mode="none"):
        super().__init__(nlist, default_r_cut, default_r_on, mode)
        params = TypeParameter("params", "particle_types",
                               TypeParameterDict(kappa=float, alpha=float, rcut=float, V_cut=float, table_ptr=int,
                                                 table_device_ptr=int, n_segs=int, len_keys=2))
        self._add_typeparam(params)
        self._spectral = spectral
        spectral._local = self

    def _attach_hook(self):
        super()._attach_hook()
        if self._spectral._attached:
            self._sync_from_spectral()

    def _sync_from_spectral(self):
        s = self._spectral
        cpp = s._cpp_obj
        self.params.default = dict(kappa=s.kappa, alpha=s.alpha, rcut=s.r_cut, V_cut=0.0, table_ptr=cpp.getTablePtr(),
                                   table_device_ptr=cpp.getTableDevicePtr(), n_segs=cpp.getTableSegments())
        self.r_cut.default = s.r_cut


def make_esp_forces(nlist, r_cut, resolution=(32, 32, 32), order=6, kappa=None, alpha=0.0, precision="fp64", fused=True):
    """(Spectral, Local or None): with fused=True one Spectral evaluates everything."""
    spectral = Spectral(nlist, r_cut, resolution=resolution, order=order, kappa=kappa, alpha=alpha, precision=precision,
                        fused=fused)
    local = None if fused else Local(nlist, r_cut, spectral)
    return spectral, local


__all__ = ["Spectral", "Local", "make_esp_forces"]
