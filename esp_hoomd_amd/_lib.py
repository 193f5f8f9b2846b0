"""ctypes binding of the C ABI declared in ``include/esp_mi355x.h``.

The shared library ``libesp_mi355x.so`` is built in-tree by ``esp_hoomd_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback for the solver.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libesp_mi355x.so"

ESP_OK = 0
ESP_ERR_ARG = -1
ESP_ERR_HIP = -2
ESP_ERR_OOM = -3
ESP_ERR_STATE = -4
ESP_ERR_RANGE = -5
ESP_ERR_UNSUPPORTED = -6

ESP_PREC_FP64 = 0
ESP_PREC_MIXED = 1
ESP_SPLIT_PSWF = 0
ESP_SPLIT_GAUSSIAN = 1

ESP_COMPUTE_PAIR = 1
ESP_COMPUTE_MESH = 2
ESP_COMPUTE_ENERGY_VIRIAL = 4
ESP_COMPUTE_ALL = 7
ESP_NOALLOC_MAX_PARTICLES = 262144
ESP_N_STAGES = 12


class EspParams(C.Structure):
    _fields_ = [
        ("nx", C.c_uint), ("ny", C.c_uint), ("nz", C.c_uint),
        ("order", C.c_uint),
        ("kappa", C.c_double), ("rcut", C.c_double), ("alpha", C.c_double),
        ("n_table", C.c_uint),
        ("precision", C.c_int), ("split", C.c_int),
        ("window_c", C.c_double), ("window_degree", C.c_uint),
        ("pair_lanes", C.c_int), ("k_cut", C.c_double), ("tile", C.c_uint),
    ]


class EspSystem(C.Structure):
    _fields_ = [
        ("N", C.c_uint),
        ("postype", C.c_void_p), ("charge", C.c_void_p),
        ("group_size", C.c_uint), ("group_idx", C.c_void_p),
        ("box", C.c_double * 6),
        ("n_neigh", C.c_void_p), ("nlist", C.c_void_p), ("head_list", C.c_void_p),
        ("n_ex", C.c_void_p), ("ex_list", C.c_void_p), ("ex_pitch", C.c_uint),
    ]


class EspOut(C.Structure):
    _fields_ = [
        ("force", C.c_void_p), ("virial", C.c_void_p), ("virial_pitch", C.c_size_t),
        ("ext_energy_virial", C.c_void_p), ("flags", C.c_uint),
    ]


class EspTableEntry(C.Structure):
    _fields_ = [
        ("potential_coeffs", C.c_double * 4), ("force_coeffs", C.c_double * 4),
        ("potential_coeff4", C.c_double), ("force_coeff4", C.c_double),
        ("r_lo", C.c_double), ("dr_inv", C.c_double),
    ]


class EspInfo(C.Structure):
    _fields_ = [
        ("nx", C.c_uint), ("ny", C.c_uint), ("nz", C.c_uint), ("order", C.c_uint),
        ("kappa", C.c_double), ("rcut", C.c_double), ("alpha", C.c_double),
        ("n_table", C.c_uint), ("precision", C.c_int), ("split", C.c_int),
        ("window_degree", C.c_uint), ("window_c", C.c_double),
        ("self_coeff", C.c_double), ("background_coeff", C.c_double),
        ("tile", C.c_uint), ("device_bytes", C.c_size_t), ("k_cut", C.c_double),
    ]


class EspPairParam(C.Structure):
    """EvaluatorPairPSWF::param_type (include/EvaluatorPairPSWF.h)."""
    _fields_ = [
        ("kappa", C.c_double), ("alpha", C.c_double), ("rcut", C.c_double), ("V_cut", C.c_double),
        ("table_host", C.c_void_p), ("table_device", C.c_void_p), ("n_segs", C.c_uint), ("pad", C.c_uint * 3),
    ]


class EspError(RuntimeError):
    """Raised for a non-zero status from the C ABI."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None

_DP = C.POINTER(C.c_double)
_VP = C.c_void_p


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def load() -> C.CDLL:
    """Load the library (once).  Raises ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"esp_hoomd_amd: {LIB_PATH.name} not found; build it with "
            "`python -m esp_hoomd_amd.build` (hipcc, gfx950)")
    lib = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
    _sig(lib, "esp_create", C.c_int, [C.POINTER(_VP), C.POINTER(EspParams)])
    _sig(lib, "esp_set_params", C.c_int, [_VP, C.POINTER(EspParams)])
    _sig(lib, "esp_invalidate", C.c_int, [_VP])
    _sig(lib, "esp_destroy", C.c_int, [_VP])
    _sig(lib, "esp_last_error", C.c_char_p, [_VP])
    _sig(lib, "esp_compute_forces", C.c_int, [_VP, C.c_uint64, C.POINTER(EspSystem), C.POINTER(EspOut), _VP])
    _sig(lib, "esp_pair_forces", C.c_int, [_VP, C.POINTER(EspSystem), C.POINTER(EspOut), _VP])
    _sig(lib, "esp_prepare", C.c_int, [_VP, C.POINTER(EspSystem), C.POINTER(EspOut), _VP])
    _sig(lib, "esp_pair_forces_noalloc", C.c_int, [_VP, C.POINTER(EspSystem), C.POINTER(EspOut), _VP])
    _sig(lib, "esp_batch_create", C.c_int, [C.POINTER(_VP)])
    _sig(lib, "esp_batch_destroy", C.c_int, [_VP])
    _sig(lib, "esp_batch_last_error", C.c_char_p, [_VP])
    _sig(lib, "esp_batch_set_options", C.c_int, [_VP, C.c_int, C.c_uint])
    _sig(lib, "esp_batch_compute", C.c_int, [_VP, C.c_uint, C.POINTER(_VP), C.POINTER(EspSystem), C.POINTER(EspOut), _VP])
    _sig(lib, "esp_batch_stats", C.c_int, [_VP, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)])
    _sig(lib, "esp_get_info", C.c_int, [_VP, C.POINTER(EspInfo)])
    _sig(lib, "esp_get_charge_sums", C.c_int, [_VP, _DP, _DP])
    _sig(lib, "esp_get_table_size", C.c_int, [_VP, C.POINTER(C.c_uint)])
    _sig(lib, "esp_get_table_entry", C.c_int, [_VP, C.c_uint, C.POINTER(EspTableEntry)])
    _sig(lib, "esp_get_split_table", C.c_int, [_VP, C.POINTER(C.c_uint), _DP, _DP, _DP, _DP])
    _sig(lib, "esp_set_split_table", C.c_int, [_VP, C.c_uint, _DP, _DP, _DP, _DP])
    _sig(lib, "esp_get_fixed_point_scale", C.c_int, [_VP, C.POINTER(C.c_float), C.POINTER(C.c_int)])
    _sig(lib, "esp_get_pair_table", C.c_int, [_VP, C.POINTER(_VP), C.POINTER(_VP), C.POINTER(C.c_uint)])
    _sig(lib, "esp_potential_pair_forces", C.c_int, [C.POINTER(EspSystem), C.POINTER(EspPairParam), C.c_double, C.c_int,
                                                     C.POINTER(EspOut), _VP])
    _sig(lib, "esp_get_window", C.c_int, [_VP, C.POINTER(C.c_uint), C.POINTER(C.c_uint), _DP])
    _sig(lib, "esp_set_profiling", C.c_int, [_VP, C.c_int])
    _sig(lib, "esp_set_graph", C.c_int, [_VP, C.c_int])
    _sig(lib, "esp_get_graph_stats", C.c_int, [_VP, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)])
    _sig(lib, "esp_get_stage_times", C.c_int, [_VP, _DP, C.POINTER(C.c_uint)])
    _sig(lib, "esp_reset_stage_times", C.c_int, [_VP])
    _sig(lib, "esp_stage_name", C.c_char_p, [C.c_int])
    _sig(lib, "esp_build_split_table", C.c_int, [C.c_int, C.c_double, C.c_double, C.c_uint, _DP, _DP, _DP, _DP])
    _sig(lib, "esp_split_eval", C.c_int, [C.c_int, C.c_double, C.c_double, C.c_int, _DP, _DP, _DP])
    _sig(lib, "esp_split_fourier", C.c_int, [C.c_int, C.c_double, C.c_double, C.c_int, _DP, _DP, _DP])
    _sig(lib, "esp_split_constants", C.c_int, [C.c_int, C.c_double, C.c_double, _DP, _DP])
    _sig(lib, "esp_prolate_eval", C.c_int, [C.c_double, C.c_int, _DP, _DP, _DP])
    _sig(lib, "esp_build_window", C.c_int, [C.c_uint, C.c_uint, C.c_double, _DP])
    _sig(lib, "esp_window_fourier", C.c_int, [C.c_uint, C.c_uint, _DP, C.c_int, _DP, _DP])
    _sig(lib, "esp_default_window_c", C.c_double, [C.c_uint, C.c_double])
    _sig(lib, "esp_default_window_degree", C.c_uint, [C.c_uint])
    _sig(lib, "esp_nlist_create", C.c_int, [C.POINTER(_VP), C.c_double, C.c_double])
    _sig(lib, "esp_nlist_destroy", C.c_int, [_VP])
    _sig(lib, "esp_nlist_last_error", C.c_char_p, [_VP])
    _sig(lib, "esp_nlist_build", C.c_int, [_VP, C.POINTER(EspSystem), _VP])
    _sig(lib, "esp_nlist_update", C.c_int, [_VP, C.POINTER(EspSystem), C.POINTER(C.c_int), _VP])
    _sig(lib, "esp_nlist_arrays", C.c_int, [_VP, C.POINTER(_VP), C.POINTER(_VP), C.POINTER(_VP), C.POINTER(C.c_uint)])
    _sig(lib, "esp_nlist_stats", C.c_int, [_VP, C.POINTER(C.c_ulonglong)])
    _sig(lib, "esp_nlist_cell_order", C.c_int, [_VP, C.POINTER(C.c_void_p), C.POINTER(C.c_uint)])
    _sig(lib, "esp_gather_rows", C.c_int, [C.c_uint, C.c_uint, _VP, _VP, _VP, _VP])
    _sig(lib, "esp_nlist_set_inner", C.c_int, [_VP, C.c_double])
    _sig(lib, "esp_nlist_inner_arrays", C.c_int, [_VP, C.POINTER(_VP), C.POINTER(_VP), C.POINTER(_VP), C.POINTER(C.c_uint)])
    _sig(lib, "esp_wca_forces", C.c_int, [C.POINTER(EspSystem), C.c_double, C.c_double, C.POINTER(EspOut), _VP])
    _sig(lib, "esp_integrate_first", C.c_int, [C.c_uint, _VP, _VP, _VP, C.c_double, _DP, _VP])
    _sig(lib, "esp_integrate_second", C.c_int, [C.c_uint, _VP, _VP, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                                C.c_uint64, _VP])
    _sig(lib, "esp_kinetic_potential", C.c_int, [C.c_uint, _VP, _VP, _DP, _DP])
    _sig(lib, "esp_md_batch_create", C.c_int, [C.POINTER(_VP)])
    _sig(lib, "esp_md_batch_destroy", C.c_int, [_VP])
    _sig(lib, "esp_md_batch_last_error", C.c_char_p, [_VP])
    _sig(lib, "esp_md_batch_set", C.c_int, [_VP, C.c_uint, C.POINTER(EspSystem), C.POINTER(_VP), C.POINTER(EspOut), C.POINTER(_VP),
                                          C.POINTER(C.c_uint64)])
    _sig(lib, "esp_md_batch_set_async", C.c_int, [_VP, C.c_uint, C.POINTER(EspSystem), C.POINTER(_VP), C.POINTER(EspOut),
                                                C.POINTER(_VP), C.POINTER(C.c_uint64), _VP])
    _sig(lib, "esp_md_batch_set_hold", C.c_int, [_VP, C.POINTER(C.c_ubyte), _VP])
    _sig(lib, "esp_md_batch_integrate_first", C.c_int, [_VP, C.c_double, _VP])
    _sig(lib, "esp_md_batch_integrate_second", C.c_int, [_VP, C.c_double, C.c_double, C.c_double, C.c_uint64, _VP])
    _sig(lib, "esp_md_batch_wca", C.c_int, [_VP, C.c_double, C.c_double, _VP])
    _sig(lib, "esp_md_batch_update_nlists", C.c_int, [_VP, C.POINTER(C.c_uint), C.POINTER(C.c_ubyte), _VP])
    _sig(lib, "esp_nlist_set_capacity", C.c_int, [_VP, C.c_uint, C.c_uint])
    _sig(lib, "esp_nlist_set_cutoff", C.c_int, [_VP, C.c_double])
    _sig(lib, "esp_md_batch_predict", C.c_int, [_VP, C.c_double, _VP])
    _sig(lib, "esp_md_batch_read_predicted", C.c_int, [_VP, C.POINTER(C.c_ubyte)])
    _sig(lib, "esp_md_batch_rebuild_async", C.c_int, [_VP, C.POINTER(C.c_ubyte), C.POINTER(C.c_ubyte), _VP])
    _sig(lib, "esp_md_batch_finish_rebuilds", C.c_int, [_VP, C.POINTER(C.c_ubyte), C.POINTER(C.c_uint)])
    _sig(lib, "esp_device_count", C.c_int, [C.POINTER(C.c_int)])
    _sig(lib, "esp_set_device", C.c_int, [C.c_int])
    _sig(lib, "esp_device_malloc", C.c_int, [C.POINTER(_VP), C.c_size_t])
    _sig(lib, "esp_device_free", C.c_int, [_VP])
    _sig(lib, "esp_memcpy_htod", C.c_int, [_VP, _VP, C.c_size_t])
    _sig(lib, "esp_memcpy_dtoh", C.c_int, [_VP, _VP, C.c_size_t])
    _sig(lib, "esp_memcpy_dtod", C.c_int, [_VP, _VP, C.c_size_t, _VP])
    _sig(lib, "esp_memcpy_htod_async", C.c_int, [_VP, _VP, C.c_size_t, _VP])
    _sig(lib, "esp_host_malloc_pinned", C.c_int, [C.POINTER(_VP), C.c_size_t])
    _sig(lib, "esp_host_free_pinned", C.c_int, [_VP])
    _sig(lib, "esp_memset", C.c_int, [_VP, C.c_int, C.c_size_t])
    _sig(lib, "esp_copy_records", C.c_int, [_VP, _VP, C.c_uint, _VP])
    _sig(lib, "esp_stream_create", C.c_int, [C.POINTER(_VP)])
    _sig(lib, "esp_stream_destroy", C.c_int, [_VP])
    _sig(lib, "esp_stream_synchronize", C.c_int, [_VP])
    _sig(lib, "esp_device_synchronize", C.c_int, [])
    _sig(lib, "esp_event_create", C.c_int, [C.POINTER(_VP)])
    _sig(lib, "esp_event_destroy", C.c_int, [_VP])
    _sig(lib, "esp_event_record", C.c_int, [_VP, _VP])
    _sig(lib, "esp_event_elapsed_ms", C.c_int, [_VP, _VP, C.POINTER(C.c_float)])
    _sig(lib, "esp_version", C.c_char_p, [])
    _lib = lib
    return lib


def check(rc: int, ctx=None) -> None:
    if rc != ESP_OK:
        lib = load()
        msg = lib.esp_last_error(ctx).decode() if lib.esp_last_error(ctx) else ""
        raise EspError(rc, msg or "error")


def check_batch(rc: int, batch) -> None:
    if rc != ESP_OK:
        msg = load().esp_batch_last_error(batch)
        raise EspError(rc, (msg.decode() if msg else "") or "error")


def dptr(a: np.ndarray):
    """ctypes double pointer of a contiguous float64 numpy array."""
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def gpu_available() -> bool:
    """True when a HIP device is visible (does not initialise anything else)."""
    if os.environ.get("ESP_FORCE_NO_GPU"):
        return False
    lib = load()
    n = C.c_int(0)
    lib.esp_device_count(C.byref(n))
    return n.value > 0
