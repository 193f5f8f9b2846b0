"""The C ABI: the library loads, exports every declared entry point, and reports argument
errors with status codes (host-side paths only; no GPU work)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from esp_hoomd_amd import _lib
from esp_hoomd_amd.solver import Params, Solver

HEADER = Path(__file__).resolve().parent.parent / "include" / "esp_mi355x.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|double|unsigned int)\s+(esp_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_version():
    assert _lib.load().esp_version().decode().startswith("esp_mi355x")


@pytest.mark.parametrize("bad,code", [
    (dict(order=1), _lib.ESP_ERR_ARG),
    (dict(order=17), _lib.ESP_ERR_ARG),
    (dict(rcut=0.0), _lib.ESP_ERR_ARG),
    (dict(kappa=-1.0), _lib.ESP_ERR_ARG),
    (dict(alpha=-0.5), _lib.ESP_ERR_ARG),
    (dict(alpha=float("nan")), _lib.ESP_ERR_ARG),
    (dict(nx=14), _lib.ESP_ERR_UNSUPPORTED),
    (dict(nx=1), _lib.ESP_ERR_ARG),
    (dict(n_table=8), _lib.ESP_ERR_ARG),
    (dict(precision=3), _lib.ESP_ERR_ARG),
    (dict(window_degree=20), _lib.ESP_ERR_ARG),
    (dict(pair_lanes=3), _lib.ESP_ERR_ARG),
])
def test_invalid_params_are_rejected(bad, code):
    p = Params(nx=32, ny=32, nz=32, order=6, kappa=3.0, rcut=3.0)
    for k, v in bad.items():
        setattr(p, k, v)
    with pytest.raises(_lib.EspError) as ei:
        Solver(p)
    assert ei.value.code == code
    assert len(_lib.load().esp_last_error(None)) > 0


def test_set_params_error_keeps_context_usable():
    s = Solver(Params(order=6, kappa=3.0, rcut=3.0))
    with pytest.raises(_lib.EspError):
        s.set_params(Params(order=40, kappa=3.0, rcut=3.0))
    s.set_params(Params(order=8, kappa=3.0, rcut=3.0))
    assert s.info().order == 8


def test_null_arguments():
    lib = _lib.load()
    assert lib.esp_create(None, None) == _lib.ESP_ERR_ARG
    assert lib.esp_compute_forces(None, 0, None, None, None) == _lib.ESP_ERR_ARG
    s = Solver(Params(kappa=3.0, rcut=3.0))
    assert lib.esp_compute_forces(s.ctx, 0, None, None, None) == _lib.ESP_ERR_ARG
    assert lib.esp_prepare(None, None, None, None) == _lib.ESP_ERR_ARG
    assert lib.esp_nlist_cell_order(None, None, None) == _lib.ESP_ERR_ARG
    assert lib.esp_gather_rows(4, 8, None, None, None, None) == _lib.ESP_ERR_ARG
    assert lib.esp_gather_rows(0, 8, None, None, None, None) == _lib.ESP_OK
    assert lib.esp_prepare(s.ctx, None, None, None) == _lib.ESP_ERR_ARG


def test_table_entries_and_range_error():
    s = Solver(Params(kappa=4.0, rcut=3.0, n_table=512))
    lib = _lib.load()
    n = C.c_uint()
    assert lib.esp_get_table_size(s.ctx, C.byref(n)) == 0 and n.value == 512
    e = s.table_entry(100)
    r_mid = (100.5 / 512) * 3.0
    from esp_hoomd_amd.solver import split_eval
    Sv, dS = split_eval(0, 4.0, 3.0, np.array([r_mid / 3.0]))
    assert abs(e.potential_coeffs[0] - Sv[0] / (4 * np.pi * r_mid)) < 1e-12
    assert abs(e.r_lo - 100 / 512 * 3.0) < 1e-15 and abs(e.dr_inv - 512 / 3.0) < 1e-9
    with pytest.raises(_lib.EspError) as ei:
        s.table_entry(512)
    assert ei.value.code == _lib.ESP_ERR_RANGE


def test_split_table_roundtrip_and_info():
    s = Solver(Params(kappa=4.0, rcut=3.0, order=8))
    scr, dscr, rlo, dri = s.split_table()
    assert scr.shape == (4096, 5) and abs(scr[0, 0] - 1.0) < 1e-14
    s.set_split_table(scr * 0.5, dscr, rlo, dri)
    scr2, *_ = s.split_table()
    np.testing.assert_array_equal(scr2, scr * 0.5)
    inf = s.info()
    assert inf.order == 8 and inf.window_degree == _lib.load().esp_default_window_degree(8)
    P, D, w = s.window()
    assert P == 8 and w.shape == (8, D + 1)
