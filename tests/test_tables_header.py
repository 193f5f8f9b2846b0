"""The PSWF_Coeffs.h-format generator and reader (host only)."""
import numpy as np
import pytest

import oracle as O
from esp_hoomd_amd import _lib, tables
from esp_hoomd_amd.solver import Params, Solver, split_eval


@pytest.mark.parametrize("c,n", [(9.0, 512), (13.5, 4096)])
def test_header_round_trip_is_exact(c, n):
    text = tables.write_header(None, c, n_segs=n, order=8, sigma=1.3)
    h = tables.read_header(text)
    scr, dscr, rlo, dri = tables.split_tables(c, n)
    assert h["PSWF_N_SEGS"] == n and h["PSWF_ORDER"] == 8 and h["PSWF_C"] == c and h["PSWF_RCUT"] == 1.0
    np.testing.assert_array_equal(h["PSWF_SCREEN_COEFFS"], scr)
    np.testing.assert_array_equal(h["PSWF_DSCREEN_COEFFS"], dscr)
    np.testing.assert_array_equal(h["PSWF_R_LO"], rlo)
    np.testing.assert_array_equal(h["PSWF_DR_INV"], dri)
    assert h["PSWF_WINDOW_COEFFS"].size == 8 * (h["PSWF_WINDOW_DEGREE"] + 1)
    assert "namespace esp_pswf_coeffs" in text and text.count("{") == text.count("}")


def test_emitted_table_evaluates_the_split():
    """The reference lookup (oracle, golden-pinned) on the emitted arrays reproduces the
    library's split S(s), dS/ds and the endpoint values S(0) = 1, S(1) = 0."""
    c = 12.0
    h = tables.read_header(tables.write_header(None, c, n_segs=1024))
    s = np.random.default_rng(1).uniform(1e-3, 0.999, 500)
    L, nd, S, dS = O.split_eval(h["PSWF_SCREEN_COEFFS"], h["PSWF_DSCREEN_COEFFS"], h["PSWF_R_LO"], h["PSWF_DR_INV"], 1.0, s)
    S_ref, dS_ref = split_eval(_lib.ESP_SPLIT_PSWF, c, 1.0, s)
    assert np.abs(S - S_ref).max() < 1e-11
    assert np.abs(dS - dS_ref).max() < 1e-9 * np.abs(dS_ref).max()
    assert abs(h["PSWF_SCREEN_COEFFS"][0, 0] - 1.0) < 1e-12


def test_reader_accepts_the_reference_layout():
    """A header written in the reference's layout (its names, namespace and comment style)."""
    n = 4
    rows = "\n".join("    " + ", ".join(f"{0.1 * (i + 1) + 0.01 * k:.17e}" for k in range(5)) + "," for i in range(n))
    text = f"""// PSWF_Coeffs.h -- AUTO-GENERATED.
#pragma once
namespace hoomd
{{
namespace md
{{
namespace esp_pswf_coeffs
{{
inline constexpr int PSWF_N_SEGS = {n};
inline constexpr int PSWF_ORDER = 6;
inline constexpr double PSWF_C = 9.00000000000000711e+00;
inline constexpr double PSWF_RCUT = 1.00000000000000000e+00;
// Row-major [seg][coeff], seg in [0,{n}), coeff in [0,5)
inline constexpr double PSWF_SCREEN_COEFFS[{5 * n}] = {{
{rows}
}};
inline constexpr double PSWF_DSCREEN_COEFFS[{5 * n}] = {{
{rows}
}};
inline constexpr double PSWF_R_LO[{n}] = {{
    0.0, 0.25, 0.5, 0.75,
}};
inline constexpr double PSWF_DR_INV[{n}] = {{
    4.0, 4.0, 4.0, 4.0,
}};
inline constexpr double PSWF_GF_DENOM_COEFFS[5] = {{
    1.0, 2.0, 3.0, 4.0, 5.0,
}};
}} // namespace esp_pswf_coeffs
}} // namespace md
}} // namespace hoomd
"""
    h = tables.read_header(text)
    assert h["PSWF_SCREEN_COEFFS"].shape == (4, 5) and h["PSWF_SCREEN_COEFFS"][2, 1] == pytest.approx(0.31)
    np.testing.assert_array_equal(h["PSWF_R_LO"], [0, 0.25, 0.5, 0.75])
    with pytest.raises(ValueError):
        tables.read_header(text.replace("PSWF_DR_INV[4]", "PSWF_DR_INV[5]"))


def test_load_into_solver():
    s = Solver(Params(order=6, kappa=3.0, rcut=3.0))
    text = tables.write_header(None, 7.5, n_segs=256)
    tables.load_into(s, text)
    scr, dscr, rlo, dri = s.split_table()
    h = tables.read_header(text)
    np.testing.assert_array_equal(np.asarray(scr).reshape(-1, 5), h["PSWF_SCREEN_COEFFS"])
    np.testing.assert_array_equal(rlo, h["PSWF_R_LO"])
