"""Pins the product's host mathematics (prolate function, split, window) against scipy,
closed forms and independent quadrature."""
import numpy as np
import pytest
from scipy import integrate
from scipy.special import erfc, pro_ang1

import oracle as O
from esp_hoomd_amd import solver as S


@pytest.mark.parametrize("c", [3.0, 9.0, 14.0])
def test_prolate_matches_scipy(c):
    x = np.linspace(-0.95, 0.95, 17)
    psi, _ = S.prolate_eval(c, x)
    ref, _ = pro_ang1(0, 0, c, x)
    ratio = ref / psi
    assert np.ptp(ratio) < 1e-9 * np.abs(ratio).mean()


def test_prolate_integral_is_antiderivative():
    c = 10.0
    x = np.linspace(0, 1, 11)
    psi, integ = S.prolate_eval(c, x)
    for xi, Ii in zip(x, integ):
        val, _ = integrate.quad(lambda t: S.prolate_eval(c, np.array([t]))[0][0], 0, xi, epsabs=1e-14)
        assert abs(val - Ii) < 1e-12


@pytest.mark.parametrize("c", [6.0, 12.0, 20.0])
def test_pswf_split_endpoints_and_derivative(c):
    rc = 2.0
    s = np.linspace(0, 1, 41)
    Sv, dS = S.split_eval(0, c / rc, rc, s)
    assert abs(Sv[0] - 1.0) < 1e-14 and abs(Sv[-1]) < 1e-14
    h = 1e-6
    sp, _ = S.split_eval(0, c / rc, rc, s[1:-1] + h)
    sm, _ = S.split_eval(0, c / rc, rc, s[1:-1] - h)
    np.testing.assert_allclose((sp - sm) / (2 * h), dS[1:-1], atol=1e-8)
    assert np.all(np.diff(Sv) <= 1e-15)  # monotone decreasing


@pytest.mark.parametrize("c", [8.0, 14.0])
def test_pswf_fourier_against_radial_quadrature(c):
    # F(k) = k^2 * FT[(1 - S(r/rc)) / (4 pi r)] = k int_0^rc (1 - S) sin(kr) dr + cos(k rc)
    rc = 3.0
    k = np.array([0.3, 1.0, 2.5, 4.0, 6.0])
    F, dF = S.split_fourier(0, c / rc, rc, k)
    for ki, Fi in zip(k, F):
        f = lambda r: (1.0 - S.split_eval(0, c / rc, rc, np.array([r / rc]))[0][0]) * np.sin(ki * r)
        val, _ = integrate.quad(f, 0, rc, limit=200, epsabs=1e-14)
        assert abs(ki * val + np.cos(ki * rc) - Fi) < 1e-10
    h = 1e-5
    Fp, _ = S.split_fourier(0, c / rc, rc, k + h)
    Fm, _ = S.split_fourier(0, c / rc, rc, k - h)
    np.testing.assert_allclose((Fp - Fm) / (2 * h), dF, atol=1e-8)
    F0, _ = S.split_fourier(0, c / rc, rc, np.array([0.0]))
    assert abs(F0[0] - 1.0) < 1e-14


def test_split_constants():
    c, rc = 12.0, 3.0
    self_c, bg = S.split_constants(0, c / rc, rc)
    _, dS0 = S.split_eval(0, c / rc, rc, np.array([0.0]))
    assert abs(self_c + dS0[0] / rc) < 1e-14
    k = np.array([1e-3])
    F, _ = S.split_fourier(0, c / rc, rc, k)
    assert abs((1 - F[0]) / k[0] ** 2 - bg) < 1e-6 * bg
    # Gaussian split closed forms
    kap = 1.3
    self_g, bg_g = S.split_constants(1, kap, rc)
    assert abs(self_g - 2 * kap / np.sqrt(np.pi)) < 1e-15 and abs(bg_g - 1 / (4 * kap * kap)) < 1e-15
    s = np.linspace(0, 1, 9)
    Sg, _ = S.split_eval(1, kap, rc, s)
    np.testing.assert_allclose(Sg, erfc(kap * rc * s), rtol=1e-15)
    kk = np.linspace(0, 5, 7)
    Fg, _ = S.split_fourier(1, kap, rc, kk)
    np.testing.assert_allclose(Fg, np.exp(-kk ** 2 / (4 * kap * kap)), rtol=1e-15)


def test_split_table_reproduces_split():
    c, rc = 12.0, 3.5
    scr, dscr, rlo, dri = S.build_split_table(0, c / rc, rc, 4096)
    rng = np.random.default_rng(0)
    r = rng.uniform(1e-3, rc * (1 - 1e-9), 400)
    L, nd, Sv, dS = O.split_eval(scr, dscr, rlo, dri, rc, r)
    Sx, dSx = S.split_eval(0, c / rc, rc, r / rc)
    np.testing.assert_allclose(Sv, Sx, atol=1e-14)
    np.testing.assert_allclose(dS, dSx / rc, atol=1e-13)
    # continuity at every segment boundary
    ends = np.array([np.polyval(scr[i][::-1], 1.0) for i in range(4095)])
    assert np.abs(ends - scr[1:, 0]).max() < 1e-14
    np.testing.assert_array_equal(rlo, np.arange(4096) / 4096)


@pytest.mark.parametrize("P", [4, 6, 8, 10, 12])
def test_window_normalisation_symmetry_and_fourier(P):
    D, cw = S.default_window(P, 1.5)
    w = S.build_window(P, D, cw)
    th = np.array([0.0, 0.4, 1.3, np.pi])
    what = S.window_fourier(w, th)
    assert abs(what[0] - 1.0) < 1e-13
    ref = np.array([np.sqrt(O.window_fourier2(w, t)) for t in th])
    np.testing.assert_allclose(what, ref, rtol=1e-13)
    # mirror symmetry: tap P-1-k at z equals tap k at -z
    z = np.linspace(-1, 1, 7)
    for k in range(P):
        a = np.polyval(w[k][::-1], z)
        b = np.polyval(w[P - 1 - k][::-1], -z)
        np.testing.assert_allclose(a, b, atol=1e-15)
    # approximate partition of unity (aliasing error of the window)
    t = np.linspace(0, 1, 33)
    tot = sum(np.polyval(w[k][::-1], 2 * t - 1) for k in range(P))
    assert np.abs(tot - 1).max() < 10 ** (-0.9 * P + 1.5)
