"""Host-side logic of the autotuner: the accuracy model, FFT sizes and candidate constraints."""
import math

import numpy as np
import pytest

from esp_hoomd_amd import _lib, tune
from esp_hoomd_amd.accuracy import fft_size_at_least, reference_params


def test_fft_sizes_are_235_smooth_and_minimal():
    for n in range(1, 400):
        m = fft_size_at_least(n, 1)
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        assert k == 1 and m >= n
        assert all(fft_size_at_least(j, 1) != j for j in range(n, m))  # nothing smooth in between


def test_window_model_reproduces_its_table_and_is_monotone():
    for P, row in tune.WINDOW_ERROR.items():
        for s, e in zip(tune.SIGMA_GRID, row):
            assert tune.window_error(P, s) == pytest.approx(e, rel=1e-9)
    for P in (6, 8, 9, 10, 12):
        v = [tune.window_error(P, s) for s in np.linspace(1.15, 3.0, 30)]
        assert all(a >= b for a, b in zip(v, v[1:]))
    assert tune.window_error(9, 1.5) < tune.window_error(8, 1.5)
    assert tune.split_error(12) < tune.split_error(10)


@pytest.mark.parametrize("precision", [_lib.ESP_PREC_FP64, _lib.ESP_PREC_MIXED])
def test_best_split_meets_goal_with_smallest_product(precision):
    for P in (6, 8, 10, 12):
        r = tune.best_split(P, 1e-5, precision)
        if r is None:
            assert P == 6
            continue
        c, sigma = r
        assert tune.predicted_error(c, sigma, P, precision) <= 1e-5 * (1 + 1e-9)
        # no other grid point with a smaller sigma * c meets the goal
        for c2 in np.arange(4.0, 32.0, 0.5):
            for s2 in np.linspace(tune.SIGMA_MIN, 4.0, 40):
                if s2 * c2 < sigma * c * (1 - 1e-3):
                    assert tune.predicted_error(c2, s2, P, precision) > 1e-5


def test_tune_candidates_respect_constraints():
    N, L = 20000, (20000 / 0.1) ** (1 / 3)
    cands = tune.tune(N, (L, L, L), 1e-5, _lib.ESP_PREC_MIXED, orders=(6, 8, 10), mesh_range=(16, 64), top=50)
    assert cands and all(a.predicted_ms <= b.predicted_ms for a, b in zip(cands, cands[1:]))
    for ch in cands:
        p = ch.params
        assert 16 <= p.nx <= 64 and p.nx == p.ny == p.nz and p.order in (6, 8, 10)
        assert p.order <= p.nx // 2
        assert p.rcut + 0.4 <= 0.5 * L
        assert ch.predicted_error <= 1e-5 / 1.25 * (1 + 1e-9)
        assert p.kappa * p.rcut == pytest.approx(ch.c)
        sigma = math.pi / ((L / p.nx) * p.kappa)
        assert sigma >= ch.sigma * (1 - 1e-9)


def test_unreachable_target_raises():
    with pytest.raises(ValueError):
        tune.tune(100, (5.0, 5.0, 5.0), 1e-12, _lib.ESP_PREC_MIXED, orders=(6,), mesh_range=(16, 20))


def test_reference_params_are_tight():
    p = reference_params((215.443,) * 3)
    assert p.order == 12 and p.precision == _lib.ESP_PREC_FP64
    h = 215.443 / p.nx
    sigma = math.pi / (h * p.kappa)
    assert sigma >= 1.5 and p.kappa * p.rcut == pytest.approx(20.0)
    assert tune.predicted_error(20.0, sigma, 12, _lib.ESP_PREC_FP64) < 4e-9
    small = reference_params((10.0, 11.0, 12.0))
    assert small.rcut <= 4.5 and min(small.nx, small.ny, small.nz) >= 24


def test_list_upkeep_enters_the_step_and_shortens_the_cutoff():
    """The amortised neighbour-list build is part of a simulation step: step = force + build /
    rebuild_every, and charging it never makes the tuner pick a longer cutoff (the list grows
    as (r_c + skin)^3; the mesh pays for a shorter one)."""
    N, L = 1_000_000, (1_000_000 / 0.1) ** (1 / 3)
    with_list = tune.CostModel()
    no_list = tune.CostModel(rebuild_every=0.0)
    f = no_list.step_ms(N, 144 ** 3, 10, 170.0, _lib.ESP_PREC_MIXED)
    assert f == pytest.approx(with_list.force_ms(N, 144 ** 3, 10, 170.0, _lib.ESP_PREC_MIXED))
    s = with_list.step_ms(N, 144 ** 3, 10, 170.0, _lib.ESP_PREC_MIXED)
    assert s == pytest.approx(f + with_list.nlist_ms(N, 170.0) / with_list.rebuild_every)
    # 1e6 particles, 178 entries: about 3 ms per build, as measured
    assert with_list.nlist_ms(N, 178.0) == pytest.approx(3.07, rel=0.05)
    for prec in (_lib.ESP_PREC_FP64, _lib.ESP_PREC_MIXED):
        a = tune.tune(N, (L, L, L), 1e-5, prec, cost=with_list)[0]
        b = tune.tune(N, (L, L, L), 1e-5, prec, cost=no_list)[0]
        assert a.params.rcut <= b.params.rcut + 1e-12
        print(prec, a.describe(), b.describe())
