"""The Python surface mirrors the reference's esp.py (src/esp.py:62-712): arguments, defaults,
validation errors, warnings, pre-attach errors and the coupling of Local to Spectral.
Host-only: nothing here touches a GPU."""
import warnings

import pytest

from esp_hoomd_amd import esp


def test_defaults_and_properties_before_attach():
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        s = esp.Spectral(nlist=None, default_r_cut=2.0)
    assert any("heuristic kappa" in str(x.message) for x in w)
    assert s.kappa == pytest.approx(1.6)
    assert s.resolution == (32, 32, 32) and s.order == 6 and s.r_cut == 2.0 and s.alpha == 0.0
    assert s._nlist_buffer == pytest.approx(0.2)
    with pytest.raises(RuntimeError):
        _ = s.q_sum
    with pytest.raises(RuntimeError):
        _ = s.q2_sum
    with pytest.raises(RuntimeError):
        s.set_params(order=4)
    s.invalidate()  # no-op before attach


@pytest.mark.parametrize("kw", [
    dict(default_r_cut=0.0), dict(default_r_cut=-1.0), dict(resolution=(32, 32)), dict(resolution=(32, 0, 32)),
    dict(order=0), dict(order=esp.MAX_ORDER + 1), dict(order=2.5), dict(kappa=-1.0), dict(alpha=-0.1),
    dict(n_table=8), dict(n_table=20.0),
])
def test_spectral_validation(kw):
    args = dict(nlist=None, default_r_cut=2.0, kappa=1.0)
    args.update(kw)
    with pytest.raises(ValueError):
        esp.Spectral(**args)


def test_local_validation_and_coupling():
    s = esp.Spectral(nlist=None, default_r_cut=2.0, kappa=1.5)
    with pytest.raises(ValueError):
        esp.Local(nlist=None, default_r_cut=2.0)
    with pytest.raises(TypeError):
        esp.Local(nlist=None, default_r_cut=2.0, spectral=object())
    with pytest.raises(ValueError):
        esp.Local(nlist=None, default_r_cut=2.0, spectral=s, mode="shift")
    with pytest.warns(DeprecationWarning):
        loc = esp.Local(nlist=None, default_r_cut=2.0, coulomb=s)
    assert s._pair is loc and loc.r_cut == 2.0
    loc.params[("A", "A")] = {}
    with pytest.warns(UserWarning):
        loc.params[("A", "B")] = {"epsilon": 1.0}
    assert loc.params[("A", "B")] == {}


def test_make_esp_forces_and_aliases():
    s, l = esp.make_esp_forces(nlist=None, r_cut=3.0, resolution=(16, 16, 16), order=8, kappa=4.0)
    assert isinstance(s, esp.Spectral) and isinstance(l, esp.Local) and s._pair is l
    assert esp.Coulomb is esp.Spectral and esp.Pair is esp.Local
    assert esp.MAX_ORDER >= 12 and esp.DEFAULT_TABLE_SEGMENTS == 512 and esp.MIN_TABLE_SEGMENTS == 16
    assert "Spectral" in repr(s) and "Local" in repr(l)
