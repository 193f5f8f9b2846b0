"""esp_pair_forces_noalloc (the pair correction with its scratch in the library's static device
array) against esp_pair_forces: the same kernels on the same inputs give the same bits, on two
calls in a row with the second system smaller; the entry point refuses an uninitialised context
and a system beyond its maximum size."""
import ctypes as C

import numpy as np
import pytest

from conftest import random_electrolyte
from esp_cases import make_params
from esp_hoomd_amd import _lib
from esp_hoomd_amd.solver import DeviceSystem, Solver

pytestmark = pytest.mark.gpu


def _noalloc(solver, sysd):
    s = sysd.c_system()
    o = sysd.c_out(_lib.ESP_COMPUTE_PAIR)
    rc = _lib.load().esp_pair_forces_noalloc(solver.ctx, C.byref(s), C.byref(o), None)
    _lib.check(rc, solver.ctx)
    _lib.check(_lib.load().esp_device_synchronize())
    return {"force": sysd.force.numpy(), "virial": sysd.virial.numpy()[:, : sysd.N]}


@pytest.mark.parametrize("precision", [_lib.ESP_PREC_FP64, _lib.ESP_PREC_MIXED])
def test_noalloc_pair_equals_pair_forces_on_two_calls(precision):
    solver = Solver(make_params(32, 8, 12.0, 3.0, precision=precision))
    for k, (N, L) in enumerate([(6000, 20.0), (900, 10.0)]):  # the second call is smaller
        pos, q = random_electrolyte(N, L, seed=40 + k, q2_frac=0.2)
        excl = [(0, 1), (2, 7)] if k == 0 else None
        sysd = DeviceSystem(pos, q, (L, L, L), r_list=3.3, exclusions=excl)
        ref = {key: v.copy() for key, v in solver.pair_forces(sysd).items()}
        sysd.force.zero()
        sysd.virial.zero()
        got = _noalloc(solver, sysd)
        print(f"N={N}: max |F| {np.abs(ref['force'][:, :3]).max():.3e}", flush=True)
        assert np.abs(ref["force"][:, :3]).max() > 0
        np.testing.assert_array_equal(got["force"], ref["force"])
        np.testing.assert_array_equal(got["virial"], ref["virial"])


def test_noalloc_refuses_uninitialised_context_and_oversized_systems():
    lib = _lib.load()
    pos, q = random_electrolyte(64, 6.0, seed=3)
    sysd = DeviceSystem(pos, q, (6.0, 6.0, 6.0), r_list=2.3)
    fresh = Solver(make_params(16, 4, 8.0, 2.0))
    s = sysd.c_system()
    o = sysd.c_out(_lib.ESP_COMPUTE_PAIR)
    assert lib.esp_pair_forces_noalloc(fresh.ctx, C.byref(s), C.byref(o), None) == _lib.ESP_ERR_STATE
    fresh.pair_forces(sysd)  # initialises the context
    big = sysd.c_system()
    big.N = _lib.ESP_NOALLOC_MAX_PARTICLES + 1  # refused before any array is read
    assert lib.esp_pair_forces_noalloc(fresh.ctx, C.byref(big), C.byref(o), None) == _lib.ESP_ERR_RANGE
    assert b"ESP_NOALLOC_MAX_PARTICLES" in lib.esp_last_error(fresh.ctx)
    assert lib.esp_pair_forces_noalloc(fresh.ctx, C.byref(s), C.byref(o), None) == _lib.ESP_OK
