"""Serving: continuous batching of MD requests.

* A request's trajectory is the same alone and in a mix of concurrently resident requests,
  bit for bit in both precisions: the mesh accumulates in fixed point (int32/int64 cells in
  mixed precision, int64 in FP64), neighbour rows list candidates in an order fixed by the
  positions, and every other sum is one thread's per particle.
* Admission by HBM footprint: with a budget that holds one request at a time, requests are
  deferred and served one after another, and every one finishes.
* Latency and throughput statistics are consistent.
"""
import numpy as np
import pytest

from conftest import rel_rms
from esp_hoomd_amd import _lib
from esp_hoomd_amd.serve import Request, Server, footprint_bytes, poisson_requests

pytestmark = pytest.mark.gpu


def _req(rid, N, steps, arrival=0.0, seed=None, density=0.1):
    return Request(rid=rid, N=N, density=density, steps=steps, arrival=arrival, seed=seed if seed is not None else 11 + rid)


@pytest.mark.parametrize("precision", [_lib.ESP_PREC_MIXED, _lib.ESP_PREC_FP64])
def test_request_result_is_the_same_alone_and_in_the_mix(precision):
    alone = Server(precision=precision, keep_results=True)
    a = _req(0, 3000, 25)
    alone.submit(a)
    alone.run()
    mix = Server(precision=precision, keep_results=True)
    a2 = _req(0, 3000, 25)
    others = [_req(1, 5000, 40, density=0.15), _req(2, 1200, 10, density=0.05), _req(3, 8000, 30, arrival=0.01)]
    for r in [a2] + others:
        mix.submit(r)
    st = mix.run()
    print(f"mix stats {st}", flush=True)
    assert st["requests_done"] == 4 and st["mean_active"] > 1.5
    print(f"precision {precision}: final forces alone vs mix, relative RMS "
          f"{rel_rms(a2.final_forces[:, :3], a.final_forces[:, :3]):.3e}", flush=True)
    assert np.array_equal(a2.first_forces, a.first_forces)
    assert np.array_equal(a2.final_forces, a.final_forces)
    assert np.array_equal(a2.final_positions, a.final_positions)


def test_admission_by_footprint_serialises_when_memory_is_short():
    reqs = [_req(i, 4000, 8, arrival=0.0) for i in range(3)]
    srv0 = Server()
    p, nbar = srv0._params_for(4000, (4000 / 0.1) ** (1 / 3))
    one = footprint_bytes(4000, p, nbar)
    srv = Server(hbm_budget=int(one * 1.5))
    for r in reqs:
        srv.submit(r)
    st = srv.run()
    assert st["requests_done"] == 3
    assert st["peak_resident_bytes"] <= 1.5 * one
    assert st["admissions_deferred_for_memory"] >= 1
    assert max(srv.tick_active) == 1


def test_poisson_stream_statistics():
    reqs = poisson_requests(12, rate=200.0, seed=3, N_range=(1000, 20000), steps_range=(5, 40))
    srv = Server()
    for r in reqs:
        srv.submit(r)
    st = srv.run()
    assert st["requests_done"] == 12
    assert st["solves"] == sum(r.steps for r in reqs)
    assert 0 < st["step_latency_ms_p50"] <= st["step_latency_ms_p99"]
    assert st["particle_steps_per_s"] > 0
    assert st["batch"]["replays"] >= 1


def test_served_results_meet_the_target_with_a_request_near_200k():
    """A stream with a request of 199,998 particles among smaller ones: the server calibrates its
    accuracy model on probe requests, and every result's force error at its final positions,
    against the tight FP64 reference, is within 1e-5."""
    srv = Server(precision=_lib.ESP_PREC_MIXED, keep_final_positions=True)
    scales = srv.calibrate()
    reqs = [_req(0, 199_998, 12, density=0.08), _req(1, 1500, 10, density=0.2),
            _req(2, 30_000, 20, density=0.05, arrival=0.005), _req(3, 6000, 15, density=0.12, arrival=0.01)]
    for r in reqs:
        srv.submit(r)
    st = srv.run()
    errs = srv.verify_results()
    print(f"calibrated scales {scales}; served errors {[f'{e:.2e}' for e in errs]}; "
          f"params {[(r.N, r.params.nx, r.params.order, round(r.params.rcut, 3)) for r in srv.done]}", flush=True)
    assert st["requests_done"] == 4
    assert max(errs) <= 1e-5


def _lists_built(srv):
    return sum(r.nlist.builds for r in srv.active) if srv.active else None


@pytest.mark.parametrize("precision", [_lib.ESP_PREC_MIXED, _lib.ESP_PREC_FP64])
def test_predicted_list_rebuilds_equal_checked_rebuilds(precision):
    """The tick decides its list rebuilds from the previous tick's prediction of the position
    its drift will produce; the trajectories equal, bit for bit, those of ticks that check the
    displacements after the drift and wait for the result (long enough for several rebuilds)."""
    out = {}
    for predict in (True, False):
        srv = Server(precision=precision, keep_results=True, predict_lists=predict)
        reqs = [_req(0, 2500, 90, density=0.2), _req(1, 4000, 70, density=0.1, arrival=0.002),
                _req(2, 1500, 80, density=0.15)]
        for r in reqs:
            srv.submit(r)
        srv.run()
        out[predict] = {r.rid: (r.final_positions.copy(), r.final_forces.copy()) for r in srv.done}
    for rid in out[True]:
        assert np.array_equal(out[True][rid][0], out[False][rid][0])
        assert np.array_equal(out[True][rid][1], out[False][rid][1])


def test_overflowing_list_rebuild_repeats_the_step():
    """A request whose list rows are set far too short: its first rebuild inside a tick
    overflows, the tick's step for that request is repeated alone from the state copied before
    the tick with grown rows, and its trajectory equals the one of a server that never
    overflowed, bit for bit."""
    class Shrinking(Server):
        shrunk = False

        def _admit(self, now):
            changed = super()._admit(now)
            for r in self.active:
                if r.rid == 0 and not self.shrunk:
                    r.nlist.set_capacity(8, 8)
                    self.shrunk = True
            return changed

    ref = Server(keep_results=True)
    a = _req(0, 3000, 60, density=0.2)
    ref.submit(a)
    ref.submit(_req(1, 2000, 40))
    ref.run()
    srv = Shrinking(keep_results=True)
    a2 = _req(0, 3000, 60, density=0.2)
    srv.submit(a2)
    srv.submit(_req(1, 2000, 40))
    st = srv.run()
    print(f"steps repeated: {st['steps_repeated_for_list_overflow']}", flush=True)
    assert st["steps_repeated_for_list_overflow"] >= 1
    assert np.array_equal(a2.final_positions, a.final_positions)
    assert np.array_equal(a2.final_forces, a.final_forces)


@pytest.mark.parametrize("precision", [_lib.ESP_PREC_MIXED, _lib.ESP_PREC_FP64])
def test_reused_solver_and_list_give_the_same_trajectory(precision):
    """A request admitted after another one retired takes over its solver and neighbour list
    (device arrays reused, parameters and mesh different); its trajectory equals, bit for bit,
    the one it has with fresh ones."""
    fresh = Server(precision=precision, keep_results=True, reuse_contexts=False)
    a = _req(0, 2600, 45, density=0.12)
    fresh.submit(a)
    fresh.run()
    srv = Server(precision=precision, keep_results=True)
    # a larger system on a larger mesh (a context too small for a request is not reused)
    srv.submit(_req(1, 5000, 20, density=0.1))
    srv.run()
    a2 = _req(0, 2600, 45, density=0.12)
    srv.submit(a2)
    st = srv.run()
    assert st["contexts_reused"] >= 1
    assert np.array_equal(a2.first_forces, a.first_forces)
    assert np.array_equal(a2.final_positions, a.final_positions)
    assert np.array_equal(a2.final_forces, a.final_forces)


def test_failing_requests_do_not_disturb_the_others():
    """A request with an invalid specification and one whose footprint exceeds the whole HBM
    budget fail with their reason; the request served beside them computes the same trajectory
    as alone."""
    ref = Server(keep_results=True)
    a = _req(0, 2000, 30)
    ref.submit(a)
    ref.run()
    srv0 = Server()
    p, nbar = srv0._params_for(2000, (2000 / 0.1) ** (1 / 3))
    budget = 4 * footprint_bytes(2000, p, nbar)
    srv = Server(keep_results=True, hbm_budget=budget)
    bad = _req(1, 2000, 30, density=-0.1)
    big = _req(2, 150_000, 10)
    a2 = _req(0, 2000, 30)
    for r in (bad, big, a2):
        srv.submit(r)
    st = srv.run()
    print(f"failures: {bad.error!r}; {big.error!r}", flush=True)
    assert st["requests_failed"] == 2 and st["requests_done"] == 1
    assert bad.error and "density" in bad.error
    assert big.error and "budget" in big.error
    assert np.array_equal(a2.final_positions, a.final_positions)
    assert np.array_equal(a2.final_forces, a.final_forces)
