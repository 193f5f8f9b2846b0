"""Step graphs (esp_set_graph): the whole single-system step captured once and replayed.

Each replay must equal an eager evaluation of the same inputs bit for bit (both precisions
are deterministic), with new positions written into the same arrays between replays and
eager work and synchronisations in between; a change of the call's arguments captures again,
and profiled or mesh-only calls stay eager."""
import numpy as np
import pytest

from conftest import random_electrolyte
from esp_cases import make_params
from esp_hoomd_amd import _lib
from esp_hoomd_amd.solver import DeviceSystem, Solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replays_equal_eager_steps(precision):
    N, L = 1500, 14.0
    pos, q = random_electrolyte(N, L, seed=23)
    params = make_params(32, 8, 12.0, 3.5, precision=precision)
    gsys = DeviceSystem(pos, q, (L, L, L), r_list=3.8)
    esys = DeviceSystem(pos, q, (L, L, L), r_list=3.8)
    graph, eager = Solver(params), Solver(params)
    graph.set_graph(True)
    rng = np.random.default_rng(3)
    for rep in range(4):
        # moves well inside the 0.3 skin keep both neighbour lists valid
        p = pos + rng.uniform(-0.04, 0.04, pos.shape) * rep
        gsys.set_positions(p)
        esys.set_positions(p)
        print(f"precision {precision} replay {rep}", flush=True)
        a = graph.compute(gsys)
        b = eager.compute(esys)
        for k in ("force", "virial", "ext"):
            assert np.array_equal(a[k], b[k]), (rep, k)
    assert graph.graph_stats() == (1, 4)
    assert eager.graph_stats() == (0, 0)
    # another output set is another call: captured again
    other = DeviceSystem(pos, q, (L, L, L), r_list=3.8)
    c = graph.compute(other)
    assert graph.graph_stats() == (2, 5)
    assert np.array_equal(c["force"], eager.compute(DeviceSystem(pos, q, (L, L, L), r_list=3.8))["force"])
    # mesh-only (clears the outputs with memsets) and profiled steps run eagerly
    graph.compute(gsys, flags=_lib.ESP_COMPUTE_MESH | _lib.ESP_COMPUTE_ENERGY_VIRIAL)
    graph.set_profiling(1)
    graph.compute(gsys)
    graph.set_profiling(0)
    assert graph.graph_stats() == (2, 5)
