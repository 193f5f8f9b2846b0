"""Per-stage HIP-event timing (esp_set_profiling): level 1 times every stage of the step, level 2
only the pair stage (the other kernels run unmarked); neither changes the results."""
import pytest

from conftest import random_electrolyte, rel_rms
from esp_cases import make_params, run_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_profiling_levels(precision):
    N, L = 400, 10.0
    pos, q = random_electrolyte(N, L, seed=9)
    params = make_params(24, 8, 12.0, 3.0, precision=precision)
    solver, sysd, base = run_gpu(params, pos, q, (L, L, L))
    for level in (1, 2):
        print(f"precision {precision} level {level}", flush=True)
        solver.set_profiling(level)
        solver.reset_stage_times()
        for _ in range(3):
            res = solver.compute(sysd)
        st = solver.stage_times()
        solver.set_profiling(0)
        assert st["pair"][1] == 3 and st["pair"][0] > 0
        others = {k: c for k, (_, c) in st.items() if k != "pair"}
        if level == 1:
            assert all(c == 3 for c in others.values()), others
        else:
            assert all(c == 0 for c in others.values()), others
        assert rel_rms(res["force"], base["force"]) < 1e-13
        assert rel_rms(res["virial"], base["virial"]) < 1e-13
    with pytest.raises(Exception):
        solver.set_profiling(3)
