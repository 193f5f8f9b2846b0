"""The C-ABI under host AddressSanitizer and LeakSanitizer: tests/asan/abi_driver (built by
esp_hoomd_amd.build.build_asan, host sources instrumented, device code as in the library)
walks the entry points -- host table math, context lifetime and introspection and argument
errors here; force, pair, batch, neighbour-list and MD paths on a GPU -- and must end with no
sanitizer report and no failed check."""
import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "tests" / "asan" / "abi_driver"
SUPP = ROOT / "tests" / "asan" / "lsan.supp"


def _driver():
    if not DRIVER.exists():
        from esp_hoomd_amd import build
        build.build_asan()
    return DRIVER


def _run(*args, timeout):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:exitcode=23"
    env["LSAN_OPTIONS"] = f"suppressions={SUPP}:print_suppressions=0"
    res = subprocess.run([str(_driver()), *args], capture_output=True, text=True, timeout=timeout, env=env)
    print(res.stdout[-2000:], res.stderr[-4000:], flush=True)
    assert "AddressSanitizer" not in res.stderr and "LeakSanitizer" not in res.stderr
    assert res.returncode == 0
    return res.stdout


def test_host_paths_clean_under_asan():
    out = _run("--host-only", timeout=120)
    assert "host paths only, 0 failed checks" in out


@pytest.mark.gpu
def test_gpu_paths_clean_under_asan():
    out = _run(timeout=120)
    assert "host and GPU paths, 0 failed checks" in out
