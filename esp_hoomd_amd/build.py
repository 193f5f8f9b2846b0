"""Build the in-tree HIP library ``esp_hoomd_amd/libesp_mi355x.so`` for gfx950.

Every ``csrc/*.hip`` and ``csrc/*.cpp`` file is compiled with hipcc into ``_build/`` and
linked into one shared library next to this file.  Run ``python -m esp_hoomd_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as _cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG.parent / "_build"
LIB = PKG / "libesp_mi355x.so"
ARCH = os.environ.get("ESP_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-O3",
    "-fPIC",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{PKG.parent / 'include'}",
    f"-I{CSRC}",
]


def _sources():
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + src.suffix.replace(".", "_") + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd.insert(1, "-xhip")
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile all sources and link the shared library; returns its path."""
    BUILD.mkdir(exist_ok=True)
    srcs = _sources()
    workers = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with _cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


# Host-only AddressSanitizer: each -fsanitize option goes to the host compilation (device code
# is built exactly as in the library).
ASAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
ASAN_DRIVER_SRC = PKG.parent / "tests" / "asan" / "abi_driver.cpp"
ASAN_DRIVER = PKG.parent / "tests" / "asan" / "abi_driver"


def build_asan(verbose: bool = False) -> Path:
    """Link tests/asan/abi_driver: the C-ABI driver over the library's host sources (csrc/*.cpp)
    compiled with host AddressSanitizer and its kernel objects (csrc/*.hip) as built by
    build().  The executable carries the ASan runtime itself, so it needs no preloading."""
    adir = BUILD / "asan"
    adir.mkdir(parents=True, exist_ok=True)
    kernels = [BUILD / (src.stem + "_hip.o") for src in sorted(CSRC.glob("*.hip"))]
    if not all(o.exists() for o in kernels):
        build()

    def compile_host(src: Path) -> Path:
        obj = adir / (src.stem + "_asan.o")
        cmd = [HIPCC, *COMMON, *ASAN_HOST, "-c", str(src), "-o", str(obj)]
        if src.parent == CSRC:
            cmd.insert(1, "-xhip")
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc (host ASan) failed for {src.name}:\n{res.stderr}")
        return obj

    srcs = sorted(CSRC.glob("*.cpp")) + [ASAN_DRIVER_SRC]
    with _cf.ThreadPoolExecutor(max_workers=min(len(srcs), 8)) as ex:
        objs = list(ex.map(compile_host, srcs))
    tmp = ASAN_DRIVER.with_suffix(".tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", *ASAN_HOST, "-o", str(tmp), *map(str, objs), *map(str, kernels)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"ASan driver link failed:\n{res.stderr}")
    os.replace(tmp, ASAN_DRIVER)
    if verbose:
        print(f"built {ASAN_DRIVER}")
    return ASAN_DRIVER


def build_oracle(verbose: bool = False) -> Path:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    odir = PKG.parent / "oracle"
    res = subprocess.run(["make", "-C", str(odir), "-s"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stderr}")
    if verbose:
        print(f"built {odir / 'liboracle.so'}")
    return odir / "liboracle.so"


if __name__ == "__main__":
    build(verbose=True)
    if "--oracle" in sys.argv:
        build_oracle(verbose=True)
    if "--asan" in sys.argv:
        build_asan(verbose=True)
