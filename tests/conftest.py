"""Shared test configuration: markers, import paths and system generators."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running test")


def _gpu_present() -> bool:
    try:
        from esp_hoomd_amd import _lib
        return _lib.gpu_available()
    except ImportError:
        return False


def pytest_collection_modifyitems(config, items):
    if any("gpu" in item.keywords for item in items):
        if not _gpu_present():
            skip = pytest.mark.skip(reason="no GPU visible")
            for item in items:
                if "gpu" in item.keywords:
                    item.add_marker(skip)


def random_electrolyte(N, L, seed=0, min_dist=0.0, q2_frac=0.0):
    """Neutral random +-1 (and optionally +-2) charges uniform in a box centred on 0."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-L / 2, L / 2, (N, 3))
    if min_dist > 0:
        from scipy.spatial import cKDTree
        for _ in range(200):
            t = cKDTree(pos + L / 2, boxsize=L)
            pairs = t.query_pairs(min_dist, output_type="ndarray")
            if len(pairs) == 0:
                break
            bad = np.unique(pairs[:, 1])
            pos[bad] = rng.uniform(-L / 2, L / 2, (len(bad), 3))
    q = np.where(np.arange(N) % 2 == 0, 1.0, -1.0)
    if q2_frac > 0:
        n2 = int(N * q2_frac) // 2 * 2
        q[:n2] *= 2.0
    rng.shuffle(q)
    return pos, q


def box_matrix(box):
    """Columns a1, a2, a3 of a HOOMD box (Lx, Ly, Lz, xy, xz, yz)."""
    Lx, Ly, Lz, xy, xz, yz = (list(box) + [0.0, 0.0, 0.0])[:6]
    return np.array([[Lx, xy * Ly, xz * Lz], [0.0, Ly, yz * Lz], [0.0, 0.0, Lz]])


def box_from_matrix(B):
    """(Lx, Ly, Lz, xy, xz, yz) of an upper-triangular lattice matrix."""
    Lx, Ly, Lz = B[0, 0], B[1, 1], B[2, 2]
    return (Lx, Ly, Lz, B[0, 1] / Ly, B[0, 2] / Lz, B[1, 2] / Lz)


def random_electrolyte_box(N, box, seed=0, min_dist=0.0):
    """Neutral random +-1 charges uniform in a (tilted) HOOMD box centred on 0."""
    from esp_hoomd_amd.nlist import min_image_tilted
    rng = np.random.default_rng(seed)
    B = box_matrix(box)
    L = np.asarray(box[:3], float)
    tilt = np.asarray((list(box) + [0.0, 0.0, 0.0])[3:6], float)
    u = rng.uniform(0.0, 1.0, (N, 3))
    for _ in range(500):
        pos = (u - 0.5) @ B.T
        if min_dist <= 0:
            break
        d = min_image_tilted(pos[:, None, :] - pos[None, :, :], L, tilt)
        r2 = (d * d).sum(-1) + np.eye(N) * 1e9
        bad = np.unique(np.nonzero(r2 < min_dist * min_dist)[1])
        if len(bad) == 0:
            break
        u[bad] = rng.uniform(0.0, 1.0, (len(bad), 3))
    q = np.where(np.arange(N) % 2 == 0, 1.0, -1.0)
    rng.shuffle(q)
    return pos, q


def rel_rms(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.sqrt((b * b).sum())
    return np.sqrt(((a - b) ** 2).sum()) / (den if den > 0 else 1.0)
