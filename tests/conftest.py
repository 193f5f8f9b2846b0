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


def rel_rms(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.sqrt((b * b).sum())
    return np.sqrt(((a - b) ** 2).sum()) / (den if den > 0 else 1.0)
