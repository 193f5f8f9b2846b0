"""The committed direct-Ewald forces on the 1M-particle lattice subset (tests/golden/
ewald_subset_N1M.npz, written by tests/golden/make_ewald_subset.py): the generator still
produces the same particles, and an Ewald sum with a different splitting parameter reproduces
the stored forces (the sum's own truncation error is far below any tolerance used against it)."""
from pathlib import Path

import numpy as np

import oracle as O
from esp_hoomd_amd import workloads as W

FIX = Path(__file__).resolve().parent / "golden" / "ewald_subset_N1M.npz"


def test_fixture_matches_generator_and_second_splitting():
    d = np.load(FIX)
    pos, q, L = W.lattice_electrolyte(1_000_000, 0.1)
    assert str(d["digest"]) == W.array_digest(pos, q)
    assert np.array_equal(d["index"], W.subset_indices(len(pos), 10_000))
    idx = d["index"][::100]  # 100 targets
    rc = 75.0
    f, _, _ = O.ewald(pos, q, (L, L, L), 5.0 / rc, rc, 50.0 / rc, targets=idx, energy=False)
    ref = d["force"][::100]
    err = np.sqrt(((f - ref) ** 2).sum() / (ref ** 2).sum())
    assert err < 1e-9, err
