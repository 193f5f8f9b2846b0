"""Host-side neighbour-list construction in the HOOMD layout (full list).

``head_list[i]`` is the offset of particle i's neighbours in ``nlist``; ``n_neigh[i]`` their
count.  Built with a periodic k-d tree; used to feed the device pair kernel and by tests.
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree


def build_full_nlist(pos: np.ndarray, box, r_list: float, exclusions=None):
    """Full neighbour list of all pairs with minimum-image distance < r_list.

    pos: (N, 3) positions in a box centred on the origin; box: (Lx, Ly, Lz).
    exclusions: optional iterable of (i, j) pairs removed from the list.
    Returns (n_neigh uint32[N], nlist uint32[M], head_list uint64[N]).
    """
    pos = np.asarray(pos, dtype=np.float64)
    N = len(pos)
    L = np.asarray(box[:3], dtype=np.float64)
    if np.any(r_list >= 0.5 * L):
        raise ValueError("r_list must be below half the box length")
    frac = pos / L + 0.5
    frac -= np.floor(frac)
    p = frac * L
    p = np.minimum(p, np.nextafter(L, 0))
    tree = cKDTree(p, boxsize=L)
    pairs = tree.query_pairs(r_list, output_type="ndarray")
    if exclusions is not None and len(pairs):
        ex = set()
        for i, j in exclusions:
            ex.add((min(i, j), max(i, j)))
        keep = np.array([(int(a), int(b)) not in ex for a, b in pairs], dtype=bool)
        pairs = pairs[keep]
    i = np.concatenate([pairs[:, 0], pairs[:, 1]]) if len(pairs) else np.zeros(0, np.int64)
    j = np.concatenate([pairs[:, 1], pairs[:, 0]]) if len(pairs) else np.zeros(0, np.int64)
    order = np.lexsort((j, i))
    i = i[order]
    j = j[order]
    n_neigh = np.bincount(i, minlength=N).astype(np.uint32)
    head = np.zeros(N, dtype=np.uint64)
    if N > 1:
        head[1:] = np.cumsum(n_neigh[:-1], dtype=np.uint64)
    return n_neigh, j.astype(np.uint32), head


def build_exclusion_list(N: int, pairs):
    """HOOMD-layout exclusion list: (n_ex uint32[N], ex_list uint32[N * pitch], pitch)."""
    per = [[] for _ in range(N)]
    for a, b in pairs:
        per[a].append(b)
        per[b].append(a)
    pitch = max(1, max((len(x) for x in per), default=1))
    n_ex = np.array([len(x) for x in per], dtype=np.uint32)
    ex = np.zeros((N, pitch), dtype=np.uint32)
    for a, lst in enumerate(per):
        ex[a, : len(lst)] = lst
    return n_ex, ex.reshape(-1), pitch
