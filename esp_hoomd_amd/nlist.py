"""Host-side neighbour-list construction in the HOOMD layout (full list).

``head_list[i]`` is the offset of particle i's neighbours in ``nlist``; ``n_neigh[i]`` their
count.  Built with a periodic k-d tree; used to feed the device pair kernel and by tests.
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree


def build_full_nlist(pos: np.ndarray, box, r_list: float, exclusions=None):
    """Full neighbour list of all pairs with minimum-image distance < r_list.

    pos: (N, 3) positions in a box centred on the origin; box: (Lx, Ly, Lz) or HOOMD's
    (Lx, Ly, Lz, xy, xz, yz) with tilt factors (a tilted box is searched pair by pair).
    exclusions: optional iterable of (i, j) pairs removed from the list.
    Returns (n_neigh uint32[N], nlist uint32[M], head_list uint64[N]).
    """
    pos = np.asarray(pos, dtype=np.float64)
    N = len(pos)
    L = np.asarray(box[:3], dtype=np.float64)
    tilt = np.asarray((list(box) + [0.0, 0.0, 0.0])[3:6], dtype=np.float64)
    if np.any(tilt != 0.0):
        pairs = _tilted_pairs(pos, L, tilt, r_list)
    else:
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


def min_image_tilted(d: np.ndarray, L: np.ndarray, tilt: np.ndarray) -> np.ndarray:
    """HOOMD's BoxDim::minImage for separations d (..., 3) in a box (Lx, Ly, Lz) with tilt
    factors (xy, xz, yz): a3 by rint(d_z / Lz), then a2 by rint(d_y / Ly), then a1."""
    d = np.array(d, dtype=np.float64, copy=True)
    xy, xz, yz = tilt
    iz = np.rint(d[..., 2] / L[2])
    d[..., 2] -= L[2] * iz
    d[..., 1] -= yz * L[2] * iz
    d[..., 0] -= xz * L[2] * iz
    iy = np.rint(d[..., 1] / L[1])
    d[..., 1] -= L[1] * iy
    d[..., 0] -= xy * L[1] * iy
    d[..., 0] -= L[0] * np.rint(d[..., 0] / L[0])
    return d


def _tilted_pairs(pos, L, tilt, r_list):
    """All pairs i < j within r_list at their nearest image in a tilted box (all pairs tested;
    meant for the small systems of tests).  r_list must be below half the box's smallest
    perpendicular width, so that the nearest image is the only one within reach."""
    xy, xz, yz = tilt
    n1 = np.sqrt(1.0 + xy * xy + (xy * yz - xz) ** 2) / L[0]
    n2 = np.sqrt(1.0 + yz * yz) / L[1]
    if r_list >= 0.5 / max(n1, n2, 1.0 / L[2]):
        raise ValueError("r_list must be below half the smallest perpendicular width of the box")
    N = len(pos)
    out = []
    for i0 in range(0, N, 512):
        blk = pos[i0:i0 + 512]
        d = min_image_tilted(blk[:, None, :] - pos[None, :, :], L, tilt)
        r2 = (d * d).sum(axis=-1)
        ii, jj = np.nonzero(r2 < r_list * r_list)
        ii = ii + i0
        keep = ii < jj
        out.append(np.stack([ii[keep], jj[keep]], axis=1))
    return np.concatenate(out) if out else np.zeros((0, 2), np.int64)


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
