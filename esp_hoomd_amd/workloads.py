"""Synthetic inputs of the benchmark workloads, seeded so that every caller (benchmark, tests,
fixture scripts) builds the same particles.

* ``small_electrolyte``: 1,000 charges (500 of +1, 500 of -1) uniform in a cubic box of edge
  12 (seed 0), with overlapping pairs (closer than 0.8 sigma) redrawn, velocities at T = 1
  (seed 1) -- the CPU-runnable case: mesh 32^3, P = 8, r_c = 3.0.
* ``lattice_electrolyte``: the 1,000,000-particle primitive-model electrolyte at number density
  0.1 (L = 215.44): a 100^3 simple-cubic lattice jittered by U(+-0.1 a) (seed 2), a random
  permutation of 500,000 +1 / 500,000 -1 charges (seed 3); ``equilibrate`` runs the Langevin
  equilibration (T = 1, gamma = 1, seed 4) on the GPU.
* ``batch_systems``: 512 independent electrolytes (seed 6), N log-uniform in [500, 50,000]
  rounded to even, +-1 with 10 % +-2 (each valence class neutral), density 0.1.
* ``lattice_positions`` / ``neutral_charges``: the generators of one served request (random
  sites of a jittered lattice, so MD starts without close contacts).
"""
from __future__ import annotations

import hashlib
import math
from typing import List, Tuple

import numpy as np

#: seeded subset on which 1M-particle forces are compared with the direct Ewald sum
SUBSET_SEED = 5


def lattice_electrolyte(N: int = 1_000_000, rho: float = 0.1, seed_pos: int = 2, seed_q: int = 3, jitter: float = 0.1):
    """Seeded jittered simple-cubic lattice of N = m^3 sites at number density rho, box centred
    on the origin, neutral +-1 charges in a random permutation.  Returns (pos, q, L)."""
    m = int(round(N ** (1.0 / 3.0)))
    if m ** 3 != N:
        raise ValueError("N must be a perfect cube")
    L = (N / rho) ** (1.0 / 3.0)
    a = L / m
    g = (np.arange(m) + 0.5) * a - L / 2
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    rng = np.random.default_rng(seed_pos)
    pos += rng.uniform(-jitter * a, jitter * a, pos.shape)
    q = np.where(np.arange(N) < N // 2, 1.0, -1.0)
    if N % 2:
        q[-1] = 0.0
    q = np.random.default_rng(seed_q).permutation(q)
    return pos, q, L


def subset_indices(N: int, n: int = 10_000, seed: int = SUBSET_SEED) -> np.ndarray:
    """Sorted seeded sample of n particle indices (the error-measurement subset)."""
    return np.sort(np.random.default_rng(seed).choice(N, size=min(n, N), replace=False)).astype(np.int64)


def array_digest(*arrays) -> str:
    """SHA-256 of the float64 bytes of the arrays (identifies a generated system)."""
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, np.float64).tobytes())
    return h.hexdigest()


def small_electrolyte(N: int = 1000, L: float = 12.0, min_dist: float = 0.8, seed: int = 0):
    """1:1 point charges uniform in a cubic box with pairs closer than min_dist redrawn (WCA at
    random overlap would diverge).  Returns (pos, q, L)."""
    from .md import random_packing
    pos = random_packing(N, L, min_dist, seed=seed)
    q = np.where(np.arange(N) % 2 == 0, 1.0, -1.0)
    return pos, q, L


def lattice_positions(N: int, density: float, seed: int, jitter: float = 0.1):
    """N particles on randomly chosen sites of a jittered simple-cubic lattice (cubic box of
    volume N / density, centred on the origin): no close contacts, so MD starts cleanly."""
    L = (N / density) ** (1.0 / 3.0)
    m = int(math.ceil(N ** (1.0 / 3.0)))
    a = L / m
    rng = np.random.default_rng(seed)
    sites = rng.permutation(m ** 3)[:N]
    idx = np.stack([sites % m, (sites // m) % m, sites // (m * m)], 1)
    pos = (idx + 0.5) * a - L / 2 + rng.uniform(-jitter * a, jitter * a, (N, 3))
    return pos, L


def neutral_charges(N: int, q2_frac: float, seed: int) -> np.ndarray:
    """+-1 charges with a fraction q2_frac of +-2, each valence class neutral (N even)."""
    rng = np.random.default_rng(seed + 7919)
    n2 = int(N * q2_frac) // 2 * 2
    n1 = N - n2
    q = np.concatenate([np.tile([2.0, -2.0], n2 // 2), np.tile([1.0, -1.0], n1 // 2), np.zeros(n1 % 2)])
    return rng.permutation(q)


def batch_sizes(n_sys: int = 512, seed: int = 6, lo: int = 500, hi: int = 50_000) -> np.ndarray:
    """Particle counts of the batch: log-uniform in [lo, hi], rounded down to even."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n_sys)).astype(int) // 2 * 2


def batch_systems(n_sys: int = 512, seed: int = 6, density: float = 0.1, q2_frac: float = 0.1,
                  which=None) -> List[Tuple[np.ndarray, np.ndarray, float]]:
    """[(pos, q, L)] of the batch (or of the members listed in `which`)."""
    Ns = batch_sizes(n_sys, seed)
    out = []
    for i in (range(n_sys) if which is None else which):
        pos, L = lattice_positions(int(Ns[i]), density, seed=seed * 1000 + i)
        out.append((pos, neutral_charges(int(Ns[i]), q2_frac, seed * 1000 + i), L))
    return out


def equilibrate(state, params, steps: int = 1000, dt: float = 0.002, kT: float = 1.0, gamma: float = 1.0, seed: int = 4,
                skin: float = 0.4):
    """Langevin dynamics (ESP with `params` + WCA, sigma = epsilon = 1, masses 1) for `steps`
    steps on the GPU, from Maxwell-Boltzmann velocities at kT; returns the Simulation."""
    from . import md
    from .solver import Solver
    state.thermalize(kT, seed=seed + 1)
    sim = md.Simulation(state, dt=dt, method=md.Langevin(kT=kT, gamma=gamma, seed=seed), nlist=md.NeighborList(skin))
    sim.add_force(md.ESPForce(Solver(params)))
    sim.add_force(md.WCA(1.0, 1.0))
    sim.run(steps)
    return sim


__all__ = ["lattice_electrolyte", "subset_indices", "array_digest", "small_electrolyte", "lattice_positions",
           "neutral_charges", "batch_sizes", "batch_systems", "equilibrate", "SUBSET_SEED"]
