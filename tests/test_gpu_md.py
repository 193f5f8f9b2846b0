"""GPU MD support: neighbour list against a k-d tree, WCA against numpy, the ESP objects in a
simulation against the oracle, NVE energy conservation and the Langevin thermostat."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from conftest import random_electrolyte, rel_rms
from esp_hoomd_amd import esp, md

pytestmark = pytest.mark.gpu


def _kdtree_pairs(pos, L, r):
    t = cKDTree((pos + L / 2) % L, boxsize=L)
    return {tuple(p) for p in t.query_pairs(r, output_type="ndarray")}


@pytest.mark.parametrize("N,L,r_cut,buf,inner", [(500, 10.0, 2.5, 0.3, 1.2), (300, 7.0, 3.0, 0.4, 0.0), (64, 6.5, 1.2, 0.0, 0.5),
                                                 (2000, 10.0, 2.5, 0.4, 1.5)])
def test_nlist_matches_kdtree(N, L, r_cut, buf, inner):
    pos, q = random_electrolyte(N, L, seed=N)
    excl = [(0, 1), (5, 9)]
    st = md.State(pos, q, (L, L, L), exclusions=excl)
    nl = md.NeighborList(buffer=buf, inner=inner)
    nl.update(st, r_cut, force_rebuild=True)
    n, lists = nl.host_arrays(st)
    got = {(min(i, int(j)), max(i, int(j))) for i in range(N) for j in lists[i]}
    exp = _kdtree_pairs(pos, L, r_cut + buf) - {(0, 1), (5, 9)}
    assert got == exp
    assert all(i not in set(lists[i].tolist()) for i in range(N))
    # symmetric full list
    assert sum(len(x) for x in lists) == 2 * len(exp)
    # candidates are listed cell by cell, each cell in ascending index (not in the order the
    # binning atomics happened to run): at most one ascending run per neighbour cell (27 cells
    # of at least r_list, or 125 of at least r_list / 2 when every axis has five of those), and
    # a second list built from the same positions is identical
    half = min(int(L // ((r_cut + buf) / 2)) for _ in range(3)) >= 5
    assert max(1 + int(np.sum(np.diff(x.astype(np.int64)) < 0)) for x in lists if len(x)) <= (125 if half else 27)
    nl2 = md.NeighborList(buffer=buf, inner=inner)
    st2 = md.State(pos, q, (L, L, L), exclusions=excl)
    nl2.update(st2, r_cut, force_rebuild=True)
    n2, lists2 = nl2.host_arrays(st2)
    assert np.array_equal(n, n2) and all(np.array_equal(a, b) for a, b in zip(lists, lists2))
    if inner == 0.0:
        assert st.inner_ptrs is None
        return
    # the inner list: the same pass, the pairs within inner + buffer
    n_in, lists_in = nl.host_arrays(st, inner=True)
    got_in = {(min(i, int(j)), max(i, int(j))) for i in range(N) for j in lists_in[i]}
    assert got_in == _kdtree_pairs(pos, L, inner + buf) - {(0, 1), (5, 9)}
    assert sum(len(x) for x in lists_in) == 2 * len(got_in)


def test_wca_over_inner_list_equals_full_list():
    """WCA summed over the inner list (single-system and batched launches) against the same
    force over the full list; only the summation order differs."""
    L, N = 12.0, 800
    pos = md.random_packing(N, L, 0.85, seed=5)
    out = {}
    for inner in (0.0, 1.2):
        st = md.State(pos, np.zeros(N), (L, L, L))
        nl = md.NeighborList(0.3, inner=inner)
        nl.update(st, 3.0, force_rebuild=True)
        assert (st.inner_ptrs is not None) == (inner > 0)
        st.force.zero()
        st.virial.zero()
        md.WCA(1.0, 1.0).add_to(st)
        single = (st.force.numpy().copy(), st.virial.numpy().copy())
        st.force.zero()
        st.virial.zero()
        b = md.BatchMD()
        b.set([(st, nl)])
        b.wca(1.0, 1.0)
        out[inner] = (single, (st.force.numpy().copy(), st.virial.numpy().copy()))
    full = out[0.0][0]
    assert np.abs(full[0][:, :3]).max() > 1.0          # the packing has overlapping WCA shells
    for f, v in out[1.2]:
        assert rel_rms(f, full[0]) < 1e-13
        assert rel_rms(v, full[1]) < 1e-13
    assert rel_rms(out[0.0][1][0], full[0]) < 1e-13


def test_wca_matches_numpy():
    L = 8.0
    pos = md.random_packing(200, L, 0.85, seed=3)
    st = md.State(pos, np.zeros(200), (L, L, L))
    sim = md.Simulation(st, dt=0.001)
    w = md.WCA(1.0, 1.0)
    sim.add_force(w)
    sim.compute_forces()
    f = st.force.numpy()
    ref = np.zeros((200, 4))
    rc2 = 2 ** (1 / 3)
    for i in range(200):
        d = pos[i] - pos
        d -= L * np.round(d / L)
        r2 = (d * d).sum(1)
        r2[i] = np.inf
        m = r2 < rc2
        sr6 = 1 / r2[m] ** 3
        fd = 24 * sr6 * (2 * sr6 - 1) / r2[m]
        ref[i, :3] = (fd[:, None] * d[m]).sum(0)
        ref[i, 3] = 0.5 * (4 * sr6 * (sr6 - 1) + 1).sum()
    assert rel_rms(f, ref) < 1e-12


def test_esp_objects_in_simulation_match_oracle():
    import oracle as O
    N, L = 200, 9.0
    pos, q = random_electrolyte(N, L, seed=21)
    st = md.State(pos, q, (L, L, L))
    sim = md.Simulation(st, dt=0.002, nlist=md.NeighborList(0.3))
    spectral, local = esp.make_esp_forces(sim.nlist, r_cut=3.0, resolution=(36, 36, 36), order=8, kappa=4.0)
    sim.add_force(spectral)
    sim.add_force(local)
    sim.compute_forces()
    f = st.force.numpy()
    solver = spectral._cpp_obj
    scr, dscr, rlo, dri = solver.split_table()
    _, _, w = solver.window()
    ref = O.OracleESP((L, L, L), (36, 36, 36), 8, w, scr, dscr, rlo, dri, 3.0, k_cut=solver.info().k_cut).compute(pos, q)
    assert rel_rms(f[:, :3], ref["force"][:, :3]) < 1e-10
    ke, pe = sim.energies()
    assert abs(pe - ref["energy"]) < 1e-10 * abs(ref["energy"])
    assert abs(spectral.q_sum) < 1e-12 and abs(spectral.q2_sum - N) < 1e-9
    # pair part alone when only Local is attached
    st2 = md.State(pos, q, (L, L, L))
    sim2 = md.Simulation(st2, dt=0.002, nlist=md.NeighborList(0.3))
    sp2, lo2 = esp.make_esp_forces(sim2.nlist, r_cut=3.0, resolution=(36, 36, 36), order=8, kappa=4.0)
    sp2._attach(sim2)  # solver needed by Local, but Spectral's mesh part not in the force list
    sim2.add_force(lo2)
    sim2.compute_forces()
    f2 = st2.force.numpy()
    # pair-only forces = oracle total minus its mesh part: check via energy per particle
    assert rel_rms(f2[:, 3], ref["force"][:, 3]) < 1e-10


def test_nve_energy_conservation_esp_plus_wca():
    N, L = 512, 12.0 * (512 / 1000) ** (1 / 3)
    pos = md.random_packing(N, L, 0.9, seed=0)
    q = np.where(np.arange(N) % 2 == 0, 1.0, -1.0)
    st = md.State(pos, q, (L, L, L))
    st.thermalize(1.0, seed=1)
    sim = md.Simulation(st, dt=0.002, nlist=md.NeighborList(0.4))
    spectral, local = esp.make_esp_forces(sim.nlist, r_cut=2.8, resolution=(32, 32, 32), order=8, kappa=14.0 / 2.8)
    for f in (spectral, local, md.WCA(1.0, 1.0)):
        sim.add_force(f)
    sim.compute_forces()
    ke0, pe0 = sim.energies()
    e0 = ke0 + pe0
    es = []
    for _ in range(10):
        sim.run(10)
        ke, pe = sim.energies()
        es.append(ke + pe)
    drift = max(abs(e - e0) for e in es) / ke0
    print(f"NVE: E0={e0:.6f} max |dE|/KE0 = {drift:.2e}, nlist builds {sim.nlist.builds}", flush=True)
    assert drift < 2e-3


def test_langevin_thermalises():
    N, L = 343, 10.0
    pos, _ = md.jittered_lattice(N, N / L ** 3, seed=2)
    q = np.zeros(N)
    st = md.State(pos, q, (L, L, L))
    sim = md.Simulation(st, dt=0.005, method=md.Langevin(kT=1.5, gamma=2.0, seed=4), nlist=md.NeighborList(0.4))
    sim.add_force(md.WCA())
    sim.run(400)
    temps = []
    for _ in range(20):
        sim.run(20)
        temps.append(sim.temperature())
    assert abs(np.mean(temps) - 1.5) < 0.15


def test_rebuild_exactly_when_a_particle_moves_half_the_skin():
    """esp_nlist_update and the batched displacement check rebuild a list when (and only when) a
    particle has moved at least buffer / 2 since the last build, by minimum image (a move across
    the periodic boundary counts by its short distance)."""
    N, L, r_cut, buf = 400, 9.0, 2.0, 0.4

    def wrap(p, box):
        # positions handed to the library lie in the box [-L/2, L/2), as the integrator keeps them
        return p - box * np.round(p / box)

    pos, q = random_electrolyte(N, L, seed=21)
    pos[7] = (L / 2 - 0.05, 0.0, 0.0)  # next to the +x face
    st = md.State(pos, q, (L, L, L))
    nl = md.NeighborList(buffer=buf)
    nl.update(st, r_cut, force_rebuild=True)
    assert nl.builds == 1
    moved = pos.copy()
    moved[7, 0] += 0.15  # crosses the +x face: -L/2 + 0.1 after wrapping, 0.15 by minimum image
    moved = wrap(moved, L)
    assert moved[7, 0] < 0
    st.set_positions(moved)
    assert not nl.update(st, r_cut) and nl.builds == 1
    moved[3, 1] += 0.19
    moved = wrap(moved, L)
    st.set_positions(moved)
    assert not nl.update(st, r_cut) and nl.builds == 1
    moved[3, 1] += 0.02  # 0.21 >= buffer / 2
    moved = wrap(moved, L)
    st.set_positions(moved)
    assert nl.update(st, r_cut) and nl.builds == 2
    n, lists = nl.host_arrays(st)
    got = {(min(i, int(j)), max(i, int(j))) for i in range(N) for j in lists[i]}
    assert got == _kdtree_pairs(moved, L, r_cut + buf)

    # the batched check: only the system that moved far enough is rebuilt
    members = []
    for k in range(3):
        p_k, q_k = random_electrolyte(300, 8.0, seed=40 + k)
        s_k = md.State(p_k, q_k, (8.0, 8.0, 8.0))
        n_k = md.NeighborList(buffer=buf, inner=1.12)
        n_k.update(s_k, r_cut, force_rebuild=True)
        members.append((s_k, n_k, p_k))
    b = md.BatchMD()
    b.set([(s_k, n_k) for s_k, n_k, _ in members])
    assert b.update_nlists() == []
    p1 = members[1][2].copy()
    p1[5, 2] += 0.25
    members[1][0].set_positions(wrap(p1, 8.0))
    p2 = members[2][2].copy()
    p2[9, 0] += 0.1
    members[2][0].set_positions(wrap(p2, 8.0))
    assert b.update_nlists() == [1]
    assert [m[1].builds for m in members] == [1, 2, 1]


def _esp_wca_sim(st, sort_every):
    sim = md.Simulation(st, dt=0.002, nlist=md.NeighborList(0.4), sort_every=sort_every)
    spectral, local = esp.make_esp_forces(sim.nlist, r_cut=2.8, resolution=(32, 32, 32), order=8, kappa=14.0 / 2.8)
    for f in (spectral, local, md.WCA(1.0, 1.0)):
        sim.add_force(f)
    return sim


def test_sorted_particles_follow_the_same_trajectory():
    """Re-sorting the particles into cell order (State.permute with the list's cell order, every
    10 steps) leaves the trajectory the one of the unsorted run, up to the order of the pair
    sums: positions (in the input order) agree to 1e-9 after 40 steps, the energies to 1e-9."""
    N, L = 512, 12.0 * (512 / 1000) ** (1 / 3)
    pos = md.random_packing(N, L, 0.9, seed=3)
    q = np.where(np.arange(N) % 2 == 0, 1.0, -1.0)
    runs = []
    for sort_every in (0, 10):
        st = md.State(pos, q, (L, L, L))
        st.thermalize(1.0, seed=5)
        sim = _esp_wca_sim(st, sort_every)
        sim.run(40)
        runs.append((sim, st))
    (a, sa), (b, sb) = runs
    print(f"sorts {b.sorts}, list builds {a.nlist.builds} / {b.nlist.builds}", flush=True)
    assert a.sorts == 0 and b.sorts == 4
    tag = sb.tag.numpy()
    assert not np.array_equal(tag, np.arange(N))  # the particles were actually reordered
    assert np.array_equal(np.sort(tag), np.arange(N))
    pa, pb = sa.positions(), sb.in_input_order(sb.positions())
    d = (pa - pb + L / 2) % L - L / 2
    assert np.abs(d).max() < 1e-9
    va = sa.vel.numpy()[:, :3]
    vb = sb.in_input_order(sb.vel.numpy()[:, :3])
    assert np.abs(va - vb).max() < 1e-9
    ea, eb = sum(a.energies()), sum(b.energies())
    assert abs(ea - eb) <= 1e-9 * abs(ea)


def test_permute_refuses_exclusions():
    st = md.State(np.zeros((4, 3)) + np.arange(4)[:, None], np.array([1.0, -1.0, 1.0, -1.0]), (10.0, 10.0, 10.0),
                  exclusions=[(0, 1)])
    with pytest.raises(ValueError):
        md.Simulation(st, dt=0.002, sort_every=5)
