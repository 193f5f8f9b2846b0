"""The benchmark workloads themselves, on the GPU, against the oracle and the direct Ewald sum.

* Small electrolyte (N = 1,000, L = 12, mesh 32^3, P = 8, r_c = 3.0, FP64): forces, energy and
  virial equal the oracle's to 1e-10; 100 NVE steps (dt = 0.002, WCA + ESP) conserve energy.
* 1,000,000-particle lattice electrolyte (density 0.1): the reference GPU configuration
  (128^3, P = 10, r_c = 3.5, FP64), the two mixed-precision accuracy settings (96^3 / P = 8 at
  1e-4, 160^3 / P = 12 at 1e-6) and the autotuned 1e-5 configurations (FP64 and mixed),
  compared with FP64 direct-Ewald forces on a seeded 10,000-particle subset (committed
  fixture), and again after 1,000 Langevin equilibration steps (Ewald evaluated here on the
  host cores for a 2,000-particle subset).
* A 16-system sample of the 512-system batch (500-50,000 particles, +-1 and +-2) in one batched
  launch: mixed-precision members against Ewald at 1e-5, FP64 members against the oracle at
  1e-10.
"""
import math
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from conftest import rel_rms
from esp_cases import oracle_for
from esp_hoomd_amd import _lib, accuracy, md, tune, workloads as W
from esp_hoomd_amd.batch import Batch
from esp_hoomd_amd.solver import Params, Solver

pytestmark = pytest.mark.gpu

FIX = Path(__file__).resolve().parent / "golden" / "ewald_subset_N1M.npz"
FP64, MIXED = _lib.ESP_PREC_FP64, _lib.ESP_PREC_MIXED


def _small_params(precision=FP64):
    # mesh 32^3 at L = 12 (h = 0.375), P = 8, r_c = 3.0; the prolate bandwidth c = kappa r_c that
    # minimises the accuracy model's error at this spacing and cutoff (1.9e-6, the closest the
    # configuration comes to its 1e-6 target): c = 13.25
    return Params(nx=32, ny=32, nz=32, order=8, kappa=13.25 / 3.0, rcut=3.0, precision=precision)


def _forces(params, st, skin=0.4):
    nl = md.NeighborList(buffer=skin)
    nl.update(st, params.rcut, force_rebuild=True)
    solver = Solver(params)
    solver.enqueue_state(st)
    _lib.check(_lib.load().esp_device_synchronize())
    return st.force.numpy()


def test_small_electrolyte_fp64_matches_oracle_and_ewald():
    pos, q, L = W.small_electrolyte()
    params = _small_params()
    st = md.State(pos, q, (L, L, L))
    nl = md.NeighborList(buffer=0.4)
    nl.update(st, params.rcut, force_rebuild=True)
    solver = Solver(params)
    solver.enqueue_state(st)
    f = st.force.numpy()
    vir = st.virial.numpy()[:, : st.N]
    ext = st.ext.numpy()[:7]
    ref = oracle_for(solver, (L, L, L), params).compute(pos, q)
    assert rel_rms(f[:, :3], ref["force"][:, :3]) < 1e-10
    assert rel_rms(f[:, 3], ref["force"][:, 3]) < 1e-10
    assert rel_rms(vir, ref["virial"]) < 1e-10
    e = f[:, 3].sum() + ext[0]
    assert abs(e - ref["energy"]) < 1e-10 * abs(ref["energy"])
    assert rel_rms(vir.sum(1) + ext[1:7], ref["total_virial"]) < 1e-10
    fe, ee, _ = O.ewald_auto(pos, q, (L, L, L), tol=1e-15)
    err = rel_rms(f[:, :3], fe)
    print(f"small electrolyte: force error vs Ewald {err:.3e}, energy {e:.8f} vs Ewald {ee:.8f}", flush=True)
    # measured 5.8e-6 (the accuracy model, fitted at density 0.1, predicts 1.9e-6 at this density 0.58)
    assert err < 1e-5
    assert abs(e - ee) < 1e-4 * abs(ee)


def _nve_drift(with_esp):
    pos, q, L = W.small_electrolyte()
    st = md.State(pos, q, (L, L, L))
    st.thermalize(1.0, seed=1)
    sim = md.Simulation(st, dt=0.002, nlist=md.NeighborList(0.4))
    if with_esp:
        sim.add_force(md.ESPForce(Solver(_small_params())))
    sim.add_force(md.WCA(1.0, 1.0))
    sim.compute_forces()
    ke0, pe0 = sim.energies()
    es = []
    for _ in range(10):
        sim.run(10)
        ke, pe = sim.energies()
        es.append(ke + pe)
    assert sim.timestep == 100
    return max(abs(x - (ke0 + pe0)) for x in es) / ke0, ke0 + pe0, sim.nlist.builds


def test_small_electrolyte_nve_100_steps():
    """100 velocity-Verlet steps at dt = 0.002 from T = 1 (seed 1), ESP + WCA.  The random start
    (pairs only kept apart to 0.8 sigma) holds a large WCA energy, whose integration error at
    this time step dominates the energy drift: measured max |dE| / KE0 = 1.35e-2 with ESP, and
    the same run with WCA alone drifts as much; the ESP forces add less than 1e-3 of KE0."""
    d_esp, e0, builds = _nve_drift(True)
    d_wca, _, _ = _nve_drift(False)
    print(f"NVE 100 steps: E0 {e0:.6f}, max |dE| / KE0 = {d_esp:.3e} (WCA alone {d_wca:.3e}), list builds {builds}",
          flush=True)
    assert d_esp < 3e-2
    assert abs(d_esp - d_wca) < 1e-3


# ---- the 1M-particle electrolyte -----------------------------------------------------------
L1M = (1_000_000 / 0.1) ** (1.0 / 3.0)


def _mesh_params(n, P, target, precision, model_precision=None):
    """Mesh n^3 and order P fixed; (c, sigma) with the smallest cutoff that meets `target` by the
    accuracy model (evaluated without its FP32 term when model_precision is FP64)."""
    c, sigma = tune.best_split(P, target, precision if model_precision is None else model_precision)
    h = L1M / n
    rc = sigma * c * h / math.pi
    return Params(nx=n, ny=n, nz=n, order=P, kappa=c / rc, rcut=rc, precision=precision)


def _settings():
    return {
        # the reference's GPU configuration (oversampling 1.3 of the band limit)
        "literal_fp64_128_P10_rc3.5": Params(nx=128, ny=128, nz=128, order=10, kappa=math.pi / (1.3 * L1M / 128), rcut=3.5,
                                             precision=FP64),
        "mixed_96_P8_1e-4": _mesh_params(96, 8, 1e-4, MIXED),
        "mixed_160_P12_1e-6": _mesh_params(160, 12, 1e-6, MIXED, model_precision=FP64),
        # the autotuner's 1e-5 picks for this lattice (measured against the tight FP64 reference)
        "tuned_fp64_180_P10": Params(nx=180, ny=180, nz=180, order=10, kappa=2.2824, rcut=5.6958, precision=FP64),
        "tuned_mixed_144_P10": Params(nx=144, ny=144, nz=144, order=10, kappa=1.81216, rcut=7.1048, precision=MIXED),
    }


def test_lattice_1M_against_ewald_fixture():
    d = np.load(FIX)
    pos, q, L = W.lattice_electrolyte(1_000_000, 0.1)
    assert str(d["digest"]) == W.array_digest(pos, q), "the generator no longer reproduces the fixture's system"
    idx = d["index"]
    ref = d["force"]
    st = md.State(pos, q, (L, L, L))
    errs = {}
    for name, p in _settings().items():
        f = _forces(p, st)
        errs[name] = rel_rms(f[idx, :3], ref)
        print(f"1M lattice, {name} (mesh {p.nx}, P {p.order}, r_c {p.rcut:.3f}, c {p.kappa * p.rcut:.3f}): "
              f"error vs Ewald {errs[name]:.3e}", flush=True)
    tight = accuracy.reference_forces(pos, q, (L, L, L))
    errs["tight_reference"] = rel_rms(tight[idx], ref)
    print(f"1M lattice, tight FP64 reference: error vs Ewald {errs['tight_reference']:.3e}", flush=True)
    assert errs["tight_reference"] < 1e-7
    # r_c / h = 2.1 caps the reference configuration's accuracy near 1e-2
    assert 1e-3 < errs["literal_fp64_128_P10_rc3.5"] < 5e-2
    assert errs["tuned_fp64_180_P10"] <= 1e-5
    assert errs["tuned_mixed_144_P10"] <= 1e-5
    # the lattice's small RMS force makes relative errors ~5x the model's random electrolytes
    assert errs["mixed_96_P8_1e-4"] < 1e-3
    assert errs["mixed_160_P12_1e-6"] < 1e-5


def test_lattice_1M_equilibrated_against_ewald():
    """1,000 Langevin steps (T = 1, gamma = 1, seed 4, dt = 0.002) with the reference GPU
    configuration + WCA, then the settings against the direct Ewald sum on 2,000 particles.

    The accuracy model is fitted on uniformly random electrolytes, whose close pairs carry large
    forces; the equilibrated WCA electrolyte has none, so its relative error at the model's
    parameters is ~3x the prediction.  The autotuner measures that on the system (against the
    tight FP64 reference) and rescales: its picks meet 1e-5 against Ewald, and the same
    measured scale gives the two fixed-mesh accuracy settings their cutoffs (96^3 / P = 8 meets
    1e-4; at 160^3 / P = 12 the FP64 method error meets 1e-6 and the mixed result is reported
    with its FP32 floor)."""
    pos, q, L = W.lattice_electrolyte(1_000_000, 0.1)
    box = (L, L, L)
    st = md.State(pos, q, box)
    sets = _settings()
    sim = W.equilibrate(st, sets["literal_fp64_128_P10_rc3.5"], steps=1000)
    print(f"equilibrated: T = {sim.temperature():.3f}, list builds {sim.nlist.builds}", flush=True)
    peq = st.positions().copy()
    del sim, st
    idx = W.subset_indices(len(peq), 2000)
    rc = 70.0
    fe, _, _ = O.ewald(peq, q, box, 5.0 / rc, rc, 50.0 / rc, targets=idx, energy=False)
    tight = accuracy.reference_forces(peq, q, box)
    errs = {"tight_reference": rel_rms(tight[idx], fe)}
    print(f"1M equilibrated, tight FP64 reference: error vs Ewald {errs['tight_reference']:.3e}", flush=True)
    assert errs["tight_reference"] < 1e-7
    runs = {"literal_fp64_128_P10_rc3.5": sets["literal_fp64_128_P10_rc3.5"],
            "model_mixed_1e-5": tune.tune(len(peq), box, 1e-5, MIXED)[0].params}
    scale = 1.0
    for name, prec in (("autotuned_fp64_1e-5", FP64), ("autotuned_mixed_1e-5", MIXED)):
        ch = tune.autotune(peq, q, box, 1e-5, prec, reference=tight)
        runs[name] = ch.params
        scale = max(scale, ch.method_scale)
    for name, n, P, target in (("96_P8_1e-4", 96, 8, 1e-4), ("160_P12_1e-6", 160, 12, 1e-6)):
        c, sigma = tune.best_split(P, target / 1.1, FP64, scale=scale)
        h = L / n
        r = sigma * c * h / math.pi
        for prec, tag in ((MIXED, "mixed"), (FP64, "fp64")):
            runs[f"{tag}_{name}"] = Params(nx=n, ny=n, nz=n, order=P, kappa=c / r, rcut=r, precision=prec)
    for name, p in runs.items():
        f = _forces(p, md.State(peq, q, box))
        errs[name] = rel_rms(f[idx, :3], fe)
        print(f"1M equilibrated, {name} (mesh {p.nx}, P {p.order}, r_c {p.rcut:.3f}, c {p.kappa * p.rcut:.3f}): "
              f"error vs Ewald {errs[name]:.3e}", flush=True)
    print(f"1M equilibrated: measured method scale {scale:.3f}", flush=True)
    assert 1e-3 < errs["literal_fp64_128_P10_rc3.5"] < 5e-2
    assert errs["autotuned_fp64_1e-5"] <= 1e-5
    assert errs["autotuned_mixed_1e-5"] <= 1e-5
    assert errs["mixed_96_P8_1e-4"] <= 1e-4
    assert errs["fp64_160_P12_1e-6"] <= 1e-6


# ---- a sample of the 512-system batch ----------------------------------------------------
def _batch_sample(n=16):
    Ns = W.batch_sizes(512, 6)
    order = np.argsort(Ns)
    which = [int(order[int(round(k * (len(order) - 1) / (n - 1)))]) for k in range(n)]
    return which, W.batch_systems(512, 6, which=which)


def test_batch_sample_mixed_against_ewald_and_fp64_against_oracle():
    which, systems = _batch_sample()
    # the model's method scale for these lattice-started systems, measured on three of them
    Ns = [len(s[0]) for s in systems]
    scale = 1.0
    for k in (0, len(systems) // 2, len(systems) - 1):
        pos, q, L = systems[k]
        ch = tune.autotune(pos, q, (L, L, L), 1e-5, MIXED, orders=(6, 8, 10), mesh_range=(16, 64),
                           cost=tune.CostModel(rebuild_every=0.0))
        scale = max(scale, ch.method_scale)
    chosen = [tune.tune(len(pos), (L, L, L), 1e-5, MIXED, orders=(6, 8, 10), mesh_range=(16, 64), method_scale=scale,
                        safety=1.1, batched=True)[0] for pos, q, L in systems]
    for precision in (MIXED, FP64):
        pairs = []
        for (pos, q, L), ch in zip(systems, chosen):
            p = Params(**{**ch.params.__dict__, "precision": precision, "tile": tune.batch_tile(ch.params.order, precision)})
            st = md.State(pos, q, (L, L, L))
            nl = md.NeighborList(buffer=0.4)
            nl.update(st, p.rcut, force_rebuild=True)
            st._nl = nl
            pairs.append((Solver(p), st))
        Batch(use_graph=True).enqueue(pairs)
        _lib.check(_lib.load().esp_device_synchronize())
        for k, ((pos, q, L), (solver, st)) in enumerate(zip(systems, pairs)):
            f = st.force.numpy()
            if precision == MIXED:
                idx = np.arange(len(pos)) if len(pos) <= 400 else W.subset_indices(len(pos), 400, seed=which[k])
                fe, _, _ = O.ewald_auto(pos, q, (L, L, L), tol=1e-15, targets=idx, energy=False)
                err = rel_rms(f[idx, :3], fe)
                print(f"batch member {which[k]} N={Ns[k]} mesh {solver.params.nx} P {solver.params.order} "
                      f"r_c {solver.params.rcut:.3f}: mixed error vs Ewald {err:.3e}", flush=True)
                assert err <= 1e-5
            else:
                ref = oracle_for(solver, (L, L, L), solver.params).compute(pos, q)
                assert rel_rms(f[:, :3], ref["force"][:, :3]) < 1e-10
                assert rel_rms(f[:, 3], ref["force"][:, 3]) < 1e-10
                assert rel_rms(st.virial.numpy()[:, : st.N], ref["virial"]) < 1e-10
    assert min(Ns) < 700 and max(Ns) > 40_000
