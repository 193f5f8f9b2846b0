"""Pins the oracle in tilted (triclinic) boxes, HOOMD's BoxDim with tilt factors xy, xz, yz
(lattice vectors a1 = (Lx, 0, 0), a2 = (xy Ly, Ly, 0), a3 = (xz Lz, yz Lz, Lz)).

The general-box Ewald sum is checked against the orthorhombic one on boxes whose tilts are
whole lattice vectors (the same periodic system written in another cell: results must agree),
against itself at two splitting parameters, and against the finite-strain derivative of its
energy, shear components included.  The mesh oracle (fractions by makeFraction, k = 2 pi B^-T m,
forces through B^-T) is then checked against that Ewald sum, against finite differences of its
energy and against finite strain."""
import numpy as np
import pytest

import oracle as O
from conftest import box_from_matrix, box_matrix, random_electrolyte, random_electrolyte_box, rel_rms
from esp_hoomd_amd import solver as S

TILTED = (7.0, 7.5, 8.0, 0.3, -0.2, 0.15)


def _esp(box, n, P, kappa, rc, split=0, alpha=0.0):
    L = np.asarray(box[:3], float)
    B = box_matrix(box)
    # coarsest axis: lattice-vector length over its mesh points
    h = max(np.linalg.norm(B[:, a]) / n for a in range(3))
    K = kappa if split == 0 else 2 * kappa * 3.7
    D, cw = S.default_window(P, np.pi / (h * K))
    w = S.build_window(P, D, cw)
    scr, dscr, rlo, dri = S.build_split_table(split, kappa, rc, 4096)
    return O.OracleESP(box, (n, n, n), P, w, scr, dscr, rlo, dri, rc, split_kind=split, kappa=kappa, alpha=alpha)


def test_ewald_box_equals_orthorhombic_ewald_without_tilt():
    pos, q = random_electrolyte(80, 8.0, seed=2)
    box = (8.0, 8.0, 8.0)
    fe, ee, we = O.ewald_auto(pos, q, box, rcut=3.9)
    fb, eb, wb = O.ewald_box_auto(pos, q, box + (0.0, 0.0, 0.0), rcut=3.9)
    assert rel_rms(fb, fe) < 1e-12
    assert abs(eb - ee) < 1e-13 * abs(ee)
    assert rel_rms(wb, we) < 1e-12


@pytest.mark.parametrize("tilt", [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, -1.0), (1.0, -1.0, 1.0)])
def test_ewald_box_whole_lattice_vector_tilts_are_the_same_system(tilt):
    """xy = 1 with Lx = Ly makes a2' = a2 + a1 (likewise for xz, yz): the same lattice, so the
    same periodic system; forces, energy and virial must equal the orthorhombic sum."""
    L = 8.0
    pos, q = random_electrolyte(70, L, seed=5)
    fe, ee, we = O.ewald_auto(pos, q, (L, L, L), rcut=3.9)
    box = (L, L, L) + tuple(tilt)
    fb, eb, wb = O.ewald_box_auto(pos, q, box)
    assert rel_rms(fb, fe) < 1e-10
    assert abs(eb - ee) < 1e-10 * abs(ee)
    assert rel_rms(wb, we) < 1e-10


def test_ewald_box_two_splitting_parameters_agree():
    pos, q = random_electrolyte_box(90, TILTED, seed=3)
    hw = O.box_half_width(TILTED)
    f1, e1, w1 = O.ewald_box_auto(pos, q, TILTED, rcut=0.99 * hw)
    f2, e2, w2 = O.ewald_box_auto(pos, q, TILTED, rcut=0.7 * hw)
    assert rel_rms(f1, f2) < 1e-11
    assert abs(e1 - e2) < 1e-11 * abs(e1)
    assert rel_rms(w1, w2) < 1e-11


def _strained(box, pos, a, b, e):
    """Box and positions after the strain r -> (I + e E_ab) r (a <= b keeps B upper triangular)."""
    Sm = np.eye(3)
    Sm[a, b] += e
    return box_from_matrix(Sm @ box_matrix(box)), pos @ Sm.T


VIRIAL_COMPONENTS = [(0, 0, 0), (0, 1, 1), (0, 2, 2), (1, 1, 3), (1, 2, 4), (2, 2, 5)]


@pytest.mark.parametrize("a,b,comp", VIRIAL_COMPONENTS)
def test_ewald_box_virial_is_finite_strain_derivative(a, b, comp):
    pos, q = random_electrolyte_box(40, TILTED, seed=8, min_dist=0.8)
    rc = 0.9 * O.box_half_width(TILTED)
    _, e0, w0 = O.ewald_box_auto(pos, q, TILTED, rcut=rc)
    eps = 1e-6
    es = []
    for sgn in (+1, -1):
        box, p = _strained(TILTED, pos, a, b, sgn * eps)
        es.append(O.ewald_box_auto(p, q, box, rcut=rc)[1])
    fd = -(es[0] - es[1]) / (2 * eps)
    assert abs(fd - w0[comp]) < 1e-7 * abs(e0)


@pytest.mark.parametrize("box", [TILTED, (8.0, 8.0, 8.0, 1.0, 0.0, 0.0), (7.0, 8.0, 9.0, -0.4, 0.25, 0.5)])
def test_oracle_gaussian_split_in_tilted_box_reduces_to_ewald(box):
    pos, q = random_electrolyte_box(100, box, seed=11)
    hw = O.box_half_width(box)
    fe, ee, we = O.ewald_box_auto(pos, q, box)
    rc = 0.95 * hw
    o = _esp(box, 64, 10, 4.2 / rc, rc, split=1)
    r = o.compute(pos, q)
    assert rel_rms(r["force"][:, :3], fe) < 2e-6
    assert abs(r["energy"] - ee) < 1e-6 * abs(ee)
    assert rel_rms(r["total_virial"], we) < 2e-6


def test_oracle_pswf_in_tilted_box_against_ewald():
    pos, q = random_electrolyte_box(120, TILTED, seed=12)
    fe, ee, we = O.ewald_box_auto(pos, q, TILTED)
    rc = 3.0
    r = _esp(TILTED, 48, 10, 14.0 / rc, rc).compute(pos, q)
    err = rel_rms(r["force"][:, :3], fe)
    assert err < 3e-6
    assert abs(r["energy"] - ee) < 3e-6 * abs(ee)
    assert rel_rms(r["total_virial"], we) < 3e-5


def test_oracle_forces_in_tilted_box_are_energy_gradient():
    box = (6.0, 6.5, 7.0, 0.35, -0.25, 0.2)
    pos, q = random_electrolyte_box(24, box, seed=14, min_dist=0.7)
    o = _esp(box, 16, 6, 10 / 2.4, 2.4)
    r = o.compute(pos, q)
    eps = 1e-5
    for j in (0, 7, 19):
        for a in range(3):
            p1 = pos.copy()
            p1[j, a] += eps
            p2 = pos.copy()
            p2[j, a] -= eps
            fd = -(o.compute(p1, q)["energy"] - o.compute(p2, q)["energy"]) / (2 * eps)
            assert abs(fd - r["force"][j, a]) < 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("a,b,comp", VIRIAL_COMPONENTS)
def test_oracle_virial_in_tilted_box_is_finite_strain_derivative(a, b, comp):
    box0 = (6.0, 6.5, 7.0, 0.35, -0.25, 0.2)
    pos, q = random_electrolyte_box(30, box0, seed=15, min_dist=0.7)
    base = _esp(box0, 16, 6, 10 / 2.4, 2.4)
    r = base.compute(pos, q)
    eps = 1e-6
    es = []
    for sgn in (+1, -1):
        box, p = _strained(box0, pos, a, b, sgn * eps)
        o = O.OracleESP(box, (16, 16, 16), 6, base.wcoef, base.screen, base.dscreen, base.r_lo, base.dr_inv, 2.4)
        es.append(o.compute(p, q)["energy"])
    fd = -(es[0] - es[1]) / (2 * eps)
    assert abs(fd - r["total_virial"][comp]) < 1e-6 * abs(r["energy"])


def test_oracle_refuses_cutoff_beyond_half_the_tilted_width():
    box = (6.0, 6.0, 6.0, 1.0, 0.0, 0.0)
    pos, q = random_electrolyte_box(10, box, seed=1)
    with pytest.raises(ValueError):
        _esp(box, 16, 6, 4.0, 0.6 * 6.0).compute(pos, q)
