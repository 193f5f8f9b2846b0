"""Helpers that build one ESP configuration for both the HIP solver and the CPU oracle."""
import numpy as np

from esp_hoomd_amd import _lib
from esp_hoomd_amd.solver import DeviceSystem, Params, Solver


def make_params(mesh, P, c, rc, split=0, precision=0, kappa=None, **kw):
    if kappa is None:
        kappa = c / rc
    nx, ny, nz = mesh if isinstance(mesh, (tuple, list)) else (mesh, mesh, mesh)
    return Params(nx=nx, ny=ny, nz=nz, order=P, kappa=kappa, rcut=rc, split=split, precision=precision, **kw)


def oracle_for(solver: Solver, box, params: Params):
    import oracle as O
    scr, dscr, rlo, dri = solver.split_table()
    Pw, D, w = solver.window()
    return O.OracleESP(box, (params.nx, params.ny, params.nz), params.order, w, scr, dscr, rlo, dri, params.rcut,
                       split_kind=params.split, kappa=params.kappa, k_cut=solver.info().k_cut, alpha=params.alpha)


def run_gpu(params: Params, pos, q, box, exclusions=None, group=None, skin=0.3):
    solver = Solver(params)
    L = np.asarray(box[:3], float)
    r_list = min(params.rcut + skin, 0.499 * L.min())
    sysd = DeviceSystem(pos, q, box, r_list=r_list, exclusions=exclusions, group=group)
    res = solver.compute(sysd)
    res["energy"] = res["force"][:, 3].sum() + res["ext"][0]
    res["total_virial"] = res["virial"].sum(axis=1) + res["ext"][1:7]
    return solver, sysd, res
