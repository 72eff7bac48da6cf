"""Shared helpers for the parity tests."""
import numpy as np

from oracle import pyoracle as po
from rkdg_heter_calculate_amd import Grid, Solver, vortex_params

FLUX_ID = {"lf": 0, "roe": 1}
DT_ID = {"reference": 0, "max": 1, "fixed": 2}


def rel_l2(Ua, Ub, mass):
    """Relative L2 difference per conserved variable (orthogonal element basis)."""
    return po.l2_norms(Ua - Ub, mass) / po.l2_norms(Ub, mass)


def vortex_grid(n, jitter=0.0, diagonal=0, seed=0):
    return Grid.splitquad(n, n, 0.0, 0.0, 10.0, 10.0, periodic=True, diagonal=diagonal, jitter=jitter, seed=seed)


def make_pair(g, P=2, flux="lf", limiter=False, dt_policy="max", cfl=0.3, tvb_M=50.0, freestream=None,
              order=None, dt_fixed=0.0, reduction_threads=512):
    ex, ey = g.element_coords()
    fs = freestream if freestream is not None else (1.0, 0.0, 0.0, 2.5)
    o = po.Oracle(ex, ey, g.neighbour, g.shared_edge, g.boundary, P=P, flux=FLUX_ID[flux], limiter=int(limiter),
                  dt_policy=DT_ID[dt_policy], cfl=cfl, tvb_M=tvb_M, freestream=fs, dt_fixed=dt_fixed,
                  reduction_threads=reduction_threads)
    s = Solver(g, order=order, degree=P, flux=flux, limiter=limiter, dt_policy=dt_policy, cfl=cfl, tvb_M=tvb_M,
               freestream=list(fs), dt_fixed=dt_fixed, reduction_threads=reduction_threads)
    return o, s


def vortex_ic(o):
    o.project(1, vortex_params())
    return o.get_state()
