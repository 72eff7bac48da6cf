"""The synthetic workloads the solver is measured on (seeded, reproducible).

  vortex_p1_64       64x64 periodic split quads on [0,10]^2, one diagonal (8192 triangles),
                     isentropic vortex (gamma 1.4, beta 5, free stream rho=u=v=p=1), P1, LF,
                     CFL 0.3, reference dt subset, 100 steps
  vortex_p2_512      512x512 jittered split quads (seed-chosen diagonals, +-0.2h), 524,288
                     triangles, P2, LF, 200 steps
  riemann_p1_1024    four-quadrant Riemann problem on [0,1]^2, 1024x1024 jittered split quads,
                     transmissive sides, P1, Roe, TVB M=50 (+ positivity), to t = 0.3
  batch_256          256 periodic vortex instances, N ~ log-uniform in [32, 256] split quads per
                     side, beta ~ U[1,5], Mach ~ U[0.1,0.8], angle ~ U[0, 2pi), P2, LF, 100 steps
  vortex_p3_2048     2048x2048 jittered split quads (8.4M triangles), P3, Roe, 50 steps
"""
import math

import numpy as np

from .solver import Grid, riemann4_params, vortex_params
from ._lib import BC_OUTFLOW

GAMMA = 1.4


def vortex_grid(n, jitter=0.2, seed=2024, diagonal=1):
    return Grid.splitquad(n, n, 0.0, 0.0, 10.0, 10.0, periodic=True, diagonal=diagonal, jitter=jitter, seed=seed)


def vortex_p1_64():
    g = Grid.splitquad(64, 64, 0.0, 0.0, 10.0, 10.0, periodic=True, diagonal=0, jitter=0.0)
    return g, dict(degree=1, flux="lf", dt_policy="reference", cfl=0.3), vortex_params(), 100


def vortex_p2(n=512):
    return vortex_grid(n), dict(degree=2, flux="lf", dt_policy="max", cfl=0.3), vortex_params(), 200


def riemann_p1(n=1024):
    g = Grid.splitquad(n, n, 0.0, 0.0, 1.0, 1.0, periodic=False, diagonal=1, jitter=0.2, seed=2025, bc=BC_OUTFLOW)
    return g, dict(degree=1, flux="roe", limiter=True, tvb_M=50.0, dt_policy="max", cfl=0.2), riemann4_params(), 0.3


def vortex_p3(n=2048):
    return vortex_grid(n, seed=2026), dict(degree=3, flux="roe", dt_policy="max", cfl=0.2), vortex_params(), 50


def batch_instances(count=256, seed=7, nmin=32, nmax=256):
    """Per-instance (n, vortex parameters, cfl) of the batched sweep."""
    rng = np.random.default_rng(seed)
    out = []
    c_inf = math.sqrt(GAMMA)
    for _ in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        beta = rng.uniform(1.0, 5.0)
        mach = rng.uniform(0.1, 0.8)
        th = rng.uniform(0.0, 2 * math.pi)
        u, v = mach * c_inf * math.cos(th), mach * c_inf * math.sin(th)
        out.append((n, vortex_params(beta=beta, uinf=u, vinf=v), 0.3))
    return out
