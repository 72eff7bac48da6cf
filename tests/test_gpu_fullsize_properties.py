"""Size-independent properties at the workloads' full sizes, where the oracle would take too
long for many steps: no oracle, the GPU state against itself.

* conservation of the periodic totals sum_e A_e U_e (mode 0 is the cell mean) over 20 steps
  of the P2 vortex on the 1024x1024 jittered mesh (2,097,152 triangles, 8,192 staged tiles),
  in every form;
* free-stream preservation over 5 steps on the P3 Roe vortex mesh family at 1024x1024
  (2,097,152 triangles), in the library's own basis, in every form.

Each edge's two contributions cancel to the rounding of its form, so the drift of a total
is a sum of those roundings over about three million edges; the bounds are about ten times
the measured drift (FP64 1.7e-15, FP32 8.2e-14, BF16 6.9e-9). Free stream: FP64 at most
2.7e-16 per mode, the P3 precision forms exactly 0.
"""
import numpy as np
import pytest

from rkdg_heter_calculate_amd import IC_VORTEX, Solver, vortex_params
from rkdg_heter_calculate_amd._lib import BASIS_NATIVE
from rkdg_heter_calculate_amd.workloads import vortex_grid, vortex_p3

pytestmark = pytest.mark.gpu

# relative drift of the totals after 20 steps (per variable, scaled by sum_e A_e |U_e|)
TOTALS_BOUND = {"fp64": 2e-14, "fp32": 1e-12, "bf16": 1e-7}


def areas(g):
    ex, ey = g.element_coords()
    return 0.5 * np.abs((ex[:, 1] - ex[:, 0]) * (ey[:, 2] - ey[:, 0]) - (ex[:, 2] - ex[:, 0]) * (ey[:, 1] - ey[:, 0]))


_GRIDS = {}


def grid(name):
    if name not in _GRIDS:
        _GRIDS[name] = vortex_grid(1024) if name == "p2" else vortex_p3(1024)[0]
    return _GRIDS[name]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
def test_totals_conserved_on_2M_triangles(prec):
    print(f"case: conservation P2 1024x1024 {prec}", flush=True)
    g = grid("p2")
    A = areas(g)
    s = Solver(g, degree=2, flux="lf", dt_policy="max", cfl=0.3, precision=prec)
    s.project_ic(IC_VORTEX, vortex_params())
    U0 = s.get_state()
    tot0 = (A * U0[:, 0, :]).sum(axis=1)
    scale = (A * np.abs(U0[:, 0, :])).sum(axis=1)
    s.advance(20)
    U = s.get_state()
    drift = np.abs((A * U[:, 0, :]).sum(axis=1) - tot0) / scale
    print("relative drift of the totals per variable", drift, flush=True)
    assert np.all(drift < TOTALS_BOUND[prec]), drift
    assert np.max(np.abs(U[0] - U0[0])) > 1e-6  # the vortex moved
    s.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
def test_free_stream_on_2M_triangles(prec):
    print(f"case: free stream P3 Roe 1024x1024 {prec}", flush=True)
    g = grid("p3")
    s = Solver(g, degree=3, flux="roe", dt_policy="max", cfl=0.15, precision=prec)
    U0 = np.zeros(s.state_shape())
    U0[:, 0, :] = np.array([1.2, 0.4, -0.3, 2.9])[:, None]
    s.set_state(U0, basis=BASIS_NATIVE)
    s.advance(5)
    dev = np.abs(s.get_state(basis=BASIS_NATIVE) - U0).max(axis=(0, 2))
    print("max deviation from the free stream per mode", dev, flush=True)
    assert dev.max() < (2e-15 if prec == "fp64" else 1e-15), dev
    s.close()
