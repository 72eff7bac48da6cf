"""rkdg_step_noalloc (the step that allocates nothing) against rkdg_stage(0, 1, 2, dt).

Both run the same stage kernels with the same dt; only the scratch differs (the library's
static device array instead of the context's own counters and residual pairs), so the
states are compared bitwise and the residual pairs exactly.  Two calls in a row, the
second on a smaller mesh; then two steps on one context; then the limiter's path.
"""
import ctypes

import numpy as np
import pytest

from rkdg_heter_calculate_amd import IC_RIEMANN4, IC_VORTEX, Solver, _lib as L, vortex_params
from rkdg_heter_calculate_amd.workloads import riemann_p1, vortex_grid

pytestmark = pytest.mark.gpu


class DeviceDoubles:
    """n doubles of device memory for the residual output (test plumbing through HIP)."""

    def __init__(self, n):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.n = n
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(8 * n)) == 0

    def read(self):
        out = np.zeros(self.n)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(8 * self.n), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def reference_and_noalloc(g, kw, ic_kind, ic, steps, res):
    """(state, residual pair) after `steps` steps through rkdg_stage and through
    rkdg_step_noalloc, from the same start and with the same dt."""
    lib = L.load()
    ref = Solver(g, **kw)
    ref.project_ic(ic_kind, ic)
    dt = ref.compute_dt()
    na = Solver(g, **kw)
    na.project_ic(ic_kind, ic)
    for _ in range(steps):
        for s in range(3):
            ref.stage(s, dt)
        L.check(lib.rkdg_step_noalloc(na.h, dt, res.ptr, None), na.h, "step_noalloc")
    na.sync()
    out = (ref.get_state(), ref.residual(), na.get_state(), res.read(), ref.flags(), na.flags())
    ref.close()
    na.close()
    return out


def test_noalloc_step_equals_stages_then_a_smaller_mesh():
    res = DeviceDoubles(2)
    kw = dict(degree=2, flux="lf", dt_policy="max", cfl=0.3)
    for n in (48, 14):  # the second call on a smaller mesh
        print(f"case: step_noalloc P2 {2 * n * n} triangles", flush=True)
        Ur, rr, Un, rn, _, _ = reference_and_noalloc(vortex_grid(n, seed=n), kw, IC_VORTEX, vortex_params(), 1, res)
        assert np.array_equal(Ur, Un)
        assert np.array_equal(rr, rn), (rr, rn)
        assert rr[0] > 0 and rr[1] > 0
    res.free()


@pytest.mark.parametrize("P,flux,prec", [(1, "lf", "fp64"), (3, "roe", "fp64"), (2, "roe", "fp64"),
                                         (2, "lf", "fp32"), (3, "roe", "bf16")])
def test_noalloc_two_steps_on_one_context(P, flux, prec):
    print(f"case: step_noalloc two steps P{P} {flux} {prec}", flush=True)
    res = DeviceDoubles(2)
    kw = dict(degree=P, flux=flux, dt_policy="max", cfl=0.3 if P < 3 else 0.15, precision=prec)
    Ur, rr, Un, rn, _, _ = reference_and_noalloc(vortex_grid(20, seed=3), kw, IC_VORTEX, vortex_params(), 2, res)
    assert np.array_equal(Ur, Un)
    assert np.array_equal(rr, rn), (rr, rn)
    res.free()


def test_noalloc_step_with_the_limiter():
    print("case: step_noalloc P1 Roe TVB", flush=True)
    res = DeviceDoubles(2)
    g, kw, ic, _ = riemann_p1(32)
    Ur, rr, Un, rn, fr, fn = reference_and_noalloc(g, kw, IC_RIEMANN4, ic, 6, res)  # limited from step 3
    assert np.array_equal(Ur, Un)
    assert np.array_equal(rr, rn)
    assert np.array_equal(fr, fn) and (fr != 0).any()
    res.free()


def test_noalloc_step_rejects_a_bad_dt():
    g = vortex_grid(8, seed=1)
    s = Solver(g, degree=2, dt_policy="max", cfl=0.3)
    s.project_ic(IC_VORTEX, vortex_params())
    assert L.load().rkdg_step_noalloc(s.h, 0.0, None, None) == L.ERR_ARG
    s.close()
