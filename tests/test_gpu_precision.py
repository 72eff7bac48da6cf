"""Precision form: FP32 and BF16 contractions on the matrix cores, split about the cell mean.

Both contract only the deviations from the FP64 cell mean, and project only flux deviations
from the flux of the means (the mean-state part is applied in FP64), so their rounding error
scales with the in-cell variation.  Checked here:
* FP32: its error against the exact isentropic vortex matches the FP64 form's (the metric's
  "matched L2 error"), and its difference to the FP64 oracle stays far below 1e-6;
* BF16 (P1-P3): single stages and multi-step runs against the FP64 oracle within a stated
  bar (bf16 keeps 8 significant bits: 2^-9 relative rounding of each deviation), and the
  batched form on the same tiles.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from rkdg_heter_calculate_amd import IC_VORTEX, Batch, Solver, vortex_params
from rkdg_heter_calculate_amd.workloads import vortex_grid as wl_vortex_grid
from helpers import make_pair, rel_l2, vortex_grid, vortex_ic

pytestmark = pytest.mark.gpu
BF16_TOL = 2e-4    # relative L2 per variable against the FP64 oracle (BF16 contractions)
FP32_TOL = 1e-7    # relative L2 per variable against the FP64 oracle (FP32 contractions)


CASES = [(P, flux, prec, tol) for P in (1, 2) for flux in ("lf", "roe")
         for prec, tol in (("bf16", BF16_TOL), ("fp32", FP32_TOL))] + \
        [(3, flux, "bf16", BF16_TOL) for flux in ("lf", "roe")]


@pytest.mark.parametrize("P,flux,prec,tol", CASES)
def test_single_stages_against_oracle(P, flux, prec, tol):
    print(f"case: {prec} single stages P{P} {flux}", flush=True)
    g = vortex_grid(8, jitter=0.2, diagonal=1, seed=3)
    o, _ = make_pair(g, P=P, flux=flux)
    U0 = vortex_ic(o)
    s = Solver(g, degree=P, flux=flux, dt_policy="max", cfl=0.3, precision=prec)
    s.set_state(U0)
    for st in range(3):
        o.stage(st, 0.01)
        s.stage(st, 0.01)
        err = rel_l2(s.get_state(), o.get_state(), o.mass())
        print(f"stage {st}: relative L2 per variable {err}", flush=True)
        assert np.all(err < tol), (st, err)
        assert np.all(err > 0)  # really computed in reduced precision
    s.close()


@pytest.mark.parametrize("P,prec,tol", [(2, "bf16", BF16_TOL), (2, "fp32", FP32_TOL), (3, "bf16", BF16_TOL)])
def test_vortex_steps_against_oracle(P, prec, tol):
    print(f"case: {prec} 50 steps P{P} on 32x32 jittered", flush=True)
    g = vortex_grid(32, jitter=0.2, diagonal=1, seed=5)
    cfl = 0.3 if P < 3 else 0.15
    o, _ = make_pair(g, P=P, flux="lf", dt_policy="max", cfl=cfl)
    s = Solver(g, degree=P, flux="lf", dt_policy="max", cfl=cfl, precision=prec)
    s.set_state(vortex_ic(o))
    s.advance(50)
    o.advance(50)
    Us, Uo = s.get_state(), o.get_state()
    err = rel_l2(Us, Uo, o.mass())
    linf = [float(np.max(np.abs(Us[v] - Uo[v])) / np.max(np.abs(Uo[v]))) for v in range(4)]
    print("relative L2 vs FP64 oracle", err, "relative Linf", linf, flush=True)
    assert np.all(err < tol), err
    s.close()


def test_fp32_error_against_the_exact_vortex_matches_fp64():
    """Same mesh, same fixed dt, 100 steps: the FP32 form's L2 error against the exact
    advected vortex is the FP64 form's to within 1% (the metric's "matched L2 error"); the
    BF16 form's is stated beside it."""
    print("case: matched L2 error, 64x64 P2", flush=True)
    g = wl_vortex_grid(64)
    d = Solver(g, degree=2, flux="lf", dt_policy="max", cfl=0.3)
    d.project_ic(IC_VORTEX, vortex_params())
    dt = d.compute_dt()
    d.close()
    errs = {}
    for prec in ("fp64", "fp32", "bf16"):
        s = Solver(g, degree=2, flux="lf", dt_policy="fixed", dt_fixed=dt, precision=prec)
        s.project_ic(IC_VORTEX, vortex_params())
        s.advance(100)
        t, _ = s.time()
        errs[prec] = np.array(s.error_vs_exact(IC_VORTEX, vortex_params(t=t))["l2"])
        s.close()
    print("L2 error vs exact:", {k: v.tolist() for k, v in errs.items()}, flush=True)
    ratio = errs["fp32"] / errs["fp64"]
    assert np.all(np.abs(ratio - 1.0) < 0.01), ratio
    # BF16 is not matched at this resolution (8 significant bits on the deviations): its
    # error is stated, and bounded here
    print("bf16 / fp64 error ratio", errs["bf16"] / errs["fp64"], flush=True)
    assert np.all(errs["bf16"] / errs["fp64"] < 10.0), errs["bf16"] / errs["fp64"]


def test_bf16_batch_equals_single():
    print("case: bf16 batch vs single P2", flush=True)
    grids = [vortex_grid(n, jitter=0.2, diagonal=1, seed=n) for n in (7, 12, 19)]
    b = Batch(grids, [dict(degree=2, dt_policy="max", cfl=0.3, precision="bf16") for _ in grids])
    for i in range(len(grids)):
        b.project_ic(i, IC_VORTEX, vortex_params())
    b.autotune()
    assert b.tiles()["staged"] == 1
    b.advance(6)
    b.sync()
    for i, gr in enumerate(grids):
        s = Solver(gr, degree=2, dt_policy="max", cfl=0.3, precision="bf16")
        s.project_ic(IC_VORTEX, vortex_params())
        s.advance(6)
        # the same staged kernel on the same tiles: the same bits
        assert np.array_equal(b.get_state(i, basis=1), s.get_state(basis=1)), i
        s.close()
    b.close()
