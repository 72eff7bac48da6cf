"""The batching form against the CPU oracle, instance by instance.

The reference runs one mesh per RkdgAdvance::advance (src/rkdgadvance.cpp:91-243) with
the dt of src/kernelfunctions.cu:3-81; a batch must reproduce, for every instance, what
the oracle computes for that mesh alone: uneven sizes (1 to 13 layout tiles), per-instance
vortex strength, Mach number, flow angle and CFL number, after the autotuner and under
every tile configuration it can pick.  Bars: 1e-12 relative L2 per variable (FP64),
1e-6 (FP32-MFMA contractions), 1e-3 (BF16-MFMA contractions: 8 significant bits on the
deviations, and the strongest vortices of the set on 7x7 to 20x20 meshes).
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from rkdg_heter_calculate_amd import BC_FARFIELD, Batch, Grid
from rkdg_heter_calculate_amd.workloads import batch_instances, vortex_grid
from helpers import rel_l2

pytestmark = pytest.mark.gpu
STEPS = 6
TOL = {"fp64": 1e-12, "fp32": 1e-6, "bf16": 1e-3}
CFL = {1: 0.25, 2: 0.22, 3: 0.14}


def instance_set(P, count=9):
    """Per-instance (grid, vortex parameters, cfl): the batched sweep's generator at small sizes."""
    out = []
    for i, (n, vp, _) in enumerate(batch_instances(count=count, seed=21, nmin=6, nmax=40)):
        out.append((vortex_grid(n, seed=1000 + i), vp, CFL[P] + 0.01 * (i % 4)))
    return out


def oracle_run(g, P, vp, cfl, fs=(1.0, 0.0, 0.0, 2.5), flux=0):
    ex, ey = g.element_coords()
    o = po.Oracle(ex, ey, g.neighbour, g.shared_edge, g.boundary, P=P, flux=flux, dt_policy=1, cfl=cfl,
                  freestream=fs)
    o.project(1, vp)
    n, t, _ = o.advance(STEPS)
    assert n == STEPS
    return o.get_state(), t, o.mass()


_ORACLE = {}


def oracle_states(P):
    if P not in _ORACLE:
        inst = instance_set(P)
        _ORACLE[P] = (inst, [oracle_run(g, P, vp, cfl) for g, vp, cfl in inst])
    return _ORACLE[P]


def make_batch(inst, P, prec):
    b = Batch([g for g, _, _ in inst], [dict(degree=P, flux="lf", dt_policy="max", cfl=cfl, precision=prec)
                                        for _, _, cfl in inst])
    for i, (_, vp, _) in enumerate(inst):
        b.project_ic(i, 1, vp)
    return b


def check_against_oracle(b, inst, ref, prec, label):
    worst = np.zeros(4)
    for i in range(len(inst)):
        Uo, to, M = ref[i]
        err = rel_l2(b.get_state(i), Uo, M)
        worst = np.maximum(worst, err)
        assert np.all(err < TOL[prec]), (label, i, inst[i][0].n_elements, err)
        # the time sums the CFL dts of the instance's own cell means: 1e-13 in FP64; in the
        # precision form the means carry the contractions' rounding
        assert abs(b.time(i) - to) <= (1e-13 if prec == "fp64" else TOL[prec]) * to, (label, i, b.time(i), to)
    print(label, "worst relative L2 per variable", worst, flush=True)


def test_instance_set_spans_tile_counts():
    inst, _ = oracle_states(2)
    tiles = {math.ceil(g.n_elements / 256) for g, _, _ in inst}
    assert len(tiles) >= 3, tiles
    assert len({round(vp[1], 6) for _, vp, _ in inst}) == len(inst)  # per-instance beta


@pytest.mark.parametrize("P,prec", [(2, "fp64"), (2, "fp32"), (1, "fp64"), (3, "fp64"), (1, "fp32"), (3, "fp32"),
                                    (2, "bf16"), (1, "bf16"), (3, "bf16")])
def test_autotuned_batch_matches_oracle(P, prec):
    print(f"case: autotuned batch P{P} {prec}", flush=True)
    inst, ref = oracle_states(P)
    b = make_batch(inst, P, prec)
    b.autotune()
    cfg = b.tiles()
    print("autotuned tiles", cfg, flush=True)
    b.advance(STEPS)
    b.sync()
    check_against_oracle(b, inst, ref, prec, f"P{P} {prec} autotuned {cfg}")
    b.close()


CONFIGS = [("fp64", 1, 256, pw) for pw in (0, 4, 8)] + \
          [("fp64", 0, blk, pw) for blk in (64, 128, 256) for pw in (0, 4, 8)] + \
          [("fp32", 1, 256, pw) for pw in (0, 4, 8)] + [("fp32", 0, 64, pw) for pw in (0, 4, 8)] + \
          [("bf16", 1, 256, pw) for pw in (0, 4, 8)]


@pytest.mark.parametrize("prec,staged,block,per_cu", CONFIGS)
def test_every_tile_configuration_matches_oracle(prec, staged, block, per_cu):
    print(f"case: batch P2 {prec} staged={staged} block={block} per_cu={per_cu}", flush=True)
    inst, ref = oracle_states(2)
    b = make_batch(inst, 2, prec)
    b.set_tiles(staged, block, per_cu)
    assert b.tiles() == {"staged": staged, "block": block, "per_cu": per_cu}
    b.advance(STEPS)
    b.sync()
    check_against_oracle(b, inst, ref, prec, f"{prec} staged={staged} block={block} per_cu={per_cu}")
    b.close()


def test_farfield_instances_match_oracle():
    """Far-field edges take the launch's free stream: two far-field meshes that share it,
    beside a periodic mesh whose (unused) free-stream parameter differs."""
    print("case: batch with far-field instances", flush=True)
    uinf, vinf = 0.5, 0.3
    fs = (1.0, uinf, vinf, 1.0 / 0.4 + 0.5 * (uinf * uinf + vinf * vinf))
    vp_open = [1.4, 3.0, uinf, vinf, 0.0, 0.0, 0.0, 0.0, 0.0]  # no periodic images
    grids = [Grid.splitquad(n, n, -5.0, -5.0, 10.0, 10.0, periodic=False, diagonal=1, jitter=0.15, seed=n,
                            bc=BC_FARFIELD) for n in (9, 23)]
    gp = vortex_grid(14, seed=5)
    vp_per = [1.4, 4.0, 1.0, -0.5, 5.0, 5.0, 10.0, 10.0, 0.0]
    inst = [(grids[0], vp_open, 0.2), (gp, vp_per, 0.22), (grids[1], vp_open, 0.24)]
    fss = [fs, (2.0, 0.0, 0.0, 9.0), fs]
    b = Batch([g for g, _, _ in inst], [dict(degree=2, dt_policy="max", cfl=c, freestream=list(f))
                                        for (_, _, c), f in zip(inst, fss)])
    for i, (_, vp, _) in enumerate(inst):
        b.project_ic(i, 1, vp)
    b.advance(STEPS)
    b.sync()
    ref = [oracle_run(g, 2, vp, c, fs=f) for (g, vp, c), f in zip(inst, fss)]
    check_against_oracle(b, inst, ref, "fp64", "far-field batch")
    b.close()


@pytest.mark.parametrize("prec,per_cu", [("fp32", 0), ("fp32", 4), ("bf16", 0), ("bf16", 8)])
def test_p3_precision_batch_runs_on_staged_tiles(prec, per_cu):
    """At P3 the precision form's batch runs on the 256-element FP32-image tiles (k_stage_pz),
    one workgroup per tile or a persistent grid, and matches the oracle within the form's bar."""
    print(f"case: batch P3 {prec} staged tiles per_cu={per_cu}", flush=True)
    inst, ref = oracle_states(3)
    b = make_batch(inst, 3, prec)
    assert b.tiles()["staged"] == 1 and b.tiles()["block"] == 256
    b.set_tiles(1, 256, per_cu)
    b.advance(STEPS)
    b.sync()
    check_against_oracle(b, inst, ref, prec, f"P3 {prec} staged per_cu={per_cu}")
    b.close()
