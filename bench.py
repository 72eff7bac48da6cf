"""Benchmark: DOF-stage updates/s of the RKDG hot path on one MI355X.

Workload (default): isentropic vortex (gamma 1.4, beta 5, free stream rho=u=v=p=1) on the
periodic [0,10]^2 square, 512 x 512 split quads with a seed-chosen diagonal per quad and
interior vertices jittered by U(+-0.2 h) (524,288 triangles), P2 basis (6 modes per
variable), 7-point volume / 3-point edge rules, Lax-Friedrichs flux, SSP-RK3 at CFL 0.3,
FP64.  A step is one SSP-RK3 step (3 stage kernels) with the dt reduced on the device.

    python bench.py [--gpus 1] [--steps 200] [--warmup 10]

Prints one JSON line.  value = sum E*Np*4*3*steps / wall time of the timed steps, with the
state resident in HBM; the timed region is bracketed by device synchronisations and HIP
events on the solver's stream (the stream the stage kernels are launched on).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def progress(msg):
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes_per_element_stage(Np):
    # The roofline yardstick: U read + U written + step-start state read in 2 of 3 stages,
    # the reference's 14 geometry doubles per element, 3 neighbour ids; neighbour re-reads
    # counted as cache hits
    return 8 * 4 * Np * (1 + 1 + 2.0 / 3.0) + 112 + 12


def moved_bytes_per_element_stage(Np):
    # What the stage kernels read and write: the same state traffic, 5 geometry doubles
    # (inverse Jacobian and area; edge normals and lengths are derived), 3 neighbour ids
    return 8 * 4 * Np * (1 + 1 + 2.0 / 3.0) + 40 + 12


def workload(n=512, P=2):
    from rkdg_heter_calculate_amd.workloads import vortex_grid
    return vortex_grid(n)


def large_p2_point(n, steps):
    """Stage-kernel roofline at a size past the 256 MB Infinity Cache (2n^2 triangles)."""
    from rkdg_heter_calculate_amd import Solver, IC_VORTEX, vortex_params
    g = workload(n)
    s = Solver(g, degree=2, flux="lf", dt_policy="max", cfl=0.3)
    s.project_ic(IC_VORTEX, vortex_params())
    s.advance(40)  # ~40 ms of sustained load first: the clocks ramp up after the host set-up
    ms = s.time_steps(steps)
    B = algorithmic_bytes_per_element_stage(6)
    E = g.n_elements
    out = {"elements": E, "steps": steps, "ms_per_step": ms / steps,
           "value": E * 6 * 4 * 3 * steps / (ms / 1e3),
           "stage_ms": ms / (3 * steps), "achieved_GBs": B * E / (ms / (3 * steps) / 1e3) / 1e9}
    out["frac"] = out["achieved_GBs"] / HBM_PEAK_GBS
    s.close()
    return out


def batch_point(count, steps, precision="fp64", degree=2):
    """Batched sweep: `count` periodic vortex instances of uneven size in one launch per stage
    (at P3 with half the instances' CFL numbers)."""
    import ctypes
    import numpy as np
    from rkdg_heter_calculate_amd import _lib as L
    from rkdg_heter_calculate_amd.solver import make_params, mesh_view
    from rkdg_heter_calculate_amd.workloads import batch_instances, vortex_grid
    lib = L.load()
    inst = batch_instances(count)
    grids = [vortex_grid(n, seed=1000 + i) for i, (n, _, _) in enumerate(inst)]
    views = [mesh_view(gr) for gr in grids]
    meshes = (L.Mesh * count)()
    params = (L.Params * count)()
    for i in range(count):
        meshes[i] = views[i]
        params[i] = make_params(degree=degree, flux="lf", dt_policy="max",
                                cfl=inst[i][2] * (0.5 if degree == 3 else 1.0), precision=precision)
    h = ctypes.c_void_p()
    L.bcheck(lib.rkdg_batch_create(count, meshes, params, ctypes.byref(h)), None, "batch_create")
    for i in range(count):
        vp = np.array(inst[i][1], dtype=np.float64)
        L.bcheck(lib.rkdg_batch_project_ic(h, i, 1, L.dptr(vp), len(vp)), h, "batch_project_ic")
    best = ctypes.c_int32()
    L.bcheck(lib.rkdg_batch_autotune(h, ctypes.byref(best)), h, "autotune")
    L.bcheck(lib.rkdg_batch_advance(h, 10, 1), h, "warmup")  # sustained load before the timed steps
    L.bcheck(lib.rkdg_batch_sync(h), h, "sync")
    t0 = time.perf_counter()
    L.bcheck(lib.rkdg_batch_advance(h, steps, 1), h, "advance")
    L.bcheck(lib.rkdg_batch_sync(h), h, "sync")
    dt = time.perf_counter() - t0
    info = np.zeros(4, np.int64)
    lib.rkdg_batch_info(h, info.ctypes.data_as(L.P_int64))
    lib.rkdg_batch_destroy(h)
    E = int(info[1])
    Np = (degree + 1) * (degree + 2) // 2
    return {"instances": count, "elements": E, "degree": degree, "steps": steps, "block": int(best.value),
            "value": E * Np * 4 * 3 * steps / dt, "ms_per_step": dt * 1e3 / steps,
            "note": f"{precision} batched form, hipGraph-replayed stages, autotuned tiles and persistence"}


def precision_point(g, P, steps, prec="fp32"):
    """A precision form (FP32 or BF16 contractions, FP64 state) on the same workload: its
    timing; then the FP64 form and it run the same `steps` steps of one fixed dt, and give
    each form's error against the exact vortex (their ratio: "matched" when within 10%) and
    the form's difference to the FP64 form (which matches the oracle to 1e-12, so that
    difference is the form's own error)."""
    import numpy as np
    from rkdg_heter_calculate_amd import Solver, IC_VORTEX, vortex_params
    from rkdg_heter_calculate_amd import _lib as L
    s = Solver(g, degree=P, flux="lf", dt_policy="max", cfl=0.3, precision=prec)
    s.project_ic(IC_VORTEX, vortex_params())
    s.advance(100)  # sustained load before the timed steps
    s.sync()
    ms = s.time_steps(steps)
    t_sim, _ = s.time()
    err_exact = s.error_vs_exact(IC_VORTEX, vortex_params(t=t_sim))
    s.close()
    E = g.n_elements
    Np = (P + 1) * (P + 2) // 2
    out = {"value": E * Np * 4 * 3 * steps / (ms / 1e3), "ms_per_step": ms / steps,
           "stage_ms": ms / (3 * steps), "l2_vs_exact_timed_run": err_exact["l2"]}
    # matched-step comparison against the FP64 form
    d = Solver(g, degree=P, flux="lf", dt_policy="max", cfl=0.3)
    d.project_ic(IC_VORTEX, vortex_params())
    dt0 = d.compute_dt()
    d.close()
    res, exact = {}, {}
    for form in ("fp64", prec):
        x = Solver(g, degree=P, flux="lf", dt_policy="fixed", dt_fixed=dt0, precision=form)
        x.project_ic(IC_VORTEX, vortex_params())
        x.advance(steps)
        t_f, _ = x.time()
        exact[form] = x.error_vs_exact(IC_VORTEX, vortex_params(t=t_f))
        res[form] = x.get_state(L.BASIS_NATIVE)  # orthonormal on the reference triangle
        x.close()
    ratio = [a / b for a, b in zip(exact[prec]["l2"], exact["fp64"]["l2"])]
    out.update({"l2_vs_exact": exact[prec]["l2"], "linf_vs_exact": exact[prec]["linf"],
                "l2_vs_exact_fp64_same_steps": exact["fp64"]["l2"], "l2_vs_exact_ratio_to_fp64": ratio,
                "matched_l2": bool(all(abs(r - 1.0) <= 0.1 for r in ratio))})
    a, b = res["fp64"], res[prec]
    area = _areas(g)
    # L2 norm of a field = sqrt(sum_e |J_e| sum_i c_i^2) in the orthonormal reference basis
    rel = _rel_l2(a, b, area)
    linf = [float(np.max(np.abs(b[v] - a[v])) / np.max(np.abs(a[v]))) for v in range(4)]
    kind = "FP32 on v_mfma_f32_4x4x1f32" if prec == "fp32" else "BF16 on v_mfma_f32_4x4x4_16b_bf16"
    out.update({"rel_l2_vs_fp64": rel, "rel_linf_vs_fp64": linf, "compare_steps": steps, "compare_dt": dt0,
                "note": f"{kind} contractions of the deviations from the FP64 cell mean (FP32 accumulation); "
                        "FP64 state, means, fluxes, alpha, RK and CFL"})
    return out


def _rel_l2(a, b, area):
    """Relative L2 difference per conserved variable of two native-basis states."""
    import numpy as np
    return [float(np.sqrt(np.sum(area * np.sum((b[v] - a[v]) ** 2, axis=0)) / np.sum(area * np.sum(a[v] ** 2, axis=0))))
            for v in range(4)]


def _areas(g):
    import numpy as np
    ex, ey = g.element_coords()
    return 0.5 * np.abs((ex[:, 1] - ex[:, 0]) * (ey[:, 2] - ey[:, 0]) - (ex[:, 2] - ex[:, 0]) * (ey[:, 1] - ey[:, 0]))


def riemann_point(n=1024, t_end=0.3):
    """Four-quadrant Riemann problem, P1, Roe, TVB M=50 (+ positivity), FP64 and the
    precision form (limiter in FP64), each run to t_end from the same cell averages."""
    from rkdg_heter_calculate_amd import Solver, IC_RIEMANN4
    from rkdg_heter_calculate_amd import _lib as L
    from rkdg_heter_calculate_amd.workloads import riemann_p1
    g, kw, ic, _ = riemann_p1(n)
    out, states = {"elements": g.n_elements, "t_end": t_end}, {}
    for prec in ("fp64", "fp32"):
        s = Solver(g, precision=prec, **kw)
        s.project_ic(IC_RIEMANN4, ic)
        s.sync()
        t0 = time.perf_counter()
        nsteps, t = 0, 0.0
        while t < t_end:
            n, t = s.advance(5000, t_end=t_end)
            nsteps += n
            progress(f"riemann {prec}: t = {t:.4f} after {nsteps} steps")
            if n == 0:
                break
        s.sync()
        wall = time.perf_counter() - t0
        states[prec] = s.get_state(L.BASIS_NATIVE)
        progress(f"riemann {prec}: {nsteps} steps to t = {t:.4f} in {wall:.1f} s")
        out[prec] = {"steps": nsteps, "t": t, "wall_s": wall,
                     "value": g.n_elements * 3 * 4 * 3 * nsteps / wall, "ms_per_step": wall * 1e3 / nsteps,
                     "limited_cells_last_stage": int((s.flags() != 0).sum())}
        s.close()
    out["rel_l2_fp32_vs_fp64"] = _rel_l2(states["fp64"], states["fp32"], _areas(g))
    out["note"] = "value counts limiter launches inside the step; wall time includes one host check per 32 steps"
    return out


def p3_point(n=2048, steps=50):
    """Large P3 vortex (Roe, CFL 0.2): FP64 and the precision form with FP32 and with BF16
    contractions, stage-kernel roofline and each precision form's difference to FP64 after
    the same fixed-dt steps."""
    from rkdg_heter_calculate_amd import Solver, IC_VORTEX, vortex_params
    from rkdg_heter_calculate_amd import _lib as L
    from rkdg_heter_calculate_amd.workloads import vortex_p3
    g, kw, vp, _ = vortex_p3(n)
    E = g.n_elements
    out = {"elements": E, "steps": steps}
    B = algorithmic_bytes_per_element_stage(10)
    dt0 = None
    states = {}
    for prec in ("fp64", "fp32", "bf16"):
        s = Solver(g, precision=prec, **kw)
        s.project_ic(IC_VORTEX, vp)
        if dt0 is None:
            dt0 = s.compute_dt()
        s.advance(2)
        ms = s.time_steps(steps)
        progress(f"p3 {prec}: {ms / steps:.2f} ms per step")
        out[prec] = {"ms_per_step": ms / steps, "stage_ms": ms / (3 * steps),
                     "value": E * 10 * 4 * 3 * steps / (ms / 1e3),
                     "achieved_GBs": B * E / (ms / (3 * steps) / 1e3) / 1e9}
        s.close()
        f = Solver(g, precision=prec, dt_policy="fixed", dt_fixed=dt0, degree=3, flux="roe")
        f.project_ic(IC_VORTEX, vp)
        f.advance(steps)
        states[prec] = f.get_state(L.BASIS_NATIVE)
        f.close()
    area = _areas(g)
    out["rel_l2_fp32_vs_fp64"] = _rel_l2(states["fp64"], states["fp32"], area)
    out["rel_l2_bf16_vs_fp64"] = _rel_l2(states["fp64"], states["bf16"], area)
    return out


def cpu_baseline(g, P, steps, threads):
    """The oracle (CPU restatement of the reference algorithm) on a bounded sample."""
    from oracle import pyoracle as po
    from rkdg_heter_calculate_amd import vortex_params
    po.set_threads(threads)
    ex, ey = g.element_coords()
    o = po.Oracle(ex, ey, g.neighbour, g.shared_edge, g.boundary, P=P, flux=0, dt_policy=1, cfl=0.3)
    o.project(1, vortex_params())
    t0 = time.perf_counter()
    o.advance(steps)
    dt = time.perf_counter() - t0
    Np = (P + 1) * (P + 2) // 2
    o.close()
    return g.n_elements * Np * 4 * 3 * steps / dt, dt


def host_cpu():
    """CPU model, logical CPUs of the host and CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "affinity_cpus": allowed}


def cpu_threads_all():
    """Threads of the all-cores run: the process's CPU share (OMP_NUM_THREADS when the
    host sets one, otherwise every CPU the process may run on)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    hc = host_cpu()
    if env.isdigit() and int(env) > 0:
        return min(int(env), hc["affinity_cpus"])
    return hc["affinity_cpus"]


def read_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512, help="split quads per side")
    ap.add_argument("--degree", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=50, help="oracle steps of the all-cores CPU baseline sample")
    ap.add_argument("--cpu-steps-1t", type=int, default=3, help="oracle steps of the one-thread CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip every extra leg")
    ap.add_argument("--extra", default="p2_2M,batch,batch_fp32,riemann,p3",
                    help="comma-separated extra legs: p2_2M, batch, batch_fp32, batch_bf16, batch_p3, riemann, p3 "
                         "(p3 builds an 8.4M-triangle mesh)")
    ap.add_argument("--no-precision", action="store_true", help="skip the FP32-MFMA leg")
    ap.add_argument("--form", choices=("fp64", "fp32", "bf16"), default="fp64",
                    help="form of the main timed leg (fp32 / bf16: the precision form with FP32 or BF16 "
                         "contractions; the comparison legs are then skipped)")
    ap.add_argument("--batch", type=int, default=256, help="instances of the batched leg")
    ap.add_argument("--riemann-t-end", type=float, default=0.01,
                    help="end time of the riemann leg (0.3, the problem's own, takes about 2 x 70k steps)")
    ap.add_argument("--copy-mib", type=int, default=1024, help="size of the stream-copy measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs the hot path on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is described, not built)", file=sys.stderr)
        sys.exit(2)

    from rkdg_heter_calculate_amd import Solver, IC_VORTEX, vortex_params

    P = args.degree
    Np = (P + 1) * (P + 2) // 2
    g = workload(args.n, P)
    E = g.n_elements
    lowp = args.form != "fp64"
    cfl = 0.3 if P <= 2 else 0.2  # P3: the large-mesh P3 workload's CFL number
    s = Solver(g, degree=P, flux="lf", dt_policy="max", cfl=cfl, precision=args.form)
    s.project_ic(IC_VORTEX, vortex_params())
    # measured copy bandwidth, right before the warm-up steps: after the host set-up the
    # GPU has been idle, and its clocks take about 20 ms of sustained load to come back up
    # (the first stage launches after an idle gap run up to 30% longer), which a few
    # warm-up steps alone do not cover
    copy_gbs = None
    if args.copy_mib > 0:
        import ctypes
        from rkdg_heter_calculate_amd import _lib as L
        gbs = ctypes.c_double()
        L.check(L.load().rkdg_measure_copy(0, args.copy_mib << 20, 20, ctypes.byref(gbs)), None, "measure_copy")
        copy_gbs = gbs.value
        progress(f"stream copy: {copy_gbs:.0f} GB/s")
    progress(f"{E} elements set up ({args.form}); {args.warmup} warm-up steps")
    s.advance(args.warmup)
    s.sync()
    # the timed region: exactly `steps` steps of the production path (hipGraph-replayed
    # step pairs, dt on the device), one HIP event before and one after on the solver's
    # stream, device synchronisations on both sides
    t0 = time.perf_counter()
    total_ms = s.time_steps(args.steps)
    s.sync()
    wall = time.perf_counter() - t0
    progress(f"timed {args.steps} steps: {total_ms / args.steps:.3f} ms per step")
    t_sim, nsteps = s.time()
    err = s.error_vs_exact(IC_VORTEX, vortex_params(t=t_sim))
    # untimed: the same steps launched one by one with events around every stage launch
    # (an upper bound of the kernel duration: each bracket includes that launch's dispatch)
    r = s.time_advance(args.steps)

    total_s = total_ms / 1e3
    dofs = E * Np * 4
    value = dofs * 3 * args.steps / total_s
    B = algorithmic_bytes_per_element_stage(Np)
    per_launch = B * E
    avg_launch_ms = total_ms / (3 * args.steps)  # the stage kernels back to back, gaps included
    achieved = per_launch / (avg_launch_ms / 1e3) / 1e9
    kfam = ("k_stage_pz" if P == 3 else "k_stage_px") if lowp else ("k_stage_rp" if P == 1 else "k_stage")
    kname = f"{kfam}_bf16_p{P}" if args.form == "bf16" else f"{kfam}_p{P}"
    traffic = read_traffic(f"{kname}_lf_{args.n}x{args.n}_hbm.md")

    cpu = None
    if not args.no_cpu:
        hc = host_cpu()
        threads = cpu_threads_all()
        v, secs = cpu_baseline(g, P, args.cpu_steps, threads)
        progress(f"cpu baseline: {args.cpu_steps} oracle steps on {threads} threads in {secs:.1f} s")
        v1, secs1 = cpu_baseline(g, P, args.cpu_steps_1t, 1)
        progress(f"cpu baseline: {args.cpu_steps_1t} oracle steps on 1 thread in {secs1:.1f} s")
        cpu = {"value": v, "unit": "DOF-stage updates/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} SSP-RK3 steps of the same {E}-triangle P{P} mesh "
                         f"({secs:.1f} s, OpenMP over elements, {threads} threads)",
               "single_thread": {"value": v1, "cores": 1,
                                 "sample": f"{args.cpu_steps_1t} SSP-RK3 steps of the same mesh ({secs1:.1f} s)"},
               "cpu_model": hc["cpu_model"], "nproc": hc["nproc"], "affinity_cpus": hc["affinity_cpus"],
               "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}

    prec, prec_bf16 = None, None
    if not args.no_precision and not lowp:
        prec = precision_point(g, P, args.steps, "fp32")
        progress("precision form (fp32) leg done")
        prec_bf16 = precision_point(g, P, args.steps, "bf16")
        progress("precision form (bf16) leg done")

    extra = {}
    if not args.no_extra:
        legs = [x for x in args.extra.split(",") if x]
        runners = {"p2_2M": ("p2_2M_fp64", lambda: large_p2_point(1024, 20)),
                   "batch": ("batch_fp64", lambda: batch_point(args.batch, 100)),
                   "batch_fp32": ("batch_fp32_mfma", lambda: batch_point(args.batch, 100, precision="fp32")),
                   "batch_bf16": ("batch_bf16_mfma", lambda: batch_point(args.batch, 100, precision="bf16")),
                   "riemann": ("riemann_p1_2M", lambda: riemann_point(1024, args.riemann_t_end)),
                   "p3": ("vortex_p3_8M", lambda: p3_point(2048, 50)),
                   "batch_p3": ("batch_p3", lambda: {p: batch_point(args.batch, 50, precision=p, degree=3)
                                                     for p in ("fp64", "fp32", "bf16")})}
        for leg in legs:
            if leg not in runners:
                sys.exit(f"bench.py: unknown extra leg {leg!r}")
            key, fn = runners[leg]
            t0 = time.perf_counter()
            extra[key] = fn()
            progress(f"extra leg {leg} done in {time.perf_counter() - t0:.1f} s")

    out = {
        "metric": "DOF-stage updates/s on 1 MI355X, FP64 vs FP32-MFMA at matched L2 error",
        "value": value,
        "unit": "DOF-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32+f64", "bf16": "bf16+f32+f64"}[args.form],
        "data": "synthetic (seeded jittered split-quad mesh, projected isentropic vortex)",
        "config": {"workload": f"isentropic vortex, periodic [0,10]^2, {args.n}x{args.n} jittered split quads "
                               f"({E} triangles), P{P}, LF, SSP-RK3 CFL {cfl}, " +
                               {"fp64": "FP64", "fp32": "FP32-MFMA", "bf16": "BF16-MFMA"}[args.form],
                   "elements": E, "degree": P, "dofs": dofs,
                   "form": {"fp64": "plain FP64", "fp32": "precision (FP32-MFMA contractions, FP64 state)",
                            "bf16": "precision (BF16-MFMA contractions, FP64 state)"}[args.form]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"{kfam}<{P},LF>", "bytes_per_launch": per_launch,
                     "bytes_moved_per_launch": moved_bytes_per_element_stage(Np) * E,
                     "avg_launch_ms": avg_launch_ms, "avg_launch_ms_event_brackets": r["stage_ms"],
                     "measured_copy_GBs": copy_gbs},
        "cpu_baseline": cpu,
        "precision_form": prec,
        "precision_form_bf16": prec_bf16,
        "wall_s": wall,
        "l2_vs_exact": err["l2"],
        "linf_vs_exact": err["linf"],
        "sim_time": t_sim,
        "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
