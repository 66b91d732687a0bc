#!/usr/bin/env python3
"""Throughput of the multigrid solve on one MI355X.

Workload (the headline configuration): Poisson on a 1025 x 1025 x 513 grid
(h = 1/1024, x, y in [0,1], z in [0,0.5]), 7 levels, F-cycles with red-black
Gauss-Seidel (omega = 1, 2 sweeps), full-weighting restriction, trilinear
prolongation, Neumann faces in x and y, Dirichlet in z, u = x^3 y^2 z,
f = 2x^3 z + 6x y^2 z, FP64, solved from u = 0 to ||r||inf/||f||inf <= 1e-10
(consistent boundary treatment: the reference's own coarse Neumann data stalls
at O(h) and never reaches that tolerance).

A step is one complete solve to the tolerance; the right-hand side and the
hierarchy stay resident in HBM, each step restarts from u = 0 on the device.
value = fine-grid DOF x steps / wall time of the timed steps.

The mixed-precision leg (same grid and cycle, all-Dirichlet faces,
u = cos(x z^2) sin(y^3)) times the FP32-cycle / FP64-refinement solver and the
plain FP64 solver on the same problem, and reports their DOF/s, the ratio, and
the max-abs difference of the mixed solution from the FP64 one and from the
analytic u; the same with levels >= 3 of the FP32 cycle stored in FP16
(--no-mixed skips it).

Prints one JSON line (see the keys below).  --gpus N > 1 exits: this build
runs on one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "multigrid-poisson_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import mgp  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_time_ms(fn, reps, stream):
    """Average duration of `fn(stream)` over `reps` launches, HIP events on the
    stream the kernel is launched on."""
    L = mgp.lib()
    e0, e1 = mgp.vp(), mgp.vp()
    mgp.check(L.mgp_event_create(ctypes.byref(e0)))
    mgp.check(L.mgp_event_create(ctypes.byref(e1)))
    fn(stream)  # warm
    mgp.check(L.mgp_event_record(e0, stream))
    for _ in range(reps):
        fn(stream)
    mgp.check(L.mgp_event_record(e1, stream))
    mgp.check(L.mgp_event_sync(e1))
    ms = ctypes.c_float()
    mgp.check(L.mgp_event_elapsed_ms(e0, e1, ctypes.byref(ms)))
    L.mgp_event_destroy(e0)
    L.mgp_event_destroy(e1)
    return ms.value / reps


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1025, help="points along x and y (z gets n//2+1)")
    ap.add_argument("--levels", type=int, default=7)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--nsmooth", type=int, default=2)
    ap.add_argument("--omega", type=float, default=1.0,
                    help="SOR factor of the red-black Gauss-Seidel smoother in the headline and mixed legs")
    ap.add_argument("--coarse-sweeps", type=int, default=32,
                    help="smoothing passes per visit of the coarsest level (17x17x9)")
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--coarse-cg", action="store_true",
                    help="solve the coarsest level by CG (MGP_SOLVER_COARSE_CG) in the FP64 and mixed legs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch kernels one by one (for profilers)")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="mgp_tuning_set before the run (e.g. sweep.target_wgs=4096)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32+FP64-refinement leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-solves leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--serve-count", type=int, default=500)
    ap.add_argument("--serve-sample", type=int, default=64, help="requests solved one at a time in FP64")
    ap.add_argument("--serve-load", type=float, default=3.0,
                    help="Poisson arrival rate as a multiple of the one-at-a-time FP64 rate")
    ap.add_argument("--serve-workers", type=int, default=12, help="solves in flight in the serving leg")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run once each: a fill (write-only) and an inf-norm (read-only) of one 4.3 GB array, "
                         "then one fine-level one-pass sweep of the workload; no JSON line (for rocprofv3 --pmc)")
    ap.add_argument("--probe-batch", action="store_true",
                    help="run the batched leg's solve launch by launch (for profilers); no JSON line")
    ap.add_argument("--probe-mixed", type=int, default=0, metavar="PROBLEM",
                    help="two FP32-refinement solves of problem 2 or 3 launch by launch (for kernel traces); "
                         "no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs the solve on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is a design note only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    mgp.require_gpu()
    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.probe_mixed:
        probe_mixed(args, args.probe_mixed)
        return
    if args.probe_batch:
        probe_batch(args)
        return
    for kv in args.tune:
        k, v = kv.split("=")
        mgp.check(mgp.lib().mgp_tuning_set(k.encode(), int(v)), "mgp_tuning_set")
    n, nz = args.n, args.n // 2 + 1
    dims = (n, n, nz)
    h = 1.0 / (n - 1)
    dof = n * n * nz
    log("bench: %dx%dx%d, %d levels, F-cycle GS(%d) FW, tol %.0e" % (n, n, nz, args.levels, args.nsmooth, args.tol))
    s = mgp.Solver(dims, args.levels, h, (0.0, 0.0, 0.0), "NNNNDD", mgp.CONSISTENT, mgp.GAUSS_SEIDEL,
                   mgp.FULL_WEIGHTING, problem=2, flags=solver_flags(args),
                   maxiter=100, miniter=1, tolerance=args.tol,
                   coarse_sweeps=args.coarse_sweeps)
    s.to_device()
    log("bench: hierarchy resident; warmup")
    elapsed, cycles, rel = timed_solves(s, args)
    ms_per_step = 1e3 * elapsed / args.steps
    value = dof * args.steps / elapsed
    log("bench: %.1f ms/solve, %d cycles, rel_res %.3e" % (ms_per_step, cycles[-1], rel[-1]))
    if max(rel) > args.tol:
        log("bench: WARNING: tolerance not reached")

    # Dominant kernels measured live on the fine level: the two-sweep pass with
    # the coarse correction added first, which the F-cycle runs twice per
    # cycle on the fine level (24 B per box point: read u and f once, write u
    # once, for two sweeps, plus 8 B per coarse point read once), the plain
    # two-sweep pass (once per cycle), the one-pass sweep (24 B per point per
    # sweep) and the residual with its max-norm (24 B per point: read u and f,
    # write r).
    L = mgp.lib()
    stream = mgp.vp()
    mgp.check(L.mgp_stream_create(ctypes.byref(stream)))
    u = s.device_array("u")
    f_ = s.device_array("f")
    r = s.device_array("r")
    u1 = s.device_array("u", 1)
    F = mgp.Faces()
    for q in range(6):
        t = s.face_type(q)
        F.type[q] = t
        host = mgp.vp()
        a = mgp.Array()
        tt = ctypes.c_int()
        mgp.check(L.mgp_solver_face(s.h, 0, q, ctypes.byref(host), ctypes.byref(a), ctypes.byref(tt)))
        F.g[q] = a.data if t == mgp.NEUMANN else None
    F.h = h
    box = [0, n, 0, n, 1, nz - 1]
    box_pts = n * n * (nz - 2)
    bx = mgp.box(box)
    scal = mgp.DeviceBuffer(__import__("numpy").zeros(1))
    # u -> r and back keeps every access inside the level's arrays; the values
    # are not used afterwards
    sweep_ms = kernel_time_ms(
        lambda st: mgp.check(L.mgp_rbgs_sweep_f64(mgp.CONSISTENT, ctypes.byref(u), ctypes.byref(r), ctypes.byref(f_),
                                                   ctypes.byref(F), bx, 1.0, st)), args.kernel_reps, stream)
    pair_ms = kernel_time_ms(
        lambda st: mgp.check(L.mgp_rbgs_sweep2_f64(mgp.CONSISTENT, ctypes.byref(u), ctypes.byref(r),
                                                    ctypes.byref(f_), ctypes.byref(F), bx, 1.0, st)),
        args.kernel_reps, stream)
    pro_ms = kernel_time_ms(
        lambda st: mgp.check(L.mgp_rbgs_sweep2_prolong_f64(mgp.CONSISTENT, ctypes.byref(u), ctypes.byref(r),
                                                            ctypes.byref(f_), ctypes.byref(F), bx, 1.0,
                                                            ctypes.byref(u1), st)), args.kernel_reps, stream)
    resid_ms = kernel_time_ms(
        lambda st: mgp.check(L.mgp_residual_f64(mgp.CONSISTENT, ctypes.byref(u), ctypes.byref(f_), ctypes.byref(r),
                                                 bx, h, scal.ptr, st)), args.kernel_reps, stream)
    L.mgp_stream_sync(stream)
    sweep_bytes = 24.0 * box_pts
    pro_bytes = 24.0 * box_pts + 8.0 * u1.shape[0] * u1.shape[1] * u1.shape[2]
    resid_bytes = 24.0 * box_pts
    sweep_gbs = sweep_bytes / (sweep_ms * 1e-3) / 1e9
    pair_gbs = sweep_bytes / (pair_ms * 1e-3) / 1e9
    pro_gbs = pro_bytes / (pro_ms * 1e-3) / 1e9
    resid_gbs = resid_bytes / (resid_ms * 1e-3) / 1e9
    log("bench: sweep %.2f ms (%.0f GB/s), two-sweep pass %.2f ms (%.0f GB/s), with the correction %.2f ms "
        "(%.0f GB/s), residual %.2f ms (%.0f GB/s)" %
        (sweep_ms, sweep_gbs, pair_ms, pair_gbs, pro_ms, pro_gbs, resid_ms, resid_gbs))
    f32_kernels = fp32_pass_times(u, u1, F, bx, box_pts, stream, args.kernel_reps)
    # fine-level launches per F-cycle: three smoothing passes, the two after a
    # correction with the correction inside where sweep2.prolong is on
    fused = ctypes.c_int64(0)
    mgp.check(L.mgp_tuning_get(b"sweep2.prolong", ctypes.byref(fused)))
    n_pro = 2 if fused.value >= 1 else 0
    pair_kernel = {"kernel": "rbgs_sweep2 (fine level, two sweeps per pass)", "achieved": pair_gbs,
                   "bytes_per_launch": sweep_bytes, "ms_per_launch": pair_ms, "launches_per_cycle": 3 - n_pro,
                   "traffic": measured_traffic(n, "rbgs_sweep2")}
    pro_kernel = {"kernel": "rbgs_sweep2_prolong (fine level, correction u += P(u_coarse) and two sweeps per pass)",
                  "achieved": pro_gbs, "bytes_per_launch": pro_bytes, "ms_per_launch": pro_ms,
                  "launches_per_cycle": n_pro, "traffic": measured_traffic(n, "rbgs_sweep2_prolong")}
    top = max((pair_kernel, pro_kernel), key=lambda k: k["ms_per_launch"] * k["launches_per_cycle"])

    s.close()
    mixed = guarded_leg("mixed", args.no_mixed, lambda: mixed_leg(dims, h, args))
    batch = guarded_leg("batch", args.no_batch, lambda: batch_leg(args))
    serve = guarded_leg("serve", args.no_serve, lambda: serve_leg(args))

    cpu = port = None
    if not args.no_cpu_baseline:
        port = cpu_baseline(dims, args.levels, h, args.nsmooth, cycles[-1], args.cpu_threads, args.coarse_sweeps)
        cpu = guarded_leg("cpu reference", False,
                          lambda: cpu_baseline_reference(dims, args.levels, args.nsmooth, cycles[-1], args.cpu_threads))
        if not cpu or "error" in cpu:
            cpu, port = port, cpu

    out = {
        "metric": "DOF/s to ||r||/||f|| <= 1e-10 (plain FP64 F-cycle, 1025x1025x513)",
        "value": value,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic u = x^3 y^2 z)",
        "config": {"workload": "poisson-fcycle-%dx%dx%d" % dims, "grid": list(dims), "levels": args.levels,
                   "cycle": "F", "smoother": "red-black GS SOR omega=%g, %d sweeps" % (args.omega, args.nsmooth),
                   "restriction": "full weighting", "bc": "Neumann x,y / Dirichlet z", "tolerance": args.tol,
                   "coarsest": "CG (FP64, one workgroup)" if args.coarse_cg else "%d GS passes" % args.coarse_sweeps,
                   "cycles_to_tol": cycles[-1], "final_rel_res": rel[-1], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": top["kernel"], "achieved": top["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": top["achieved"] / HBM_PEAK_GBS, "traffic": top["traffic"],
                     "bytes_per_launch": top["bytes_per_launch"], "ms_per_launch": top["ms_per_launch"],
                     "launches_per_cycle": top["launches_per_cycle"]},
        "kernels": {"rbgs_sweep2": {"ms": pair_ms, "GB/s": pair_gbs, "frac": pair_gbs / HBM_PEAK_GBS,
                                    "launches_per_cycle": pair_kernel["launches_per_cycle"]},
                    "rbgs_sweep2_prolong": {"ms": pro_ms, "GB/s": pro_gbs, "frac": pro_gbs / HBM_PEAK_GBS,
                                            "launches_per_cycle": n_pro},
                    "rbgs_sweep": {"ms": sweep_ms, "GB/s": sweep_gbs, "frac": sweep_gbs / HBM_PEAK_GBS},
                    "residual_norm": {"ms": resid_ms, "GB/s": resid_gbs, "frac": resid_gbs / HBM_PEAK_GBS},
                    **f32_kernels},
        "fp32_refine": mixed,
        "batched": batch,
        "serving": serve,
        "cpu_baseline": cpu,
        "cpu_port": port,
    }
    print(json.dumps(out), flush=True)


def fp32_pass_times(u, u1, F, bx, box_pts, stream, reps):
    """The FP32 two-sweep passes of the mixed-precision cycles on arrays of the
    fine level's layout (u, f, an output and level 1's u, FP32, homogeneous
    Neumann data like the correction problem's): the plain pass (12 B per box
    point) and the pass with the correction inside (12 B per box point + 4 B per
    level-1 point), each timed like the FP64 kernels."""
    import numpy as np
    L = mgp.lib()
    shape = [u.shape[q] for q in range(3)]
    halo = [u.halo[q] for q in range(6)]
    cshape = [u1.shape[q] for q in range(3)]
    chalo = [u1.halo[q] for q in range(6)]
    a, b, f = (mgp.DeviceArray(shape, halo, np.float32) for _ in range(3))
    c = mgp.DeviceArray(cshape, chalo, np.float32)
    for arr, v in ((a, 0.25), (b, 0.0), (f, 1.0), (c, 0.0)):
        mgp.check(L.mgp_fill_f32(arr.ref(), v, None))
    F32 = mgp.Faces()
    keep = []
    for q in range(6):
        F32.type[q] = F.type[q]
        if F.type[q] == mgp.NEUMANN:
            m = list(shape)
            m[{0: 1, 1: 1, 2: 0, 3: 0, 4: 2, 5: 2}[q]] = 1
            g = mgp.DeviceBuffer(np.zeros(int(np.prod(m)), np.float32))
            keep.append(g)
            F32.g[q] = g.ptr
    F32.h = F.h
    R = ctypes.byref
    plain = kernel_time_ms(lambda st: mgp.check(L.mgp_rbgs_sweep2_f32(mgp.CONSISTENT, a.ref(), b.ref(), f.ref(),
                                                                       R(F32), bx, 1.0, st)), reps, stream)
    fused = kernel_time_ms(lambda st: mgp.check(L.mgp_rbgs_sweep2_prolong_f32(mgp.CONSISTENT, a.ref(), b.ref(),
                                                                               f.ref(), R(F32), bx, 1.0, c.ref(),
                                                                               st)), reps, stream)
    L.mgp_stream_sync(stream)
    cpts = cshape[0] * cshape[1] * cshape[2]
    out = {}
    for name, ms, nbytes in (("rbgs_sweep2_f32", plain, 12.0 * box_pts),
                             ("rbgs_sweep2_prolong_f32", fused, 12.0 * box_pts + 4.0 * cpts)):
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[name] = {"ms": ms, "GB/s": gbs, "frac": gbs / HBM_PEAK_GBS}
    log("bench: FP32 two-sweep pass %.2f ms, with the correction %.2f ms" % (plain, fused))
    for arr in (a, b, f, c):
        arr.free()
    return out


def measured_traffic(n, kernel):
    """HBM bytes per fine-level launch of `kernel` (rbgs_sweep2 or
    rbgs_sweep2_prolong) from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    summarised in profiles/ (scaled by calibration kernels of known size), for
    the 1025x1025x513 workload; None for other sizes or without a summary."""
    path = os.path.join(ROOT, "profiles", "%s_%dx%dx%d_traffic.md" % (kernel, n, n, n // 2 + 1))
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def solver_flags(args):
    return (mgp.QUIET | (mgp.NO_GRAPHS if args.no_graphs else 0) |
            (mgp.COARSE_CG if args.coarse_cg else 0))


def timed_solves(s, args):
    """Warmup solves, then args.steps timed solves from u = 0; returns (seconds,
    cycles per solve, final relative residuals)."""
    for _ in range(args.warmup):
        s.reset()
        s.solve("Fcycle", args.nsmooth, args.omega)
    s.sync()
    cycles, rel = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.reset()
        it, rr, _t = s.solve("Fcycle", args.nsmooth, args.omega)
        cycles.append(it + 1)
        rel.append(rr)
    s.sync()
    return time.perf_counter() - t0, cycles, rel


def interior(a, shape, halo):
    """(nx, ny, nz) view of a halo-padded flat host array."""
    hN, hS, hE, hW, hT, hB = halo
    p = mgp.padded_shape(shape, halo)
    return a.reshape(p)[hW:hW + shape[0], hS:hS + shape[1], hB:hB + shape[2]]


def mixed_leg(dims, h, args):
    """FP32 cycles inside FP64 refinement vs plain FP64, same grid, levels,
    F-cycle and coarsest passes.  The all-Dirichlet problem
    u = cos(x z^2) sin(y^3) at tol 1e-10 and at 2e-10 (also with levels >= 3
    of the FP32 cycle stored in FP16), and the Neumann-x/y problem
    u = x^3 y^2 z at 1e-10; max-abs differences of the all-Dirichlet
    solutions from each other and from the analytic u."""
    import numpy as np

    dof = dims[0] * dims[1] * dims[2]
    out = {"workload": "poisson-fcycle-mixed-%dx%dx%d" % tuple(dims)}
    sols = {}
    cases = (("cos_dirichlet", 3, "DDDDDD", 1e-10), ("cubic_neumann_xy", 2, "NNNNDD", 1e-10))
    for label, prob, bc, tol in cases:
        key = "%s_tol%.0e" % (label, tol)
        out[key] = {}
        variants = [("fp64", mgp.F64, 0), ("fp32_refine", mgp.MIXED, 0)]
        if bc == "DDDDDD":  # FP16 storage of levels >= 3 (Dirichlet faces only)
            variants.append(("fp32_refine_fp16_coarse", mgp.MIXED, 3))
        for name, dt, half_from in variants:
            log("bench: mixed leg, %s %s" % (key, name))
            s = mgp.Solver(dims, args.levels, h, (0.0, 0.0, 0.0), bc, mgp.CONSISTENT, mgp.GAUSS_SEIDEL,
                           mgp.FULL_WEIGHTING, problem=prob, dtype=dt,
                           flags=solver_flags(args), maxiter=30, miniter=1,
                           tolerance=tol, coarse_sweeps=args.coarse_sweeps, half_from=half_from)
            s.to_device()
            el, cyc, rel = timed_solves(s, args)
            out[key][name] = {"DOF/s": dof * args.steps / el, "ms_per_solve": 1e3 * el / args.steps,
                              "iterations": cyc[-1], "final_rel_res": rel[-1], "reached_tol": rel[-1] <= tol}
            log("bench:   %.1f ms/solve, %d iterations, rel_res %.3e" % (1e3 * el / args.steps, cyc[-1], rel[-1]))
            s.to_host()
            host, a = s._arr(0, "u")
            sols[(label, name)] = np.array(interior(s.array("u"), list(a.shape), list(a.halo)), copy=True)
            s.close()
        o = out[key]
        o["speedup"] = o["fp32_refine"]["DOF/s"] / o["fp64"]["DOF/s"]
        o["ms_per_iteration"] = {n: o[n]["ms_per_solve"] / max(o[n]["iterations"], 1) for n, _, _ in variants}
        if len(variants) > 2:
            o["speedup_fp16_coarse"] = o["fp32_refine_fp16_coarse"]["DOF/s"] / o["fp64"]["DOF/s"]
    x = np.arange(dims[0]) * h
    y = np.arange(dims[1]) * h
    z = np.arange(dims[2]) * h

    def exact(label, i0, i1):
        xs = x[i0:i1][:, None, None]
        if label == "cos_dirichlet":
            return np.cos(xs * (z ** 2)[None, None, :]) * np.sin(y ** 3)[None, :, None]
        return xs ** 3 * (y ** 2)[None, :, None] * z[None, None, :]

    for label, _, _, _ in cases:
        names = [n for (lb, n) in sols if lb == label]
        ref = sols[(label, "fp64")]
        diff = {n: 0.0 for n in names if n != "fp64"}
        err = {n: 0.0 for n in names}
        for i0 in range(0, dims[0], 64):
            ua = exact(label, i0, i0 + 64)
            for n in names:
                u = sols[(label, n)][i0:i0 + 64]
                err[n] = max(err[n], float(np.max(np.abs(u - ua))))
                if n != "fp64":
                    diff[n] = max(diff[n], float(np.max(np.abs(u - ref[i0:i0 + 64]))))
        out["%s_max_abs_diff_vs_fp64" % label] = diff
        out["%s_max_abs_err_vs_analytic" % label] = err
        log("bench: mixed %s: |u - u_fp64| %s, errors vs analytic %s" % (label, diff, err))
    return out


def traffic_probe(args):
    """Kernels for HBM-counter runs, each launched once on the headline grid:
    calibration kernels of known size -- a fill of r (8 B/point written), an
    inf-norm of f (8 B/point read), r += f (mgp_add_f64: 16 B/point read,
    8 B/point written) -- then the fine-level operators of the cycle: the
    one-pass sweep u -> r and the two-sweep pass u -> r (24 B per box point),
    the residual with its norm (24 B), the norm-only residual (16 B), the fused
    residual + full weighting into level 1's f (16 B per fine point read,
    1 B written), the prolongation + correction from level 1 (16 B per fine
    point, 1 B read), and the hi + lo update + residual of mixed refinement
    (40 B per point: u, f 8 B each, lo, e 4 B each read; u, lo, rhs written)."""
    import numpy as np

    n, nz = args.n, args.n // 2 + 1
    h = 1.0 / (n - 1)
    s = mgp.Solver((n, n, nz), args.levels, h, (0.0, 0.0, 0.0), "NNNNDD", mgp.CONSISTENT, mgp.GAUSS_SEIDEL,
                   mgp.FULL_WEIGHTING, problem=2, flags=mgp.QUIET | mgp.NO_GRAPHS, coarse_sweeps=args.coarse_sweeps)
    s.to_device()
    L = mgp.lib()
    u, f_, r = s.device_array("u"), s.device_array("f"), s.device_array("r")
    u1, f1 = s.device_array("u", 1), s.device_array("f", 1)
    scal = mgp.DeviceBuffer(np.zeros(1))
    F = mgp.Faces()
    for q in range(6):
        t = s.face_type(q)
        F.type[q] = t
        host, a, tt = mgp.vp(), mgp.Array(), ctypes.c_int()
        mgp.check(L.mgp_solver_face(s.h, 0, q, ctypes.byref(host), ctypes.byref(a), ctypes.byref(tt)))
        F.g[q] = a.data if t == mgp.NEUMANN else None
    F.h = h
    L.mgp_solver_sync(s.h)
    shape, halo = [u.shape[0], u.shape[1], u.shape[2]], [u.halo[q] for q in range(6)]
    lo, ln, e, out = (mgp.DeviceArray(shape, halo, np.float32) for _ in range(4))
    un = mgp.DeviceArray(shape, halo)
    for a in (lo, ln, e, out):
        mgp.check(L.mgp_fill_f32(a.ref(), 0.0, None))
    box = mgp.box([0, n, 0, n, 1, nz - 1])
    R = ctypes.byref
    mgp.check(L.mgp_fill_f64(R(r), 0.0, None))
    mgp.check(L.mgp_norm_inf_f64(R(f_), scal.ptr, None))
    mgp.check(L.mgp_add_f64(R(r), R(f_), None))
    mgp.check(L.mgp_rbgs_sweep_f64(mgp.CONSISTENT, R(u), R(r), R(f_), R(F), box, 1.0, None))
    mgp.check(L.mgp_rbgs_sweep2_f64(mgp.CONSISTENT, R(u), R(r), R(f_), R(F), box, 1.0, None))
    mgp.check(L.mgp_rbgs_sweep2_prolong_f64(mgp.CONSISTENT, R(u), R(r), R(f_), R(F), box, 1.0, R(u1), None))
    mgp.check(L.mgp_residual_f64(mgp.CONSISTENT, R(u), R(f_), R(r), box, h, scal.ptr, None))
    mgp.check(L.mgp_residual_norm_f64(mgp.CONSISTENT, R(u), R(f_), R(r), box, h, scal.ptr, None))
    mgp.check(L.mgp_residual_restrict_fw_f64(mgp.CONSISTENT, R(u), R(f_), R(r), R(f1), box, h, None))
    mgp.check(L.mgp_prolong_add_f64(R(u1), R(u), None))
    mgp.check(L.mgp_update_residual_hilo_f64_f32(R(u), un.ref(), lo.ref(), ln.ref(), e.ref(), R(f_), R(r), out.ref(),
                                                 R(F), box, scal.ptr, None))
    L.mgp_stream_sync(None)
    pts = n * n * nz
    bpts = n * n * (nz - 2)
    cpts = u1.shape[0] * u1.shape[1] * u1.shape[2]
    log("traffic probe (algorithmic bytes): fill %d, norm %d, add %d, sweep %d, two-sweep pass %d, two-sweep pass "
        "with the correction %d, residual %d, norm-only residual %d, residual + FW %d, prolongation + add %d, "
        "hi+lo update + residual %d" %
        (8 * pts, 8 * pts, 24 * pts, 24 * bpts, 24 * bpts, 24 * bpts + 8 * cpts, 24 * bpts, 16 * bpts,
         16 * pts + 8 * pts // 8, 16 * pts + 8 * pts // 8, 40 * pts))
    s.close()


def probe_mixed(args, problem):
    """One warm-up and one measured FP32-refinement solve (launch by launch,
    for kernel traces) of the mixed leg's problem `problem` on the headline
    grid: 2 = u = x^3 y^2 z with Neumann x/y faces, 3 = cos(x z^2) sin(y^3)
    with Dirichlet faces; no JSON line."""
    n, nz = args.n, args.n // 2 + 1
    h = 1.0 / (n - 1)
    bc = "NNNNDD" if problem == 2 else "DDDDDD"
    s = mgp.Solver((n, n, nz), args.levels, h, (0.0, 0.0, 0.0), bc, mgp.CONSISTENT, mgp.GAUSS_SEIDEL,
                   mgp.FULL_WEIGHTING, problem=problem, dtype=mgp.MIXED, flags=mgp.QUIET | mgp.NO_GRAPHS,
                   maxiter=30, miniter=1, tolerance=args.tol, coarse_sweeps=args.coarse_sweeps)
    s.to_device()
    for k in range(2):
        s.reset()
        s.sync()
        t0 = time.perf_counter()
        it, rel, _ = s.solve("Fcycle", args.nsmooth, 1.0)
        log("probe mixed p%d solve %d: %.1f ms, %d iterations, rel_res %.3e" %
            (problem, k, 1e3 * (time.perf_counter() - t0), it, rel))
    s.close()


def probe_batch(args):
    """The batched leg's solve (same problems and settings as batch_leg), one
    warm-up and one measured solve, launch by launch (for kernel traces and
    counters); no JSON line."""
    import numpy as np

    rng = np.random.default_rng(args.seed + 4)
    n = args.batch_count
    dims = [tuple(int(v) for v in rng.choice([33, 65, 129], size=3)) for _ in range(n)]
    modes = [modes_of(rng) for _ in range(n)]
    b = mgp.Batch(dims, relaxation=mgp.JACOBI, nsmooth=3, omega=6.0 / 7.0, maxiter=60, tolerance=args.tol,
                  flags=mgp.BATCH_NO_GRAPHS)
    for p in range(n):
        b.set_rhs(p, modes_rhs(dims[p], modes[p]))
    for k in range(2):
        t = b.solve()
        st = [b.stats(p) for p in range(n)]
        log("probe batch solve %d: %.1f ms, iterations %d..%d, worst rel_res %.2e" %
            (k, 1e3 * t, min(i for i, _ in st), max(i for i, _ in st), max(r for _, r in st)))
    b.close()


def guarded_leg(name, skip, fn):
    """A secondary leg; its failure is reported in the line, not fatal."""
    if skip:
        return None
    try:
        return fn()
    except Exception as e:  # noqa: BLE001
        log("bench: %s leg failed: %r" % (name, e))
        return {"error": repr(e)}


def max_levels(d):
    L = 1
    while all(((n - 1) >> (L - 1)) % 2 == 0 and ((n - 1) >> (L - 1)) // 2 >= 2 for n in d):
        L += 1
    return L


def modes_of(rng, count=4):
    return [(int(a), int(b), int(c), float(rng.standard_normal())) for a, b, c in rng.integers(1, 9, size=(count, 3))]


def modes_rhs(d, modes):
    """Host f = sum A sin(a pi x) sin(b pi y) sin(c pi z), x = i/(nx-1), separable."""
    import numpy as np

    h = 1.0 / (d[0] - 1)
    x, y, z = (np.arange(n) * h for n in d)
    f = np.zeros(d)
    for a, b, c, amp in modes:
        f += amp * np.sin(a * np.pi * x)[:, None, None] * (np.sin(b * np.pi * y)[:, None] * np.sin(c * np.pi * z))
    return f


def batch_leg(args):
    """Batched FP32-cycle / FP64-refinement solves (256 problems by default; Jacobi 6/7, 3+3,
    V-cycle, FW, trilinear, zero Dirichlet) vs the same problems solved one at
    a time (mixed and plain FP64, hierarchies resident, a timed sample)."""
    import numpy as np

    rng = np.random.default_rng(args.seed + 4)
    n = args.batch_count
    dims = [tuple(int(v) for v in rng.choice([33, 65, 129], size=3)) for _ in range(n)]
    modes = [modes_of(rng) for _ in range(n)]
    dof = sum(a * b * c for a, b, c in dims)
    log("bench: batch leg, %d problems, %.3g DOF" % (n, dof))
    b = mgp.Batch(dims, relaxation=mgp.JACOBI, nsmooth=3, omega=6.0 / 7.0, maxiter=60, tolerance=args.tol)
    for p in range(n):
        b.set_rhs(p, modes_rhs(dims[p], modes[p]))
    b.solve()  # records the iteration graphs
    t = b.solve()
    st = [b.stats(p) for p in range(n)]
    worst = max(r for _, r in st)
    log("bench:   batched %.1f ms, iterations %d..%d, worst rel_res %.2e" %
        (1e3 * t, min(i for i, _ in st), max(i for i, _ in st), worst))
    b.close()
    # one-at-a-time solves of a sample of the same problems
    k = min(n, 24)
    seq = {}
    for name, dt in (("fp32_refine_single", mgp.MIXED), ("fp64_single", mgp.F64)):
        tot = 0.0
        for p in range(k):
            d = dims[p]
            s = mgp.Solver(d, max_levels(d), 1.0 / (d[0] - 1), (0.0, 0.0, 0.0), "DDDDDD", mgp.CONSISTENT, mgp.JACOBI,
                           mgp.FULL_WEIGHTING, problem=-1, dtype=dt, maxiter=60, miniter=1, tolerance=args.tol)
            s.array("f")[:] = modes_rhs(d, modes[p]).ravel()
            s.to_device()
            s.solve("Vcycle", 3, 6.0 / 7.0)  # records the cycle graph
            s.reset()
            s.sync()
            t0 = time.perf_counter()
            s.solve("Vcycle", 3, 6.0 / 7.0)
            s.sync()
            tot += time.perf_counter() - t0
            s.close()
        seq[name] = {"solves/s": k / tot, "sample": "%d of %d problems, solve() only" % (k, n)}
    return {"workload": "batched-%d-dirichlet-{33,65,129}^3" % n, "problems": n, "dof": dof, "seconds": t,
            "solves/s": n / t, "DOF/s": dof / t, "worst_rel_res": worst,
            "iterations": [min(i for i, _ in st), max(i for i, _ in st)], "one_at_a_time": seq,
            "speedup_vs_fp64_single": (n / t) / seq["fp64_single"]["solves/s"]}


def serve_requests(rng, count):
    out = []
    while len(out) < count:
        d = tuple(int(2 ** k + 1) for k in rng.integers(5, 10, size=3))
        bc = "".join("N" if v else "D" for v in rng.integers(0, 2, size=6))
        if "D" not in bc:  # all-Neumann is singular: redrawn
            continue
        out.append((d, bc, modes_of(rng)))
    return out


def serve_leg(args):
    """The serving loop (FP32-cycle / FP64-refinement F-cycles, tol 1e-8) on
    requests of 2^k+1 points per axis (k in 5..9), face types by seed, Fourier
    right-hand sides.  Against one-at-a-time plain FP64 F-cycle solves of a
    sample of the same requests (fresh hierarchy each, as the server's
    requests of new shapes get): the burst capacity over all requests, then
    Poisson arrivals offered at args.serve_load x the one-at-a-time rate,
    with the delivered rate and p50 / p99 latency over all of them.  Both
    solve the coarsest level by CG (the server always; the one-at-a-time
    solver with MGP_SOLVER_COARSE_CG)."""
    import numpy as np

    rng = np.random.default_rng(args.seed + 5)
    reqs = serve_requests(rng, args.serve_count)
    tol = 1e-8
    L = mgp.lib()
    # one-at-a-time FP64 on a sample, with and without cycle graphs (the
    # faster counts)
    sample = reqs[:min(len(reqs), args.serve_sample)]
    log("bench: serving leg, %d requests; one-at-a-time FP64 on %d" % (len(reqs), len(sample)))
    seq = {}
    for name, gflag in (("graphs", 0), ("no_graphs", mgp.NO_GRAPHS)):
        t0 = time.perf_counter()
        for d, bc, modes in sample:
            s = mgp.Solver(d, max_levels(d), 1.0 / (d[0] - 1), (0.0, 0.0, 0.0), bc, mgp.CONSISTENT,
                           mgp.GAUSS_SEIDEL, mgp.FULL_WEIGHTING, problem=None, dtype=mgp.F64,
                           flags=mgp.QUIET | mgp.COARSE_CG | gflag, maxiter=100, miniter=1, tolerance=tol,
                           coarse_sweeps=32)
            mgp.check(L.mgp_solver_to_device_zero(s.h))
            fa = s.device_array("f")
            abc = (ctypes.c_int32 * 48)(*[v for a, b_, c, _ in modes for v in (a, b_, c)])
            amp = (ctypes.c_double * 16)(*[m[3] for m in modes])
            mgp.check(L.mgp_fill_modes_f64(ctypes.byref(fa), len(modes), abc, amp, 1.0 / (d[0] - 1), None))
            it, rr, _ = s.solve("Fcycle", 2, 1.0)
            ua = s.device_array("u")
            host = np.empty(int(np.prod(mgp.padded_shape(list(ua.shape), list(ua.halo)))))
            mgp.check(L.mgp_memcpy_d2h(host.ctypes.data, ua.data, host.nbytes, None))
            s.close()
        seq[name] = len(sample) / (time.perf_counter() - t0)
        log("bench:   one at a time (%s): %.1f solves/s" % (name, seq[name]))
    seq_rate = max(seq.values())

    def run(arrival_rate):
        srv = mgp.Server(hbm_budget=96 << 30, max_active=args.serve_workers, tolerance=tol, coarse_sweeps=32)
        start = time.perf_counter()
        ids = []
        t_next = start
        arr = np.random.default_rng(args.seed + 55)
        for d, bc, modes in reqs:
            if arrival_rate:
                t_next += arr.exponential(1.0 / arrival_rate)
                dt = t_next - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
            ids.append(srv.submit(d, bc, modes, want_solution=True))
        offered_span = time.perf_counter() - start
        res = [srv.wait(i)[0] for i in ids]
        wall = time.perf_counter() - start
        _, peak, resident = srv.stats()
        srv.close()
        lat = np.array([r.latency_s for r in res])
        bad = [r for r in res if r.status != 0 or r.rel_res > tol]
        return {"requests": len(res), "solves/s": len(reqs) / wall, "offered/s": len(reqs) / offered_span,
                "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                "max_latency_s": float(lat.max()), "peak_active": peak, "failed": len(bad),
                "mean_setup_s": float(np.mean([r.setup_s for r in res])),
                "mean_run_s": float(np.mean([r.run_s for r in res])),
                "mean_fetch_s": float(np.mean([r.fetch_s for r in res])),
                "mean_queue_s": float(np.mean([r.queue_s for r in res]))}

    log("bench:   server burst")
    burst = run(None)
    log("bench:   burst %.1f solves/s" % burst["solves/s"])
    rate = args.serve_load * seq_rate
    log("bench:   server at Poisson arrivals %.1f/s (%.1f x one-at-a-time FP64)" % (rate, args.serve_load))
    poisson = run(rate)
    poisson["arrival_rate"] = rate
    log("bench:   delivered %.1f solves/s, p50 %.3f s, p99 %.3f s" %
        (poisson["solves/s"], poisson["p50_latency_s"], poisson["p99_latency_s"]))
    return {"workload": "serving-%d-requests-2^k+1-mixed-bc" % len(reqs), "tolerance": tol,
            "workers": args.serve_workers,
            "one_at_a_time_fp64_solves/s": seq_rate,
            "one_at_a_time_fp64_by_launch_mode": seq,
            "one_at_a_time_sample": "%d of %d requests" % (len(sample), len(reqs)),
            "burst": burst, "poisson": poisson,
            "capacity_vs_fp64_one_at_a_time": burst["solves/s"] / seq_rate,
            "delivered_vs_fp64_one_at_a_time": poisson["solves/s"] / seq_rate}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_reference(params, threads, timeout):
    """oracle/_ref/ref_timing (the reference's own sources, built with its own
    host flags -O3 -ffast-math -fopenmp) in timing mode; returns the per-cycle
    times the reference prints and its set-up / solve wall times."""
    import subprocess
    import tempfile

    exe = os.path.join(ROOT, "oracle", "_ref", "ref_timing")
    with tempfile.TemporaryDirectory() as tmp:
        par = os.path.join(tmp, "p.txt")
        with open(par, "w") as fp:
            fp.write("".join("%s %s\n" % kv for kv in params.items()))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
        out = subprocess.run([exe, par, tmp, tmp], capture_output=True, text=True, timeout=timeout, env=env,
                             check=True).stdout
    cyc = [float(l.split()[2]) for l in out.splitlines() if l.endswith(" seconds") and " took " in l]
    tail = [l for l in out.splitlines() if l.startswith("harness ")][-1].split()
    kv = dict(zip(tail[1::2], tail[2::2]))
    return cyc, float(kv["setup_s"]), float(kv["solve_s"]), float(kv["rel_res"])


def cpu_baseline_reference(dims, levels, nsmooth, cycles, threads):
    """The reference's own CPU path timed on the host cores (OpenMP target
    regions run on the host): one F-cycle of the headline workload's shape
    (1025x1025x513, 7 levels, red-black GS omega 1, 2 sweeps, full weighting,
    Neumann x/y and Dirichlet z requested, u = x^3 y^2 z) in the reference's own
    boundary treatment (north/south and bottom forced Dirichlet, stalls at
    O(h): it never reaches 1e-10), scaled to the GPU's cycle count; plus the
    10 V-cycles of the 65^3 configuration in full."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_timing")
    if not os.path.exists(exe):
        return None
    if threads <= 0:
        threads = len(os.sched_getaffinity(0))
    log("bench: CPU baseline (reference, %d threads): 65^3 V-cycles" % threads)
    c1 = dict(mode="time", nx=65, ny=65, nz=65, levels=4, bc="DDDDDD", problem=0, restriction="inj",
              relaxation="gs", cycle="Vcycle", nsmooth=2, omega=1.0, ncycles=10, lengthx=1.0)
    cyc1, set1, sol1, rr1 = run_reference(c1, threads, 300)
    log("bench: CPU baseline (reference): %dx%dx%d F-cycle" % dims)
    c2 = dict(mode="time", nx=dims[0], ny=dims[1], nz=dims[2], levels=levels, bc="NNNNDD", problem=2,
              restriction="fw", relaxation="gs", cycle="Fcycle", nsmooth=nsmooth, omega=1.0, ncycles=1, lengthx=1.0)
    cyc2, set2, sol2, rr2 = run_reference(c2, threads, 900)
    spc = sum(cyc2) / len(cyc2)
    dof = dims[0] * dims[1] * dims[2]
    log("bench: CPU baseline (reference) %.2f s per F-cycle (set-up %.1f s)" % (spc, set2))
    return {"value": dof / (spc * cycles), "unit": "DOF/s", "cores": threads, "kind": "reference",
            "sample": "the reference's own F-cycle at %dx%dx%d (%d levels, GS(%d), FW), 1 cycle of %.1f s as it "
                      "prints it, scaled to the GPU's %d cycles to 1e-10; set-up %.1f s not counted"
                      % (dims + (levels, nsmooth, spc, cycles, set2)),
            "s_per_cycle": spc, "cpu_model": cpu_model(),
            "flags": "g++ -O3 -ffast-math -fopenmp (the reference's config/g++.mkf host flags)",
            "config_65cube_vcycles": {"cycles": len(cyc1), "s_per_cycle": sum(cyc1) / max(len(cyc1), 1),
                                      "solve_s": sol1, "final_rel_res": rr1,
                                      "DOF*cycles/s": 65 ** 3 * len(cyc1) / max(sol1, 1e-9)}}


def cpu_baseline(dims, levels, h, nsmooth, cycles, threads, coarse_sweeps):
    """The oracle (this build's C restatement, OpenMP) on the host cores: one
    F-cycle plus the residual check at the full grid, scaled to the GPU's
    cycle count to the tolerance."""
    import oracle_lib as O

    if threads <= 0:
        threads = len(os.sched_getaffinity(0))
    log("bench: CPU baseline (oracle, %d threads): setup" % threads)
    o = O.Solver(dims, levels, h, (0.0, 0.0, 0.0), "NNNNDD", O.CONSISTENT, O.GAUSS_SEIDEL, O.FULL_WEIGHTING, 2,
                 threads=threads, coarse_sweeps=coarse_sweeps)
    o.to_device()
    t0 = time.perf_counter()
    o.solve(O.FCYCLE, nsmooth, 1.0, maxiter=1, miniter=1, tol=0.0)
    t = time.perf_counter() - t0
    o.close()
    dof = dims[0] * dims[1] * dims[2]
    log("bench: CPU baseline %.2f s per F-cycle" % t)
    return {"value": dof / (t * cycles), "unit": "DOF/s", "cores": threads, "kind": "port",
            "sample": "1 F-cycle + residual at %dx%dx%d (%.1f s), scaled to %d cycles" % (dims + (t, cycles)),
            "s_per_cycle": t, "cpu_model": cpu_model()}


if __name__ == "__main__":
    main()
