"""Benchmark: FP64 projected-Jacobi contact updates per second on one MI355X, plus the batched and
differentiable forms.

Headline (`value`): the large single scene -- 100k spheres, radius log-uniform in [0.02, 0.08] m,
density 1000 kg/m^3, random packing settled in a 4x4x8 m box, mu ~ U[0.3, 0.8] per body
(min-combined), ~4e5 frictional contacts, h = 5e-4 s.  One step = one full projected-Jacobi solve
of that contact problem from a cold start (no warm start) to the reference residual test with
tol_abs = tol_rel = 1e-8 (checked every 10 sweeps), everything resident in HBM.
contact-updates/s = contacts x sweeps / time.  Timed with chunks of sweeps replayed as HIP graphs
(the production path).

`roofline`: the whole sweep (range pass + body pass + the residual every 10th sweep): the
algorithmic bytes of every sweep over the timed solves' stream-event time; `kernels` gives each
kernel's average duration from start/stop HIP events carried in every dispatch on the solve's own
stream during one more step of the same workload right after the timed ones; `traffic` is the HBM
bytes per sweep from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary when it was taken on
this workload and sweep form; `copy_GBs` is a device-to-device copy measured in the same run (a
second denominator).  The pile's contact points go to the library (cdx_set_contact_points), so the
range pass rebuilds the angular rows instead of reading them (--no-geometry-rows: read them).

`cpu_baseline`: the CPU oracle (a C restatement of the reference PJ sweep: single-thread SpMV as
the reference pins Eigen to one thread, OpenMP projection) on a bounded sample of the same
workload, and a fully threaded variant as an upper bound.

`colored_gs`: the coloured Gauss-Seidel solve of the same pile to the same tolerance (residual
every sweep).  `batched`: 4096 independent sphere/box scenes dropped into a unit box, one batched
solve (one launch) per Moreau step, against the same scenes solved one by one through the plain
solver on sampled steps.  `differentiable`: 256 scenes of 64 boxes on a plane, forward and
backward of a 100-step rollout (step states kept by the forward), against the same scenes rolled out one by one.

--copy-mb N also runs one device-to-device copy of N MiB, the calibration kernel for the
FETCH_SIZE / WRITE_SIZE counters collected under rocprofv3 --pmc.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B_CONTACT_DYN = 328  # algorithmic bytes per frictional contact-update, both sides dynamic
B_CONTACT_STATIC = 256  # ... with a static side
B_BODY = 128  # per body per sweep (6 u read + 6 u write + 4 M^-1)


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


def committed_traffic(contacts):
    """HBM bytes per projected-Jacobi sweep from the committed rocprofv3 counter summary (the sum
    over the sweep's kernels of FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by a device-copy
    calibration), if it was taken on a problem with the same contact count; None otherwise."""
    path = os.path.join(ROOT, "profiles", "pj_sweep_pile100k_hbm_traffic.txt")
    try:
        per_sweep, n = None, None
        for line in open(path):
            f = line.split()
            if len(f) == 2 and f[0] == "traffic_bytes_per_sweep":
                per_sweep = int(f[1])
            if len(f) == 2 and f[0] == "traffic_contacts":
                n = int(f[1])
        if per_sweep is not None and n == contacts:
            return per_sweep, os.path.relpath(path, ROOT)
    except (OSError, ValueError):
        pass
    return None, None


def algorithmic_bytes(prob):
    both = int(((prob.body_a >= 0) & (prob.body_b >= 0)).sum())
    one = prob.nc - both
    return both * B_CONTACT_DYN + one * B_CONTACT_STATIC, prob.nb * B_BODY


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class Hip:
    """The few HIP runtime calls the copy measurements need."""

    def __init__(self):
        self.l = ctypes.CDLL("libamdhip64.so")

    def malloc(self, n):
        p = ctypes.c_void_p()
        if self.l.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
        return p

    def event(self):
        e = ctypes.c_void_p()
        if self.l.hipEventCreate(ctypes.byref(e)) != 0:
            raise RuntimeError("hipEventCreate failed")
        return e

    def record(self, e, stream):
        if self.l.hipEventRecord(e, ctypes.c_void_p(stream or None)) != 0:
            raise RuntimeError("hipEventRecord failed")

    def elapsed_ms(self, e0, e1):
        self.l.hipEventSynchronize(e1)
        ms = ctypes.c_float(0)
        self.l.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        return float(ms.value)

    def destroy(self, *events):
        for e in events:
            self.l.hipEventDestroy(e)

    def copy_once(self, mb):
        n = mb << 20
        src, dst = self.malloc(n), self.malloc(n)
        self.l.hipMemset(src, 1, ctypes.c_size_t(n))
        self.l.hipMemcpy(dst, src, ctypes.c_size_t(n), 3)  # device to device
        self.l.hipDeviceSynchronize()
        self.l.hipFree(src)
        self.l.hipFree(dst)

    def copy_bandwidth(self, mb=1024, reps=10):
        """Device-to-device copy rate in GB/s, counting the bytes read and the bytes written."""
        n = mb << 20
        src, dst = self.malloc(n), self.malloc(n)
        self.l.hipMemset(src, 1, ctypes.c_size_t(n))
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        self.l.hipEventCreate(ctypes.byref(e0))
        self.l.hipEventCreate(ctypes.byref(e1))
        self.l.hipMemcpyAsync(dst, src, ctypes.c_size_t(n), 3, None)
        self.l.hipEventRecord(e0, None)
        for _ in range(reps):
            self.l.hipMemcpyAsync(dst, src, ctypes.c_size_t(n), 3, None)
        self.l.hipEventRecord(e1, None)
        self.l.hipEventSynchronize(e1)
        ms = ctypes.c_float(0)
        self.l.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        self.l.hipEventDestroy(e0)
        self.l.hipEventDestroy(e1)
        self.l.hipFree(src)
        self.l.hipFree(dst)
        return 2.0 * n * reps / (ms.value * 1e-3) / 1e9


# --------------------------------------------------------------------------------------------
# large single scene (headline)
# --------------------------------------------------------------------------------------------

def profiled_kernels(ctx, solve):
    """Runs `solve` once with start/stop HIP events carried in every kernel dispatch on the solve's
    stream (sweeps then launch one by one instead of as graph replays) and returns
    ({kernel: (calls, total ms)}, the solve's result)."""
    ctx.set_profiling(True)
    ctx.kernel_stats()
    for k in list(ctx.kernel_stats()):
        ctx.kernel_time(k)
    res = solve()
    ctx.synchronize()
    stats = ctx.kernel_stats()
    for k in stats:
        ctx.kernel_time(k)
    ctx.set_profiling(False)
    return stats, res


def colored_gs_leg(native, ctx, prob, dt, args, metrics, v, comp_pj, elapsed, bytes_sweep):
    """Coloured Gauss-Seidel on the pile already on the device, same tolerance (residual every
    sweep): time to solution with graph-replayed sweeps, then one more solve with kernel events."""
    pg = native.default_params()
    pg.tol_abs = pg.tol_rel = args.tol
    pg.max_iter = args.gs_max_iter
    pg.warmstart = 0
    ctx.synchronize()
    tg0 = time.perf_counter()
    ginfo = ctx.gs_colored_solve(dt, 1.0, pg, raise_on_divergence=False)
    ctx.synchronize()
    tg = time.perf_counter() - tg0
    vg = ctx.velocities()
    ig = ctx.impulses()
    ig[:, 1:] *= -1.0  # the Gauss-Seidel solver stores tangents with the opposite sign
    log(f"colored GS: {ginfo.iters} sweeps, {ginfo.ncolors} colours, {tg:.3f}s")
    kern = None
    if not args.no_kernel_events:
        stats, pinfo = profiled_kernels(ctx, lambda: ctx.gs_colored_solve(dt, 1.0, pg, raise_on_divergence=False))
        sw = max(pinfo.iters, 1)
        calls, ms = stats.get("gs_color", (0, 0.0))
        rc, rms = stats.get("gs_body_residual", (0, 0.0))
        fc, fms = stats.get("gs_residual", (0, 0.0))
        sweep_us = (ms + rms + fms) / sw * 1e3
        kern = {"gs_color_avg_us": ms / max(calls, 1) * 1e3, "colour_launches_per_sweep": calls / sw,
                "gs_body_residual_avg_us": rms / max(rc, 1) * 1e3, "gs_residual_avg_us": fms / max(fc, 1) * 1e3,
                "kernel_us_per_sweep": sweep_us, "wall_us_per_sweep": tg / max(ginfo.iters, 1) * 1e6,
                "achieved_GBs": bytes_sweep / (sweep_us * 1e-6) / 1e9,
                "frac": bytes_sweep / (sweep_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                "timing": "HIP events in every dispatch during one more solve of the same problem (direct launches)"}
        log(f"colored GS kernels: {kern['colour_launches_per_sweep']:.1f} colour launches of "
            f"{kern['gs_color_avg_us']:.2f} us per sweep, {sweep_us:.1f} us of kernels per sweep")
    return {"value": prob.nc * ginfo.iters / tg, "unit": "contact-updates/s", "sweeps": ginfo.iters,
            "colors": ginfo.ncolors, "residual": ginfo.residual, "seconds_to_solution": tg,
            "pj_seconds_to_solution": elapsed / args.steps, "complementarity": metrics.complementarity(prob, vg, ig),
            "pj_complementarity": comp_pj, "kinetic_energy": float(0.5 * (vg ** 2 / prob.minv).sum()),
            "pj_kinetic_energy": float(0.5 * (v ** 2 / prob.minv).sum()), "kernels": kern}


def cpu_threads():
    """Threads for the CPU baseline: the CPUs this process may run on, capped by OMP_NUM_THREADS
    when the host sets it (a GPU box's CPU share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(n, 1)


def pile_leg(args, native, scenes, metrics, hip):
    t0 = time.time()
    pile = scenes.sphere_pile(n=args.spheres, settle_steps=args.settle, seed=0)
    prob = pile.problem()
    log(f"scene: {prob.nb} bodies, {prob.nc} contacts, generated in {time.time() - t0:.1f}s")
    dt = 5e-4
    bc, bb = algorithmic_bytes(prob)

    ctx = native.Context(0)
    if args.no_graphs:
        ctx.set_graphs(False)
    ctx.set_range_sweep(args.range_mode, args.range_contacts, args.range_block)
    ctx.set_geometry_rows(0 if args.no_geometry_rows else 1)
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    ctx.synchronize()
    # host buffers in and out across the C ABI (PCIe), timed once; not part of `value`
    tu = time.perf_counter()
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    ctx.set_contact_points(prob.point)  # the geometry form of the rows, where they reproduce bit for bit
    ctx.synchronize()
    t_upload = time.perf_counter() - tu
    p = native.default_params()
    p.tol_abs = p.tol_rel = args.tol
    p.max_iter = args.max_iter
    p.warmstart = 0

    for _ in range(args.warmup):
        info = ctx.pj_solve(dt, 1.0, p)
        log(f"warmup solve: {info.iters} sweeps, residual {info.residual:.3e}")

    stream = ctx.stream()
    e0, e1 = hip.event(), hip.event()
    ctx.synchronize()
    sweeps = 0
    residuals = []
    t1 = time.perf_counter()
    hip.record(e0, stream)
    for _ in range(args.steps):
        info = ctx.pj_solve(dt, 1.0, p)
        sweeps += info.iters
        residuals.append(info.residual)
    hip.record(e1, stream)
    ctx.synchronize()
    elapsed = time.perf_counter() - t1
    ev_s = hip.elapsed_ms(e0, e1) * 1e-3
    hip.destroy(e0, e1)
    cups = prob.nc * sweeps / elapsed
    r_used, r_n, r_parts, r_per = ctx.range_sweep_info()
    g_ok, g_bad, g_used = ctx.geometry_rows_info()
    sweep_form = {"range_sweep": r_used, "ranges": r_n, "partials_per_sweep": r_parts, "contacts_per_range": r_per,
                  "mode": args.range_mode, "block": args.range_block or 512,
                  "geometry_rows": {"used": g_used, "usable": g_ok, "mismatching_contacts": g_bad}}
    log("sweep: " + (f"{r_n} contact ranges of {r_per}, {r_parts} partial sums" if r_used else "two passes")
        + (", angular rows rebuilt from the contact points" if g_used else
           f", angular rows read (points usable: {g_ok}, mismatching contacts: {g_bad})"))
    log(f"timed: {args.steps} solves, {sweeps} sweeps, {elapsed:.3f}s ({ev_s:.3f}s by stream events), "
        f"{cups:.3e} contact-updates/s")
    td = time.perf_counter()
    v = ctx.velocities()
    imp = ctx.impulses()
    t_download = time.perf_counter() - td
    cone_ok = bool((imp[:, 0] >= 0).all() and (np.linalg.norm(imp[:, 1:], axis=1) <= prob.mu * imp[:, 0] * (1 + 1e-12) + 1e-300).all())
    comp_pj = metrics.complementarity(prob, v, imp)

    # the whole sweep over the timed region: algorithmic bytes of every sweep / stream-event time
    bytes_sweep = bc + bb
    ach_sweep = bytes_sweep * sweeps / ev_s / 1e9
    # per-kernel durations: one more solve with start/stop events carried in every dispatch
    kernels = None
    if not args.no_kernel_events:
        stats, pinfo = profiled_kernels(ctx, lambda: ctx.pj_solve(dt, 1.0, p))
        # launches enqueued after convergence return at once; average over the sweeps that did work
        psweeps = max(pinfo.iters, 1)
        kernels = {}
        # the sweep's kernels: the range pass and the body pass over its partial sums, or the two
        # passes (contacts, bodies); the residual reduction every 10th sweep
        sweep_bytes = {"pj_range": bc, "pj_bodies_parts": bb, "pj_contacts": bc, "pj_bodies": bb}
        for name in ("pj_range", "pj_bodies_parts", "pj_contacts", "pj_bodies", "pj_residual"):
            calls, ms = stats.get(name, (0, 0.0))
            if not calls:
                continue
            n_eff = psweeps if name != "pj_residual" else max(1, (psweeps + 9) // 10)
            avg = ms / 1e3 / n_eff
            nbytes = sweep_bytes.get(name)
            kernels[name] = {"avg_us": avg * 1e6, "launches_per_sweep": n_eff / psweeps,
                             "algorithmic_bytes": nbytes,
                             "GBs": nbytes / avg / 1e9 if nbytes and avg > 0 else None,
                             "frac": nbytes / avg / 1e9 / HBM_PEAK_GBS if nbytes and avg > 0 else None}
        kern_sum = sum(k["avg_us"] * k["launches_per_sweep"] for k in kernels.values())
        kernels["kernel_us_per_sweep"] = kern_sum
        kernels["timing"] = ("start/stop HIP events carried in every kernel dispatch (hipExtLaunchKernelGGL) on the "
                             "solve stream, one more solve of the same workload right after the timed ones")
    # the committed counters were taken with the geometry form of the rows
    traffic, traffic_src = committed_traffic(prob.nc) if g_used else (None, None)

    # accelerated loops on the same problem and tolerance: time to solution
    accel = None if args.no_accel else accel_leg(native, ctx, prob, dt, args, metrics, elapsed / args.steps,
                                                 sweeps / args.steps, bytes_sweep)

    # coloured Gauss-Seidel on the same problem and tolerance (residual every sweep)
    if args.no_gs:
        gs = None
    else:
        gs = colored_gs_leg(native, ctx, prob, dt, args, metrics, v, comp_pj, elapsed, bytes_sweep)
    ctx.close()

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        threads = cpu_threads()
        prm = oracle.params(max_iter=args.cpu_sweeps, tol_abs=0, tol_rel=0, warmstart=0)
        tc = time.perf_counter()
        st, v_cpu, imp_cpu, res = oracle.pj_solve(prob, dt, prm, nthreads=threads)
        tcpu = time.perf_counter() - tc
        tc = time.perf_counter()
        st2, _, _, res2 = oracle.pj_solve(prob, dt, prm, nthreads=-threads)
        tcpu2 = time.perf_counter() - tc
        cpu = {"value": prob.nc * res.iters / tcpu, "unit": "contact-updates/s", "cores": threads, "kind": "port",
               "cpu": cpu_model(), "host_cpus": os.cpu_count(),
               "sample": f"oracle PJ (C restatement of projected_jacobi.cpp: single-thread SpMV as the reference pins "
                         f"Eigen to one thread, OpenMP projection on {threads} threads = every CPU this process may use) "
                         f"on the same {prob.nc}-contact problem, {res.iters} cold-start sweeps incl. W assembly, "
                         f"{tcpu:.1f}s",
               "fully_threaded": {"value": prob.nc * res2.iters / tcpu2, "cores": threads,
                                  "sample": f"same sweeps with both SpMVs and the vector updates on {threads} OpenMP "
                                            f"threads (upper bound, not the reference's structure), {tcpu2:.1f}s"},
               "gpu_over_cpu": cups / (prob.nc * res.iters / tcpu)}
        log(f"cpu baseline: {cpu['value']:.3e} (threaded {cpu['fully_threaded']['value']:.3e}) contact-updates/s "
            f"on {threads} threads")

    out = {
        "metric": "FP64 contact-updates/s at res<=1e-8 (projected Jacobi)",
        "value": cups,
        "unit": "contact-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "sphere pile: 100k spheres r~logU[0.02,0.08] m, rho=1000, 4x4x8 m box, "
                               "mu~U[0.3,0.8], h=5e-4 s, cold-start PJ to tol 1e-8",
                   "spheres": prob.nb, "contacts": prob.nc, "rows": int(info.nrows), "sweeps_per_step": sweeps / args.steps,
                   "residual": float(np.max(residuals)), "alpha": 0.3, "cone_invariant": cone_ok,
                   "complementarity": comp_pj,
                   "pcie_inclusive": {"upload_ms": t_upload * 1e3, "download_ms": t_download * 1e3,
                                      "value": prob.nc * sweeps / (elapsed + args.steps * (t_upload + t_download)),
                                      "note": "each step with its host bodies/contacts uploaded and velocities/impulses "
                                              "read back through the C ABI"}},
        "roofline": {"bound": "hbm", "kernel": "pj_sweep", "achieved": ach_sweep, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach_sweep / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "bytes_per_sweep": bytes_sweep, "avg_sweep_us": ev_s / sweeps * 1e6, "sweep_form": sweep_form,
                     "definition": "one projected-Jacobi sweep (contact pass + body pass, residual every 10th): "
                                   "algorithmic bytes per sweep (328 B per contact with two dynamic sides, 256 B with a "
                                   "static side, 128 B per body) x sweeps / time of the timed solves measured by HIP "
                                   "events recorded on the solve stream around them (setup kernels and launch gaps "
                                   "included)",
                     "kernels": kernels},
        "cpu_baseline": cpu,
        "accelerated": accel,
        "colored_gs": gs,
    }
    return out


def accel_leg(native, ctx, prob, dt, args, metrics, t_std, sweeps_std, bytes_sweep):
    """Chebyshev (pj.chebyshev), Nesterov (pj.nesterov) and Anderson (pj.anderson, history
    pj.anderson_m) loops on the pile, cold start, to the same tolerance as the standard loop:
    sweeps and wall time of one solve each.  Anderson is capped at 3x the standard loop's sweeps:
    on sphere piles it stalls at this tolerance (as the oracle's restatement of the reference loop
    does), and an uncapped run would spend the whole iteration budget."""
    out = {"standard": {"sweeps": sweeps_std, "s": t_std}}
    for name, mode in (("chebyshev", native.CDX_ACCEL_CHEBYSHEV), ("nesterov", native.CDX_ACCEL_NESTEROV),
                       ("anderson", native.CDX_ACCEL_ANDERSON)):
        p = native.default_params()
        p.tol_abs = p.tol_rel = args.tol
        p.max_iter = args.max_iter if mode != native.CDX_ACCEL_ANDERSON else min(args.max_iter, int(3 * max(sweeps_std, 1)))
        p.warmstart = 0
        p.accel = mode
        p.anderson_m = args.anderson_m
        ctx.synchronize()
        t0 = time.perf_counter()
        info = ctx.pj_solve(dt, 1.0, p)
        ctx.synchronize()
        t = time.perf_counter() - t0
        v, imp = ctx.velocities(), ctx.impulses()
        out[name] = {"sweeps": int(info.iters), "s": t, "ms_per_sweep": t / max(info.iters, 1) * 1e3,
                     **({"history": args.anderson_m, "max_iter": p.max_iter} if mode == native.CDX_ACCEL_ANDERSON else {}),
                     "converged": bool(info.residual <= 1.0), "residual": float(info.residual),
                     "speedup_vs_standard": t_std / t, "complementarity": metrics.complementarity(prob, v, imp)}
        log(f"{name}: {info.iters} sweeps, {t:.3f}s ({t_std / t:.2f}x the standard loop), residual {info.residual:.3e}")
        if name == "chebyshev" and not args.no_kernel_events:
            stats, pinfo = profiled_kernels(ctx, lambda: ctx.pj_solve(dt, 1.0, p))
            sw = max(pinfo.iters, 1)
            # the range sweep's kernels (pjc_range, pjc_bodies_parts) or the two passes'
            per = {k: stats.get(k, (0, 0.0))[1] / sw * 1e3
                   for k in ("pjc_range", "pjc_bodies_parts", "pjc_contacts", "pjc_bodies", "pjc_residual")}
            per = {k: v for k, v in per.items() if v > 0 or k == "pjc_residual"}
            us = sum(per.values())
            out[name]["kernels"] = {**{f"{k}_us_per_sweep": v for k, v in per.items()}, "kernel_us_per_sweep": us,
                                    "achieved_GBs_standard_bytes": bytes_sweep / (us * 1e-6) / 1e9,
                                    "timing": "HIP events in every dispatch, one more solve (direct launches)"}
            log(f"chebyshev kernels: {us:.1f} us per sweep (" + ", ".join(f"{k} {v:.1f}" for k, v in per.items()) + ")")
    return out


# --------------------------------------------------------------------------------------------
# batched: thousands of independent scenes, one solve launch per step
# --------------------------------------------------------------------------------------------

def _scene_problem(bd, con, warm, v, q, b0, b1, c0, c1):
    """Scene [b0, b1) / contacts [c0, c1) of the stacked state as a standalone problem."""
    def rebase(ids):
        ids = ids[c0:c1].copy()
        ids[ids >= 0] -= b0
        return ids
    bodies = (np.full(b1 - b0, 6, np.int32), bd["minv"][6 * b0:6 * b1], v[6 * b0:6 * b1], bd["fext"][6 * b0:6 * b1],
              None, q[7 * b0:7 * b1])
    contacts = (rebase(con["body_a"]), rebase(con["body_b"]), con["mu"][c0:c1], con["frame"][c0:c1], con["angA"][c0:c1],
                con["angB"][c0:c1], con["rest"][c0:c1], con["vbias"][c0:c1], warm[c0:c1])
    return bodies, contacts


def settled_args(args):
    """The batched leg on the configured contact load: the same scenes settled into piles first
    (untimed, 4e-3 s steps at a loose tolerance; about 1.5M contacts, up to ~4,800 per scene),
    then a few timed steps, one step compared scene by scene, the default launch shape."""
    a = argparse.Namespace(**vars(args))
    a.batch_settle, a.batch_steps, a.obo_steps, a.tune_step = args.settled_settle, args.settled_steps, 1, -1
    return a


def batched_leg(args, native, scenes):
    t0 = time.time()
    b = scenes.batched_drop(nscenes=args.batch_scenes, seed=0)
    bd = b.bodies()
    ns, nb = b.nscenes, len(b.mass)
    log(f"batched: {ns} scenes, {nb} bodies, generated in {time.time() - t0:.1f}s")
    dt = 1e-3
    accel = {"none": native.CDX_ACCEL_NONE, "chebyshev": native.CDX_ACCEL_CHEBYSHEV}[args.batch_accel]
    p = native.default_params()
    p.tol_abs = p.tol_rel = args.batch_tol
    p.max_iter = args.batch_max_iter
    p.warmstart = 1
    p.accel = accel
    p.block_precond = 1 if args.batch_precond == "block" else 0
    ctx = native.Context(0)
    ctx.set_bodies(bd["dof"], bd["minv"], bd["v"], bd["fext"], None, bd["q"])
    scene_of = np.repeat(np.arange(b.nscenes), np.diff(b.body_off))
    one = native.Context(0)
    rng = np.random.default_rng(1)
    steps = args.batch_steps
    sample_steps = set(int(s) for s in np.linspace(steps // 5, steps - 1, args.obo_steps).round()) if args.obo_steps else set()
    q = bd["q"].copy()
    v = bd["v"].copy()
    prev = None

    def contacts_of(q, prev):
        con = b.contacts(q=q.reshape(-1, 7))
        nc = len(con["body_a"])
        warm = scenes.match_warmstart(prev, con) if prev is not None and nc else np.zeros((nc, 3))
        return con, warm

    def upload(con, warm):
        ctx.set_contacts(con["body_a"], con["body_b"], con["mu"], con["frame"], con["angA"], con["angB"], con["rest"],
                         con["vbias"], warm)
        ctx.set_scenes(b.body_off, con["contact_off"])

    def advance(h, q_prev, need_v):
        """Moreau update on the device; a scene whose state went non-finite (projected Jacobi is not
        guaranteed to converge) is put back at rest where it was.  The positions (the next contact
        generation) and impulses (the next warm start) come down every step, the velocities only
        when the next step compares scenes one by one or a scene is reset (None otherwise); the
        finiteness test runs on the device.  Returns q, v, impulses, resets."""
        ctx.moreau_update(h, 1.0)
        q, imp = ctx.positions(), ctx.impulses()
        nbad = 0
        if ctx.nonfinite_bodies() == 0:
            return q, (ctx.velocities() if need_v else None), imp, nbad
        v = ctx.velocities()
        bad_body = ~(np.isfinite(v.reshape(-1, 6)).all(axis=1) & np.isfinite(q.reshape(-1, 7)).all(axis=1))
        if bad_body.any():
            bad = np.unique(scene_of[bad_body])
            nbad = len(bad)
            mask = np.isin(scene_of, bad)
            q = q.reshape(-1, 7).copy()
            v = v.reshape(-1, 6).copy()
            q[mask] = q_prev.reshape(-1, 7)[mask]
            v[mask] = 0.0
            q, v = q.reshape(-1), v.reshape(-1)
            imp = np.where(np.isfinite(imp), imp, 0.0)
            ctx.set_bodies(bd["dof"], bd["minv"], v, bd["fext"], None, q)
        return q, v, imp, nbad

    # settle (input preparation, not timed): larger steps at a loose tolerance until the scenes
    # lie in piles, so the timed steps run on the configured contact load
    settle = None
    if args.batch_settle > 0:
        ts0 = time.perf_counter()
        ps = native.default_params()
        ps.tol_abs = ps.tol_rel = 1e-4
        ps.max_iter = 1000
        ps.warmstart = 1
        ps.block_precond = 1  # converges on the box manifolds where the row-scaled loops stagnate or diverge
        resets = 0
        for s in range(args.batch_settle):
            con, warm = contacts_of(q, prev)
            upload(con, warm)
            ctx.batch_pj_solve(args.batch_settle_dt, 1.0, ps)
            q, v, imp, nbad = advance(args.batch_settle_dt, q, s == args.batch_settle - 1 and 0 in sample_steps)
            resets += nbad
            prev = dict(body_a=con["body_a"], body_b=con["body_b"], point=con["point"], imp=imp)
        per = np.diff(con["contact_off"])
        settle = {"steps": args.batch_settle, "h": args.batch_settle_dt, "tol": 1e-4, "max_iter": 1000,
                  "loop": "standard, block preconditioner", "scene_resets": resets, "s": time.perf_counter() - ts0,
                  "contacts_after": int(len(con["body_a"])), "max_contacts_per_scene_after": int(per.max())}
        log(f"batched: settled {args.batch_settle} steps of {args.batch_settle_dt:g} s in {settle['s']:.1f}s, "
            f"{settle['contacts_after']} contacts (max {settle['max_contacts_per_scene_after']}/scene)")
    t_cpu = t_dev = t_solve = 0.0
    cu = 0  # contact-updates (sum over scenes of contacts x sweeps)
    contacts_total = 0
    capped = 0
    tune = None
    obo_time = batch_time_sampled = 0.0
    obo_solves = 0
    worst_dv = 0.0
    per_scene_max = 0
    per_scene_all = []
    blown = 0  # scene-steps whose velocities or positions went non-finite (reset to rest)
    for s in range(steps):
        ta = time.perf_counter()
        con, warm = contacts_of(q, prev)
        nc = len(con["body_a"])
        tb = time.perf_counter()
        upload(con, warm)
        if s == args.tune_step:
            tune = ctx.batch_autotune(dt, 1.0, p)
            log(f"batched: autotuned (block, groups) = ({tune[0]}, {tune[1]}), {tune[2]:.3f} ms at step {s}")
        ctx.synchronize()
        ts = time.perf_counter()
        iters, res = ctx.batch_pj_solve(dt, 1.0, p)
        t_step_solve = time.perf_counter() - ts
        t_solve += t_step_solve
        coff = con["contact_off"]
        per = np.diff(coff)
        per_scene_all.append(per.copy())
        per_scene_max = max(per_scene_max, int(per.max()) if len(per) else 0)
        cu += int((per.astype(np.int64) * iters).sum())
        contacts_total += nc
        capped += int((iters >= p.max_iter).sum())
        t_sample = 0.0  # the one-by-one comparison below is not part of the batched device path
        if s in sample_steps:
            # the same scenes one by one through the plain solver, on a sample of scenes
            tsa = time.perf_counter()
            vb = ctx.velocities()
            pick = rng.choice(ns, size=min(args.obo_scenes, ns), replace=False)
            t_one = 0.0
            for k in pick:
                b0, b1 = int(b.body_off[k]), int(b.body_off[k + 1])
                c0, c1 = int(coff[k]), int(coff[k + 1])
                bodies, contacts = _scene_problem(bd, con, warm, v, q, b0, b1, c0, c1)
                one.set_bodies(*bodies)
                one.set_contacts(*contacts)
                one.synchronize()
                t1 = time.perf_counter()
                one.pj_solve(dt, 1.0, p)
                t_one += time.perf_counter() - t1
                v1 = one.velocities()
                scale = max(np.abs(vb[6 * b0:6 * b1]).max(), 1e-30)
                worst_dv = max(worst_dv, float(np.abs(v1 - vb[6 * b0:6 * b1]).max() / scale))
            obo_time += t_one * ns / len(pick)
            batch_time_sampled += t_step_solve
            obo_solves += len(pick)
            t_sample = time.perf_counter() - tsa
        q, v, imp, nbad = advance(dt, q, s + 1 in sample_steps)
        if nbad:
            blown += nbad
            log(f"batched step {s}: {nbad} scenes went non-finite; reset to rest")
        tc = time.perf_counter()
        prev = dict(body_a=con["body_a"], body_b=con["body_b"], point=con["point"], imp=imp)
        t_cpu += tb - ta
        t_dev += tc - tb - t_sample
        if s % 50 == 0 or s == steps - 1:
            log(f"batched step {s}: {nc} contacts (max {int(per.max()) if len(per) else 0}/scene), "
                f"max sweeps {int(iters.max())}, solve {t_step_solve * 1e3:.2f} ms")
    ctx.close()
    one.close()
    allper = np.concatenate(per_scene_all) if per_scene_all else np.zeros(1, int)
    dist = {"p10": float(np.percentile(allper, 10)), "p50": float(np.percentile(allper, 50)),
            "p90": float(np.percentile(allper, 90)), "p99": float(np.percentile(allper, 99)), "max": int(allper.max()),
            "share_10_to_3000": float(((allper >= 10) & (allper <= 3000)).mean()),
            "last_step_max": int(per_scene_all[-1].max()) if per_scene_all else 0}
    out = {
        "metric": "batched env-steps/s (scenes x Moreau steps per second)",
        "workload": f"{ns} scenes of 8-512 spheres/boxes (50/50, box edges U[0.05,0.2] m) dropped into a unit box"
                    f"{f' and settled ({args.batch_settle} steps of {args.batch_settle_dt:g} s, untimed)' if settle else ''}, "
                    f"h=1e-3 s, {steps} steps, warm-started PJ ({args.batch_accel} loop, {args.batch_precond} preconditioner) to tol {args.batch_tol:g} "
                    f"per scene (reference default), one batched solve launch per step",
        "loop": args.batch_accel, "preconditioner": args.batch_precond, "settle": settle,
        "contacts_per_scene": dist,
        "steps": steps, "scenes": ns, "bodies": nb,
        "value": ns * steps / t_dev, "unit": "env-steps/s",
        "definition": "device path per step: contact upload, batched solve, position update, finiteness test on the "
                      "device, positions and impulses downloaded (velocities too when a scene is reset or the next "
                      "step compares scenes one by one)",
        "solve_only_env_steps_per_s": ns * steps / t_solve,
        "end_to_end_env_steps_per_s": ns * steps / (t_cpu + t_dev),
        "cpu_contact_generation_s": t_cpu, "device_s": t_dev, "solve_s": t_solve,
        "contact_updates_per_s": cu / t_solve, "mean_contacts_per_step": contacts_total / steps,
        "max_contacts_per_scene": per_scene_max, "scene_solves_at_iteration_cap": capped,
        "capped_rate": capped / float(ns * steps), "scene_steps_reset_after_divergence": blown,
        "reset_rate": blown / float(ns * steps),
        "launch": {"block": tune[0], "groups": tune[1],
                   "note": "the tuner's shape; with one scene per workgroup and every scene within 60 KiB of LDS the "
                           "solve keeps each scene in LDS and raises the block to one thread per contact and per body "
                           "of the largest scene (up to 512)"} if tune else None,
        "one_by_one": None,
    }
    if obo_solves:
        out["one_by_one"] = {
            "sampled_steps": sorted(sample_steps), "sampled_scene_solves": obo_solves,
            "estimated_solve_s_per_step": obo_time / len(sample_steps),
            "batched_solve_s_per_step": batch_time_sampled / len(sample_steps),
            "speedup": obo_time / max(batch_time_sampled, 1e-12),
            "max_rel_velocity_difference": worst_dv,
            "method": "at each sampled step a random sample of scenes is solved alone with cdx_pj_solve (same inputs, "
                      "same per-scene tolerance); solve time scaled to all scenes, compared with the batched solve of "
                      "that step"}
    return out


# --------------------------------------------------------------------------------------------
# jointed multibody: capsule chains with spherical joints on a heightfield, batched
# --------------------------------------------------------------------------------------------

def _scene_rows(rows, lam, b0, b1):
    """Joint rows of the bodies [b0, b1) (a whole chain), rebased, with their Lambda_prev."""
    n = rows["ns"] + rows["nd"]
    sel = (rows["body_a"] >= b0) & (rows["body_a"] < b1)
    spring = np.arange(n) < rows["ns"]
    sub = {k: np.asarray(rows[k])[sel] for k in ("body_a", "body_b", "wa", "wb", "comp", "vsrc", "gerr")}
    for k in ("body_a", "body_b"):
        sub[k] = np.where(sub[k] >= 0, sub[k] - b0, -1).astype(np.int32)
    sub["ns"], sub["nd"] = int((sel & spring).sum()), int((sel & ~spring).sum())
    return sub, (None if lam is None else lam[sel])


def chains_leg(args, native, scenes):
    from cardillocxx_amd import joints

    t0 = time.time()
    terrain = scenes.fractal_heightfield(n=256, amp=0.1, seed=0)
    b, jt = scenes.capsule_chains(nchains=args.chain_scenes, links=32, terrain=terrain, seed=0, drop=args.chain_drop,
                                  local=True)
    frames = joints.joint_frames(b.q, jt)
    bd = scenes.shape_bodies(b.shape, b.dims, b.mass, b.q, b.v.reshape(-1))
    ns, nb = b.nscenes, len(b.mass)
    log(f"chains: {ns} chains x 32 capsules, {len(jt['body_a'])} joints, generated in {time.time() - t0:.1f}s")
    dt, damping = 1e-3, args.chain_damping
    p = native.default_params()
    p.tol_abs = p.tol_rel = args.batch_tol
    p.max_iter = args.batch_max_iter
    p.warmstart = 1
    ctx = native.Context(0)
    ctx.set_bodies(bd["dof"], bd["minv"], bd["v"], bd["fext"], None, bd["q"])
    one = native.Context(0)
    rng = np.random.default_rng(2)
    steps = args.chain_steps
    sample_steps = set(int(s) for s in np.linspace(steps // 2, steps - 1, args.obo_steps).round()) if args.obo_steps else set()
    q, v = bd["q"].copy(), bd["v"].copy()
    prev, lam = None, None
    t_cpu = t_dev = t_solve = 0.0
    cu = contacts_total = capped = 0
    obo_time = batch_time_sampled = 0.0
    obo_solves, worst_dv, tune = 0, 0.0, None
    for s in range(steps):
        ta = time.perf_counter()
        con = b.contacts(q=q.reshape(-1, 7))
        nc = len(con["body_a"])
        warm = scenes.match_warmstart(prev, con) if prev is not None and nc else np.zeros((nc, 3))
        rows = joints.spherical_rows(q, jt, frames, damping)
        tb = time.perf_counter()
        ctx.set_contacts(con["body_a"], con["body_b"], con["mu"], con["frame"], con["angA"], con["angB"], con["rest"],
                         con["vbias"], warm)
        ctx.set_joints(rows, lam_prev=lam)
        ctx.set_scenes(b.body_off, con["contact_off"])
        if s == min(args.tune_step, steps - 1):
            tune = ctx.batch_autotune(dt, 1.0, p)
            log(f"chains: autotuned (block, groups) = ({tune[0]}, {tune[1]}), {tune[2]:.3f} ms at step {s}")
        ctx.synchronize()
        ts = time.perf_counter()
        iters, res = ctx.batch_pj_solve(dt, 1.0, p)
        t_step = time.perf_counter() - ts
        t_solve += t_step
        coff = con["contact_off"]
        per = np.diff(coff)
        cu += int((per.astype(np.int64) * iters).sum())
        contacts_total += nc
        capped += int((iters >= p.max_iter).sum())
        lam_new = ctx.joint_impulses()
        t_sample = 0.0  # the one-by-one comparison below is not part of the batched device path
        if s in sample_steps:  # the same chains one by one through cdx_pj_solve
            tsa = time.perf_counter()
            vb = ctx.velocities()
            pick = rng.choice(ns, size=min(args.obo_scenes, ns), replace=False)
            t_one = 0.0
            for k in pick:
                b0, b1 = int(b.body_off[k]), int(b.body_off[k + 1])
                c0, c1 = int(coff[k]), int(coff[k + 1])
                bodies, contacts = _scene_problem(bd, con, warm, v, q, b0, b1, c0, c1)
                sub, sl = _scene_rows(rows, lam, b0, b1)
                one.set_bodies(*bodies)
                one.set_contacts(*contacts)
                one.set_joints(sub, lam_prev=sl)
                one.synchronize()
                t1 = time.perf_counter()
                one.pj_solve(dt, 1.0, p)
                t_one += time.perf_counter() - t1
                v1 = one.velocities()
                scale = max(np.abs(vb[6 * b0:6 * b1]).max(), 1e-30)
                worst_dv = max(worst_dv, float(np.abs(v1 - vb[6 * b0:6 * b1]).max() / scale))
            obo_time += t_one * ns / len(pick)
            batch_time_sampled += t_step
            obo_solves += len(pick)
            t_sample = time.perf_counter() - tsa
        ctx.moreau_update(dt, 1.0)
        if ctx.nonfinite_bodies():
            raise RuntimeError(f"chains: non-finite state at step {s}")
        q, imp = ctx.positions(), ctx.impulses()
        v = ctx.velocities() if s + 1 in sample_steps else None  # the one-by-one comparison's inputs
        lam = lam_new
        tc = time.perf_counter()
        prev = dict(body_a=con["body_a"], body_b=con["body_b"], point=con["point"], imp=imp)
        t_cpu += tb - ta
        t_dev += tc - tb - t_sample
        if s % 50 == 0 or s == steps - 1:
            log(f"chains step {s}: {nc} contacts, max sweeps {int(iters.max())}, solve {t_step * 1e3:.2f} ms")
    ctx.close()
    one.close()
    out = {
        "metric": "batched jointed env-steps/s (chains x Moreau steps per second)",
        "workload": f"{ns} chains of 32 capsules (r=0.02 m, length 0.1 m) joined by damped spherical joints "
                    f"(D_rot={damping:g}), dropped {args.chain_drop:g} m onto a 256x256 fractal heightfield (amp 0.1 m), "
                    f"h=1e-3 s, {steps} steps, joint rows and contacts in one warm-started PJ per chain to tol "
                    f"{args.batch_tol:g}, one batched solve launch per step",
        "steps": steps, "scenes": ns, "bodies": nb, "joints": int(len(jt["body_a"])),
        "joint_rows": int(rows["ns"] + rows["nd"]),
        "value": ns * steps / t_dev, "unit": "env-steps/s",
        "definition": "device path per step: contact and joint-row upload, batched solve, position update, finiteness "
                      "test on the device, positions, impulses and joint impulses downloaded (velocities when the next "
                      "step compares chains one by one)",
        "solve_only_env_steps_per_s": ns * steps / t_solve,
        "end_to_end_env_steps_per_s": ns * steps / (t_cpu + t_dev),
        "cpu_contacts_and_joint_rows_s": t_cpu, "device_s": t_dev, "solve_s": t_solve,
        "contact_updates_per_s": cu / t_solve, "mean_contacts_per_step": contacts_total / steps,
        "scene_solves_at_iteration_cap": capped,
        "launch": {"block": tune[0], "groups": tune[1]} if tune else None,
        "one_by_one": None,
    }
    if obo_solves:
        out["one_by_one"] = {
            "sampled_steps": sorted(sample_steps), "sampled_scene_solves": obo_solves,
            "estimated_solve_s_per_step": obo_time / len(sample_steps),
            "batched_solve_s_per_step": batch_time_sampled / len(sample_steps),
            "speedup": obo_time / max(batch_time_sampled, 1e-12),
            "max_rel_velocity_difference": worst_dv,
            "method": "at each sampled step a random sample of chains is solved alone with cdx_pj_solve and its joint "
                      "rows (same inputs, same per-chain tolerance); time scaled to all chains, compared with the "
                      "batched solve of that step"}
    return out


# --------------------------------------------------------------------------------------------
# differentiable: forward + backward of batched rollouts
# --------------------------------------------------------------------------------------------

def diff_leg(args, native, scenes):
    from cardillocxx_amd import diff

    batch = scenes.boxes_on_plane(nscenes=args.diff_scenes, per_scene=64, seed=0)
    rng = np.random.default_rng(0)
    target = rng.normal(scale=0.1, size=(batch.nscenes, 3)) + np.array([0.875, 0.875, 0.05])
    kw = dict(steps=args.diff_steps, dt=1e-3, checkpoint_every=10, tol=args.tol, adjoint_tol=args.adjoint_tol,
              max_iter=args.batch_max_iter, adjoint_mode=native.CDX_ADJ_AUTO, adjoint_max_iter=args.adjoint_max_iter)

    def run(b, tgt, batched):
        r = diff.Rollout(b, batched=batched, **kw)
        v0, mu, mass = b.v.reshape(-1).copy(), b.mu_body.copy(), b.mass.copy()
        t0 = time.perf_counter()
        pos = r.forward(v0, mu, mass)
        t1 = time.perf_counter()
        loss, gpos, gm = diff.com_loss(pos, mass, b.body_off, tgt)
        g = r.backward(gpos)
        t2 = time.perf_counter()
        r.close()
        return r, loss, g, gm, t1 - t0, t2 - t1

    r, loss, g, gm, tf, tb = run(batch, target, True)
    ns, steps = batch.nscenes, args.diff_steps
    log(f"differentiable: {ns} scenes, forward {tf:.2f}s, backward {tb:.2f}s, "
        f"{r.stats['forward_solves']} solves, {r.stats['adjoint_iters']} adjoint applications, "
        f"{r.stats.get('least_squares_adjoints', 0)} least-squares adjoint steps, "
        f"worst adjoint relative residual {r.stats.get('worst_adjoint_residual', 0.0):.2e}")
    pick = rng.choice(ns, size=min(args.diff_obo, ns), replace=False)
    t_one = 0.0
    worst = worst_grid = 0.0

    def rel_diff(g1, gm1, b0, b1):
        # the masses' gradient through the solve is ~1e-18 here (equal masses): measured against the
        # scene's other gradients
        scale = max(np.abs(g.v0[b0:b1]).max(), np.abs(g.mu[b0:b1]).max())
        w = 0.0
        for a, c in ((g1.v0, g.v0[b0:b1]), (g1.mu, g.mu[b0:b1]), (g1.mass + gm1, (g.mass + gm)[b0:b1])):
            w = max(w, float(np.abs(a - c).max() / max(np.abs(c).max(), 1e-3 * scale, 1e-300)))
        return w

    for k in pick:
        k = int(k)
        b0, b1 = batch.body_off[k], batch.body_off[k + 1]
        # timed: the scene alone through the plain grid-wide solve and adjoint
        _, _, g1, gm1, tf1, tb1 = run(batch.scene(k), target[k][None, :], False)
        t_one += tf1 + tb1
        worst_grid = max(worst_grid, rel_diff(g1, gm1, b0, b1))
        # the same scene alone through the same per-scene solve as the batch (equal arithmetic)
        _, _, g2, gm2, _, _ = run(batch.scene(k), target[k][None, :], True)
        worst = max(worst, rel_diff(g2, gm2, b0, b1))
    t_obo = t_one * ns / len(pick)
    log(f"differentiable one by one: {len(pick)} scenes in {t_one:.2f}s; gradients vs batched {worst:.2e} "
        f"(same per-scene solve), {worst_grid:.2e} (grid-wide solve)")
    return {
        "metric": "fwd+bwd env-steps/s (scenes x steps through forward and backward per second)",
        "workload": f"{ns} scenes x 64 boxes (edge 0.1 m, 1 kg) on a plane, v0~N(0,0.5^2), {steps} steps at h=1e-3 s, "
                    f"loss sum_s ||COM_s(final) - target_s||^2, grads wrt v0, mu, mass; implicit FP64 adjoint by "
                    f"restarted GMRES (least-squares where inconsistent, at most {args.adjoint_max_iter} applications), "
                    f"step states kept by the forward, PJ tol {args.tol:g} per scene, adjoint tol {args.adjoint_tol:g}",
        "value": ns * steps / (tf + tb), "unit": "env-steps/s",
        "forward_s": tf, "backward_s": tb, "loss": loss,
        "forward_solves": r.stats["forward_solves"], "adjoint_applications": r.stats["adjoint_iters"],
        "unrolled_fallbacks": r.stats["unrolled_fallbacks"],
        "least_squares_adjoint_steps": r.stats.get("least_squares_adjoints", 0),
        "worst_adjoint_relative_residual": r.stats.get("worst_adjoint_residual", 0.0),
        "one_by_one": {"sampled_scenes": [int(k) for k in pick], "estimated_s": t_obo, "speedup": t_obo / (tf + tb),
                       "max_rel_gradient_difference": worst,
                       "max_rel_gradient_difference_grid_solve": worst_grid,
                       "method": "sampled scenes rolled out alone (forward + backward through the grid-wide solve and "
                                 "adjoint), time scaled to all scenes; gradients compared with the batched ones for the "
                                 "scene rolled out alone through the same per-scene solve as the batch (a box face on "
                                 "four contacts has non-unique impulses, so the grid-wide solve's other summation order "
                                 "follows a slightly different trajectory: its difference is reported beside)"},
    }


def chains_diff_leg(args, native, scenes):
    """Forward + backward of the jointed capsule chains (the forward keeps every step's state, so
    the backward recomputes nothing): gradient of the COM loss
    w.r.t. every joint's rotational damping (and v0) through the implicit joint adjoint."""
    from cardillocxx_amd import diff

    terrain = scenes.fractal_heightfield(n=256, amp=0.1, seed=0)
    batch, jt = scenes.capsule_chains(nchains=args.chain_diff_scenes, links=32, terrain=terrain, seed=1, drop=0.0005,
                                      local=True)
    rng = np.random.default_rng(3)
    batch.v[:] = rng.normal(scale=0.2, size=batch.v.shape)
    ns, steps = batch.nscenes, args.chain_diff_steps
    com0 = np.stack([batch.q[batch.body_off[s]:batch.body_off[s + 1], :3].mean(0) for s in range(ns)])
    target = com0 + rng.normal(scale=0.05, size=(ns, 3))
    D0 = np.full(len(jt["body_a"]), args.chain_damping)
    kw = dict(steps=steps, dt=1e-3, checkpoint_every=10, tol=args.tol, adjoint_tol=args.adjoint_tol,
              max_iter=args.max_iter, adjoint_max_iter=args.chain_adjoint_max_iter)

    def run(b, spec, D, tgt, batched):
        r = diff.ChainRollout(b, spec, batched=batched, **kw)
        t0 = time.perf_counter()
        pos = r.forward(b.v.reshape(-1).copy(), D)
        t1 = time.perf_counter()
        loss, gpos, _ = diff.com_loss(pos, b.mass, b.body_off, tgt)
        g = r.backward(gpos)
        t2 = time.perf_counter()
        r.close()
        return r, loss, g, t1 - t0, t2 - t1

    r, loss, g, tf, tb = run(batch, jt, D0, target, True)
    log(f"chains differentiable: {ns} chains, forward {tf:.2f}s, backward {tb:.2f}s, {r.stats['forward_solves']} solves, "
        f"{r.stats['adjoint_iters']} adjoint iterations")
    pick = rng.choice(ns, size=min(args.chain_diff_obo, ns), replace=False)
    t_one, worst = 0.0, 0.0
    nj_per = 31
    for k in pick:
        k = int(k)
        b0 = int(batch.body_off[k])
        js = slice(k * nj_per, (k + 1) * nj_per)
        spec = {key: (np.asarray(x)[js] - (b0 if key in ("body_a", "body_b") else 0)) for key, x in jt.items()}
        _, _, g1, tf1, tb1 = run(batch.scene(k), spec, D0[js], target[k][None, :], False)
        t_one += tf1 + tb1
        for a, c in ((g1.damping, g.damping[js]), (g1.v0, g.v0[b0:int(batch.body_off[k + 1])])):
            worst = max(worst, float(np.abs(a - c).max() / max(np.abs(c).max(), 1e-30)))
    t_obo = t_one * ns / len(pick)
    log(f"chains differentiable one by one: {len(pick)} chains in {t_one:.2f}s")
    return {
        "metric": "fwd+bwd jointed env-steps/s (chains x steps through forward and backward per second)",
        "workload": f"{ns} chains of 32 capsules (r=0.02 m, length 0.1 m), damped spherical joints (D_rot="
                    f"{args.chain_damping:g}), dropped onto a 256x256 fractal heightfield (amp 0.1 m), v0~N(0,0.2^2), "
                    f"{steps} steps at h=1e-3 s, loss sum_s ||COM_s(final) - target_s||^2, grads wrt every joint's damping "
                    f"and v0; implicit FP64 adjoint through joints and contacts, step states kept by the forward, PJ tol "
                    f"{args.tol:g} per chain, adjoint tol {args.adjoint_tol:g}",
        "value": ns * steps / (tf + tb), "unit": "env-steps/s",
        "forward_s": tf, "backward_s": tb, "loss": loss,
        "forward_solves": r.stats["forward_solves"], "adjoint_applications": r.stats["adjoint_iters"],
        "least_squares_adjoint_steps": r.stats.get("inexact_adjoints", 0),
        "worst_adjoint_relative_residual": r.stats.get("worst_adjoint_residual", 0.0),
        "grad_damping_norm": float(np.linalg.norm(g.damping)),
        "one_by_one": {"sampled_chains": [int(k) for k in pick], "estimated_s": t_obo, "speedup": t_obo / (tf + tb),
                       "max_rel_gradient_difference": worst,
                       "method": "sampled chains rolled out alone (forward + backward through the grid-wide joint solve "
                                 "and adjoint), time scaled to all chains"},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spheres", type=int, default=100_000)
    ap.add_argument("--settle", type=int, default=600)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-iter", type=int, default=200_000)
    ap.add_argument("--gs-max-iter", type=int, default=50_000)
    ap.add_argument("--no-gs", action="store_true", help="skip the coloured Gauss-Seidel solve of the pile")
    ap.add_argument("--no-accel", action="store_true", help="skip the Chebyshev / Nesterov / Anderson solves of the pile")
    ap.add_argument("--anderson-m", type=int, default=2, help="Anderson history window (pj.anderson_m)")
    ap.add_argument("--cpu-sweeps", type=int, default=900, help="sweeps in the CPU baseline sample")
    ap.add_argument("--no-graphs", action="store_true",
                    help="pile leg: launch every sweep's kernels directly instead of replaying graph-captured chunks "
                         "(profiler runs)")
    ap.add_argument("--range-mode", type=int, default=1, choices=(0, 1, 2),
                    help="range sweep of projected Jacobi: 0 two passes, 1 from 32,768 contacts, 2 always")
    ap.add_argument("--range-contacts", type=int, default=0, help="contacts per range (0 = 128)")
    ap.add_argument("--range-block", type=int, default=0, help="threads per range workgroup (0 = 512, or 256)")
    ap.add_argument("--no-geometry-rows", action="store_true",
                    help="range sweep reads the angular rows instead of rebuilding them from the contact points")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-mb", type=int, default=0, help="also run one device copy of this size (counter calibration)")
    ap.add_argument("--no-kernel-events", action="store_true",
                    help="skip the extra solves timed with per-dispatch events (for profiler runs)")
    ap.add_argument("--pile-only", action="store_true", help="skip the batched and differentiable legs")
    ap.add_argument("--batch-scenes", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--batch-max-iter", type=int, default=20_000)
    ap.add_argument("--batch-tol", type=float, default=1e-5, help="batched PJ tolerance (the reference default)")
    ap.add_argument("--tune-step", type=int, default=100)
    ap.add_argument("--batch-accel", choices=("none", "chebyshev"), default="none",
                    help="loop of the batched solve (pj.chebyshev); the one-by-one comparison uses the same")
    ap.add_argument("--batch-precond", choices=("row", "block"), default="block",
                    help="batched PJ preconditioner: the row scaling alpha / D_ii (reference default) or the "
                         "per-contact inverse Delassus blocks (pj.rdiag_true_delassus)")
    ap.add_argument("--batch-settle", type=int, default=0, help="untimed settling steps before the batched steps")
    ap.add_argument("--batch-settle-dt", type=float, default=4e-3)
    ap.add_argument("--batched-only", action="store_true", help="run only the batched leg and print it")
    ap.add_argument("--settled-steps", type=int, default=20,
                    help="timed steps of the batched leg again on the settled piles (0: skip)")
    ap.add_argument("--settled-settle", type=int, default=400, help="untimed settling steps before them")
    ap.add_argument("--diff-only", action="store_true", help="run only the differentiable legs and print them")
    ap.add_argument("--obo-steps", type=int, default=4, help="steps at which scenes are also solved one by one")
    ap.add_argument("--obo-scenes", type=int, default=128, help="scenes solved one by one per sampled step")
    ap.add_argument("--no-chains", action="store_true", help="skip the jointed capsule-chain leg")
    ap.add_argument("--chains-only", action="store_true",
                    help="run only the two jointed-chain legs and print them (for profiling; not the bench line)")
    ap.add_argument("--chain-scenes", type=int, default=2048)
    ap.add_argument("--chain-steps", type=int, default=150)
    ap.add_argument("--chain-drop", type=float, default=0.03, help="drop height above the terrain under each chain")
    ap.add_argument("--chain-damping", type=float, default=0.05, help="rotational joint damping D_rot")
    ap.add_argument("--chain-diff-scenes", type=int, default=2048)
    ap.add_argument("--chain-diff-steps", type=int, default=40)
    ap.add_argument("--chain-diff-obo", type=int, default=4, help="chains rolled out one by one")
    ap.add_argument("--chain-adjoint-max-iter", type=int, default=20_000)
    ap.add_argument("--diff-scenes", type=int, default=256)
    ap.add_argument("--diff-steps", type=int, default=100)
    ap.add_argument("--diff-obo", type=int, default=4, help="scenes rolled out one by one")
    ap.add_argument("--adjoint-tol", type=float, default=1e-10)
    ap.add_argument("--adjoint-max-iter", type=int, default=2_000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); scenes and the pile are "
              "not sharded across GPUs here.", file=sys.stderr)
        sys.exit(2)

    from cardillocxx_amd import metrics, native, scenes

    hip = Hip()
    if args.copy_mb:
        hip.copy_once(args.copy_mb)
        log(f"calibration copy: {args.copy_mb} MiB device to device")
    copy_gbs = hip.copy_bandwidth()
    log(f"device copy: {copy_gbs:.0f} GB/s (read + write)")

    if args.diff_only:
        out = {"differentiable": diff_leg(args, native, scenes)}
        if not args.no_chains:
            out["jointed_chains_differentiable"] = chains_diff_leg(args, native, scenes)
        print(json.dumps(out), flush=True)
        return
    if args.batched_only:
        out = {"batched": batched_leg(args, native, scenes)}
        if args.settled_steps > 0:
            out["batched_settled"] = batched_leg(settled_args(args), native, scenes)
        print(json.dumps(out), flush=True)
        return
    if args.chains_only:
        print(json.dumps({"jointed_chains": chains_leg(args, native, scenes),
                          "jointed_chains_differentiable": chains_diff_leg(args, native, scenes)}), flush=True)
        return
    out = pile_leg(args, native, scenes, metrics, hip)
    out["roofline"]["copy_GBs"] = copy_gbs
    out["roofline"]["frac_of_copy"] = out["roofline"]["achieved"] / copy_gbs
    if not args.pile_only:
        out["batched"] = batched_leg(args, native, scenes)
        if args.settled_steps > 0:
            out["batched_settled"] = batched_leg(settled_args(args), native, scenes)
        out["differentiable"] = diff_leg(args, native, scenes)
        if not args.no_chains:
            out["jointed_chains"] = chains_leg(args, native, scenes)
            out["jointed_chains_differentiable"] = chains_diff_leg(args, native, scenes)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
