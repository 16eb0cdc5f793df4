"""Benchmark: FP64 projected-Jacobi contact updates per second on one MI355X.

Workload (the large single scene): 100k spheres, radius log-uniform in [0.02, 0.08] m, density
1000 kg/m^3, random packing settled in a 4x4x8 m box, mu ~ U[0.3, 0.8] per body (min-combined),
~4e5 frictional contacts, h = 5e-4 s.  One step = one full projected-Jacobi solve of that contact
problem from a cold start (no warm start) to the reference residual test with
tol_abs = tol_rel = 1e-8 (checked every 10 sweeps), i.e. the hot path over one batch of input
with everything resident in HBM.  contact-updates/s = contacts x sweeps / time.

Prints one JSON line.  `value` comes from the timed steps (production path: chunks of sweeps
replayed as HIP graphs).  `roofline` is the dominant kernel (pj_contacts): its average launch
duration from start/stop HIP events carried in every dispatch on the solve's own stream during one
more step of the same workload right after the timed ones; `traffic` is the HBM bytes per launch
from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary when it was taken on this workload.  `cpu_baseline` is the CPU oracle (a
restatement of the reference algorithm) timed on a bounded sample of the same workload.

--copy-mb N also runs one device-to-device copy of N MiB, the calibration kernel for the
FETCH_SIZE / WRITE_SIZE counters collected under rocprofv3 --pmc.
"""
from __future__ import annotations

import argparse
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


def committed_traffic(kernel, contacts):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary, if it was
    taken on a problem with the same contact count (FETCH_SIZE/WRITE_SIZE passes scaled by a
    256 MiB copy calibration); None otherwise."""
    path = os.path.join(ROOT, "profiles", "pj_contacts4_pile100k_hbm_traffic.txt")
    try:
        vals = {}
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] in ("traffic_bytes_per_launch", "traffic_contacts") and f[1] == kernel:
                vals[f[0]] = int(f[2])
        if vals.get("traffic_contacts") == contacts:
            return vals["traffic_bytes_per_launch"], os.path.relpath(path, ROOT)
    except (OSError, ValueError, KeyError):
        pass
    return None, None


def algorithmic_bytes(prob):
    both = int(((prob.body_a >= 0) & (prob.body_b >= 0)).sum())
    one = prob.nc - both
    return both * B_CONTACT_DYN + one * B_CONTACT_STATIC, prob.nb * B_BODY


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spheres", type=int, default=100_000)
    ap.add_argument("--settle", type=int, default=600)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-iter", type=int, default=200_000)
    ap.add_argument("--cpu-sweeps", type=int, default=100, help="sweeps in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-mb", type=int, default=0, help="also run one device copy of this size (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); a single scene does not "
              "shard across GPUs here -- see DESIGN.md for how it would.", file=sys.stderr)
        sys.exit(2)

    from cardillocxx_amd import native, scenes

    if args.copy_mb:
        import ctypes

        hip = ctypes.CDLL("libamdhip64.so")
        nbytes = args.copy_mb << 20
        src, dst = ctypes.c_void_p(), ctypes.c_void_p()
        hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes))
        hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes))
        hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
        hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)  # device to device
        hip.hipDeviceSynchronize()
        hip.hipFree(src)
        hip.hipFree(dst)
        print(f"# calibration copy: {args.copy_mb} MiB device to device", file=sys.stderr, flush=True)

    t0 = time.time()
    pile = scenes.sphere_pile(n=args.spheres, settle_steps=args.settle, seed=0)
    prob = pile.problem()
    t_scene = time.time() - t0
    print(f"# scene: {prob.nb} bodies, {prob.nc} contacts, generated in {t_scene:.1f}s", file=sys.stderr, flush=True)
    dt = 5e-4
    bc, bb = algorithmic_bytes(prob)

    ctx = native.Context(0)
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    p = native.default_params()
    p.tol_abs = p.tol_rel = args.tol
    p.max_iter = args.max_iter
    p.warmstart = 0

    for _ in range(args.warmup):
        info = ctx.pj_solve(dt, 1.0, p)
        print(f"# warmup solve: {info.iters} sweeps, residual {info.residual:.3e}", file=sys.stderr, flush=True)

    ctx.synchronize()
    sweeps = 0
    residuals = []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        info = ctx.pj_solve(dt, 1.0, p)
        sweeps += info.iters
        residuals.append(info.residual)
    ctx.synchronize()
    t2 = time.perf_counter()
    elapsed = t2 - t1
    cups = prob.nc * sweeps / elapsed

    # Kernel durations: one more step of the same workload with start/stop HIP events carried in
    # every dispatch on the solve's stream (graph replay is off while events are recorded).
    ctx.set_profiling(True)
    for k in ("pj_contacts", "pj_bodies", "pj_residual"):
        ctx.kernel_time(k)
    tp0 = time.perf_counter()
    pinfo = ctx.pj_solve(dt, 1.0, p)
    ctx.synchronize()
    tp = time.perf_counter() - tp0
    ncall_c, ms_c = ctx.kernel_time("pj_contacts")
    ncall_b, ms_b = ctx.kernel_time("pj_bodies")
    ctx.set_profiling(False)
    psweeps = max(pinfo.iters, 1)
    # average over the sweeps that did work (launches enqueued after convergence return at once
    # and are included in the total, so this slightly overstates the duration)
    avg_c = ms_c / 1e3 / psweeps
    avg_b = ms_b / 1e3 / psweeps
    ach_c = bc / avg_c / 1e9
    ach_sweep = (bc + bb) * sweeps / elapsed / 1e9

    traffic, traffic_src = committed_traffic("k_pj_contacts4", prob.nc)

    v = ctx.velocities()
    imp = ctx.impulses()
    cone_ok = bool((imp[:, 0] >= 0).all() and (np.linalg.norm(imp[:, 1:], axis=1) <= prob.mu * imp[:, 0] * (1 + 1e-12) + 1e-300).all())

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        threads = min(os.cpu_count() or 1, 16)
        prm = oracle.params(max_iter=args.cpu_sweeps, tol_abs=0, tol_rel=0, warmstart=0)
        tc = time.perf_counter()
        st, v_cpu, imp_cpu, res = oracle.pj_solve(prob, dt, prm, nthreads=threads)
        tcpu = time.perf_counter() - tc
        cpu = {"value": prob.nc * res.iters / tcpu, "unit": "contact-updates/s", "cores": threads, "kind": "port",
               "sample": f"oracle PJ (C restatement of projected_jacobi.cpp, OpenMP projection, single-thread "
                         f"SpMV) on the same {prob.nc}-contact problem, {res.iters} sweeps incl. W assembly, "
                         f"{tcpu:.1f}s"}

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
                   "residual": float(np.max(residuals)), "alpha": 0.3, "cone_invariant": cone_ok},
        "roofline": {"bound": "hbm", "kernel": "pj_contacts", "achieved": ach_c, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach_c / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "bytes_per_launch": bc, "avg_launch_us": avg_c * 1e6,
                     "pj_bodies_avg_us": avg_b * 1e6, "pj_bodies_GBs": bb / avg_b / 1e9 if avg_b > 0 else None,
                     "whole_sweep_GBs": ach_sweep, "whole_sweep_frac": ach_sweep / HBM_PEAK_GBS,
                     "timing": "start/stop HIP events carried in every kernel dispatch (hipExtLaunchKernelGGL) on the "
                               "solve stream, one profiled step of the same workload run right after the timed steps",
                     "profiled_step_ms": tp * 1e3, "profiled_sweeps": pinfo.iters},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
