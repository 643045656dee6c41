#!/usr/bin/env python3
"""Benchmark of the MI355X apply path: GMRES solves/s, p50/p99 latency to rel. residual 1e-8,
FP32 local factors / FP64 Krylov against the plain FP64 GPU path.

Workload (default): 3D Poisson 7-point on 64^3 (n = 262,144), 64 subdomains of uneven size
(recursive coordinate bisection; overlap 2 makes n_i range over 5.7k-7.4k), local LL^T factors,
nev = 24, deflated two-level RAS, GMRES(50) in FP64 to rtol 1e-8. A step is one complete GMRES
solve of one right-hand side (uniform[-1, 1] from a seed) whose b and x are resident in HBM.

One JSON line on stdout (see the field list at the end of main()).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)


def _load():
    import importlib.util
    pkg = os.path.join(REPO, "algebraic-multilevel-preconditioner_amd")
    if "amp_amd" not in sys.modules:
        spec = importlib.util.spec_from_file_location("amp_amd", os.path.join(pkg, "__init__.py"),
                                                      submodule_search_locations=[pkg])
        mod = importlib.util.module_from_spec(spec)
        sys.modules["amp_amd"] = mod
        spec.loader.exec_module(mod)


HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (generator, kwargs, subdomains, overlap, nev, restart, description)
    "poisson3d-64": ("poisson3d", dict(nx=64), 64, 2, 24, 50,
                     "3D Poisson 7-point 64^3, 64 subdomains, overlap 2, nev 24, GMRES(50), rtol 1e-8"),
    "poisson2d-256": ("poisson2d", dict(nx=256), 8, 1, 20, 30,
                      "2D Poisson 5-point 256^2, 8 subdomains, overlap 1, nev 20, additive, GMRES(30), rtol 1e-8"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="poisson3d-64", choices=sorted(WORKLOADS))
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--tasks", type=int, default=1, help="tasks per subdomain of the batched local solve")
    ap.add_argument("--profile-applies", type=int, default=0,
                    help="only run this many FP32 applies after setup (for rocprofv3 runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the path would shard by subdomain across GPUs, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    _load()
    from amp_amd import device, hostsetup as hs

    gen, kw, N, overlap, nev, restart, desc = WORKLOADS[args.workload]
    correction = "additive" if args.workload == "poisson2d-256" else "deflated"
    t0 = time.perf_counter()
    A, _, xyz, sym, bs = hs.generate(gen, **kw)
    n = A.shape[0]
    b0 = hs.random_rhs(n, args.seed)
    workers = min(16, os.cpu_count() or 1)
    st = hs.build_setup(A, b0, N, xyz=xyz, symmetric=sym, overlap=overlap, nev=nev, eig_workers=workers,
                        eig_tol=1e-6, correction=correction,
                        factorization="lu" if args.workload == "poisson2d-256" else "auto")
    t_setup = time.perf_counter() - t0
    log(f"setup {t_setup:.1f} s {st.times}; n0 = {st.coarse.n0}")

    M32 = device.HipTwoLevelSchwarz(st, "single", tasks_per_subdomain=args.tasks)
    ctx = M32.ctx
    stats = ctx.stats()
    log(f"factor entries fwd {stats.factor_entries_fwd} bwd {stats.factor_entries_bwd}, levels "
        f"{stats.levels_fwd_max}/{stats.levels_bwd_max}, bytes/local solve {stats.bytes_local_solve / 1e6:.1f} MB")

    if args.profile_applies > 0:
        r = ctx.alloc(n)
        z = ctx.alloc(n)
        r.upload(st.dec.b)
        for _ in range(args.profile_applies):
            ctx.apply_device(r, z)
        # calibration point of the HBM traffic counters: a plain copy of 2^25 doubles (512 MiB moved)
        big_a, big_b = ctx.alloc(1 << 25), ctx.alloc(1 << 25)
        for _ in range(3):
            ctx.copy_device(big_a, big_b)
        ctx.synchronize()
        log(f"profile run: {args.profile_applies} applies done")
        return

    M64 = device.HipTwoLevelSchwarz(st, "double", tasks_per_subdomain=args.tasks)
    ctx64 = M64.ctx
    nsteps = args.steps + args.warmup
    # right-hand sides in the renumbered order, resident on the device before timing
    B = np.stack([hs.random_rhs(n, args.seed + k)[st.dec.new_to_old] for k in range(nsteps)])
    dB = [ctx.alloc(n) for _ in range(nsteps)]
    dX = [ctx.alloc(n) for _ in range(nsteps)]
    for k in range(nsteps):
        dB[k].upload(B[k])
    zero = np.zeros(n)

    def run(c, k):
        dX[k].upload(zero)
        t = time.perf_counter()
        res = c.gmres_device(dB[k], dX[k], 1, restart, args.rtol, 2000)[0]
        return time.perf_counter() - t, res

    # ---- FP32 local factors / FP64 Krylov (the headline)
    for k in range(args.warmup):
        run(ctx, k)
    ctx.synchronize()
    ctx.timing(True)
    ctx.timing_read(reset=True)
    lat32, its32, conv32 = [], [], []
    t_start = time.perf_counter()
    for k in range(args.warmup, nsteps):
        dt, res = run(ctx, k)
        lat32.append(dt)
        its32.append(res.iterations)
        conv32.append(bool(res.converged))
    ctx.synchronize()
    t_total = time.perf_counter() - t_start
    launches, ms_local = ctx.timing_read(reset=True)
    ctx.timing(False)
    x32 = dX[args.warmup].download()

    # ---- plain FP64
    for k in range(args.warmup):
        run(ctx64, k)
    lat64, its64 = [], []
    t64s = time.perf_counter()
    for k in range(args.warmup, nsteps):
        dt, res = run(ctx64, k)
        lat64.append(dt)
        its64.append(res.iterations)
    t64 = time.perf_counter() - t64s
    x64 = dX[args.warmup].download()

    # ---- solution error against a tight FP64 solve
    xr, rr = ctx64.gmres(B[args.warmup], None, restart, 1e-12, 4000)
    err32 = float(np.linalg.norm(x32 - xr) / np.linalg.norm(xr))
    err64 = float(np.linalg.norm(x64 - xr) / np.linalg.norm(xr))

    # ---- batched persistent launch vs one launch per subdomain (FP32 applies)
    r = ctx.alloc(n)
    z = ctx.alloc(n)
    r.upload(B[0])
    mode_ms = {}
    for mode in (0, 1):
        ctx.set_launch_mode(mode)
        for _ in range(3):
            ctx.apply_device(r, z)
        ctx.synchronize()
        ctx.timing(True)
        ctx.timing_read(reset=True)
        t = time.perf_counter()
        for _ in range(20):
            ctx.apply_device(r, z)
        ctx.synchronize()
        wall = (time.perf_counter() - t) / 20
        nl, ms = ctx.timing_read(reset=True)
        ctx.timing(False)
        mode_ms[mode] = (ms / max(nl, 1), wall * 1e3)
    ctx.set_launch_mode(0)

    # ---- roofline of the dominant kernel (batched local solve, FP32 factors)
    avg_ms = ms_local / max(launches, 1)
    achieved = stats.bytes_local_solve / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_local_solve<float>", "launches": int(launches), "avg_us": round(avg_ms * 1e3, 2),
                "bytes_per_launch": int(stats.bytes_local_solve)}

    # ---- CPU baseline: the oracle (FP64 restatement) on the host cores, one solve
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        threads = args.cpu_threads or min(N, os.cpu_count() or 1, 16)
        O = orc.OracleSchwarz(st, "double", threads=threads)
        t = time.perf_counter()
        xo, ro = orc.gmres(O.spmv, O.apply, B[args.warmup], restart=restart, rtol=args.rtol, maxit=2000)
        tc = time.perf_counter() - t
        cpu = {"value": round(1.0 / tc, 4), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": f"1 GMRES({restart}) solve of the same workload with FP64 local factors, "
                         f"{N} subdomains on {threads} threads ({ro.iterations} its, {tc:.1f} s)",
               "iterations": ro.iterations,
               "x_rel_diff_vs_gpu_fp64": float(np.linalg.norm(xo - x64) / np.linalg.norm(x64))}

    solves_s = args.steps / t_total
    out = {
        "metric": "GMRES solves/s + p50/p99 latency to rel. residual 1e-8; FP32-local vs FP64 plain",
        "value": round(solves_s, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 local factors / f64 Krylov",
        "data": "synthetic (generated matrix, seeded uniform[-1,1] right-hand sides)",
        "config": {"workload": desc, "n": n, "subdomains": N, "overlap": overlap, "nev": nev, "n0": st.coarse.n0,
                   "restart": restart, "rtol": args.rtol, "parallelism": "single GPU, all subdomains resident"},
        "p50_ms": round(1e3 * float(np.percentile(lat32, 50)), 3),
        "p99_ms": round(1e3 * float(np.percentile(lat32, 99)), 3),
        "iterations_fp32": float(np.mean(its32)),
        "converged_fp32": all(conv32),
        "fp64_plain": {"value": round(args.steps / t64, 3), "unit": "solves/s",
                       "p50_ms": round(1e3 * float(np.percentile(lat64, 50)), 3),
                       "p99_ms": round(1e3 * float(np.percentile(lat64, 99)), 3),
                       "iterations": float(np.mean(its64))},
        "speedup_fp32_vs_fp64": round((args.steps / t_total) / (args.steps / t64), 3),
        "solution_rel_err_fp32": err32,
        "solution_rel_err_fp64": err64,
        "local_solve_us": {"batched_one_launch": round(mode_ms[0][0] * 1e3, 2),
                           "one_launch_per_subdomain": round(mode_ms[1][0] * 1e3, 2),
                           "apply_wall_us_batched": round(mode_ms[0][1] * 1e3, 2),
                           "apply_wall_us_per_subdomain": round(mode_ms[1][1] * 1e3, 2)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "setup_s": round(t_setup, 2),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
