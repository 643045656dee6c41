#!/usr/bin/env python3
"""Benchmark of the MI355X apply path: GMRES solves/s, p50/p99 latency to rel. residual 1e-8,
FP32 local factors / FP64 Krylov against the plain FP64 GPU path.

The headline value stores the local factors in FP32 and applies them with FP64 sums (the apply stays
linear to FP64 rounding, so the reference's GMRES takes the FP64 iteration count); the reference's
own single precision (solveIn<float>, float arithmetic) and flexible GMRES are reported beside it.

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

# the bubbly pressure sequence: coefficient 1 in the liquid and 1e-3 in the gas (rho_ratio 1e-3:
# beta = 1 / (alpha rho_g + 1 - alpha), rho_g = 1 / rho_ratio), 200 bubbles of radii uniform in
# 3-8 cells rising 1 cell per step (cfl 1), bubble seed 2024
BUBBLES_128 = dict(nx=128, nbubbles=200, rho_ratio=1e-3, rmin_cells=3.0, rmax_cells=8.0, cfl=1.0, seed=2024)

WORKLOADS = {
    # name: (generator, kwargs, subdomains, overlap, nev, restart, description)
    "poisson3d-64": ("poisson3d", dict(nx=64), 64, 2, 24, 50,
                     "3D Poisson 7-point 64^3, 64 subdomains, overlap 2, nev 24, GMRES(50), rtol 1e-8"),
    "poisson2d-256": ("poisson2d", dict(nx=256), 8, 1, 20, 30,
                      "2D Poisson 5-point 256^2, 8 subdomains, overlap 1, nev 20, additive, GMRES(30), rtol 1e-8"),
    "bubbles3d-128": ("bubbles3d", dict(BUBBLES_128), 256, 1, 8, 50,
                      "bubbly pressure 7-point 128^3 (coefficient 1 in the liquid, 1e-3 inside 200 seeded bubbles "
                      "of radii uniform in 3-8 cells), 256 subdomains, overlap 1, nev 8, GMRES(50), rtol 1e-8"),
    "highcontrast3d-96": ("highcontrast3d", dict(nx=96, seed=2024, contrast=1e6, channel_fraction=0.1), 128, 1, 32, 50,
                          "3D diffusion 7-point 96^3 with seeded axis-aligned channels of coefficient 1e6 (10 % of "
                          "the volume), 128 subdomains, overlap 1, nev 32, 16 right-hand sides solved together, "
                          "GMRES(50), rtol 1e-8"),
}
# right-hand sides solved together by default (a step solves them all)
DEFAULT_NRHS = {"highcontrast3d-96": 16}


COPY_COUNT = 1 << 25   # doubles of the calibration copy (256 MiB read + 256 MiB written)
PMC_APPLIES = 3        # applies of a counter pass
PROFILE_WARM = 2       # untimed applies before the timed ones in --profile-applies runs
LATENCY_SOLVES = 100   # solves of the latency pass (p50 / p99)
# arithmetic and factor-entry types of the local-solve kernels per precision of the local factors:
# "single" is the reference's solveIn<float> (float arithmetic, float entries), "mixed" FP32 entries
# with FP64 arithmetic, "double" FP64 throughout
KERNEL_TYPES = {"single": ("float", "float"), "mixed": ("double", "float"), "double": ("double", "double")}


def local_kernels(precision):
    """Kernel names of one apply's local solve as rocprofv3 reports them: the task-queue kernel's name
    also carries its workgroup size (k_local_solve<T, TF, 1024> or <T, TF, 512>); the dense tops of an
    apply run in k_dense_top<T, TF, NB> (NB column blocks per batch of loads); the level-synchronous
    kernels in k_ls_*<T>."""
    t, tf = KERNEL_TYPES[precision]
    return (f"k_ls_forward<{t}>", f"k_ls_back_partial<{t}>", f"k_ls_backward<{t}>", f"k_local_solve<{t}, {tf},",
            f"k_dense_top<{t}, {tf},")


def time_applies(ctx, r, z, count, warm=3):
    """(average ms of the local-solve launch set per apply from HIP events on its stream,
    wall ms per apply)"""
    for _ in range(warm):
        ctx.apply_device(r, z)
    ctx.synchronize()
    ctx.timing(True)
    ctx.timing_read(reset=True)
    t = time.perf_counter()
    for _ in range(count):
        ctx.apply_device(r, z)
    ctx.synchronize()
    wall = (time.perf_counter() - t) / count
    nl, ms = ctx.timing_read(reset=True)
    ctx.timing(False)
    return ms / max(nl, 1), wall * 1e3


def roofline(bytes_per_apply, avg_ms, applies, threads=None, precision="mixed"):
    """The dominant kernel's roofline: the local solve's launches of one apply (k_local_solve<T, TF,
    threads> for the subtree forward and backward sweeps, k_dense_top<T, TF> for the top tasks when
    they are all dense tops), timed together with HIP events on their stream; avg_us is their time
    per apply, bytes the algorithmic bytes of one apply's local solves."""
    achieved = bytes_per_apply / (avg_ms * 1e-3) / 1e9
    t, tf = KERNEL_TYPES[precision]
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": f"k_local_solve<{t}, {tf}, {threads or 1024}> + k_dense_top<{t}, {tf}, NB> (task queue, the launches of one apply)",
            "applies_timed": int(applies),
            "avg_us": round(avg_ms * 1e3, 2),
            "bytes_per_apply": int(bytes_per_apply)}


def measure_traffic(args):
    """HBM bytes of the local-solve launches of one apply from two rocprofv3 --pmc passes (FETCH_SIZE,
    then WRITE_SIZE) over a child run of this script with --profile-applies, each scaled by what the
    counter reads for the calibration copy of COPY_COUNT doubles. The child skips the coarse space (the
    local solve then reads no x0: 8 n bytes less than in the two-level apply). Returns (bytes, note)."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not found"
    kib = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="s2l_pmc_", dir="/tmp")
        cmd = [exe, "--pmc", counter, "-d", d, "-o", "run", "--output-format", "csv", "--", sys.executable,
               os.path.abspath(__file__), "--profile-applies", str(PMC_APPLIES), "--profile-one-level", "--setup-workers", "1",
               "--workload", args.workload, "--tasks", str(args.tasks), "--no-autotune",
               "--profile-precision", args.profile_precision] + \
              ([] if args.dense_top is None else ["--dense-top", str(args.dense_top)]) + \
              ([] if args.local_threads is None else ["--local-threads", str(args.local_threads)])
        log(f"traffic pass {counter}")
        try:
            p = subprocess.run(cmd, capture_output=True, text=True, timeout=300, cwd="/tmp",
                               env={**os.environ, "TMPDIR": "/tmp"})
        except subprocess.TimeoutExpired:
            return None, f"{counter} pass timed out"
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if p.returncode != 0 or not files:
            return None, f"{counter} pass failed (exit {p.returncode})"
        local, copy = 0.0, []
        for f in files:
            for row in csv.DictReader(open(f)):
                if row.get("Counter_Name") != counter:
                    continue
                name = row["Kernel_Name"]
                if any(k in name for k in local_kernels(args.profile_precision)):
                    local += float(row["Counter_Value"])
                elif "k_copy" in name:
                    copy.append(float(row["Counter_Value"]))
        shutil.rmtree(d, ignore_errors=True)
        if local == 0.0 or not copy:
            return None, f"{counter}: kernels not found in the counter output"
        # the child runs PMC_APPLIES + PROFILE_WARM applies
        kib[counter] = (local / (PMC_APPLIES + PROFILE_WARM), float(np.mean(copy)))
    # the copy moves COPY_COUNT * 8 bytes each way
    ref_kib = COPY_COUNT * 8 / 1024.0
    ff = ref_kib / kib["FETCH_SIZE"][1]
    wf = ref_kib / kib["WRITE_SIZE"][1]
    traffic = (kib["FETCH_SIZE"][0] * ff + kib["WRITE_SIZE"][0] * wf) * 1024.0
    note = (f"FETCH_SIZE {kib['FETCH_SIZE'][0]:.0f} KiB x {ff:.3f}, WRITE_SIZE {kib['WRITE_SIZE'][0]:.0f} KiB x "
            f"{wf:.3f} per apply's local solve (factors from a copy of {COPY_COUNT} doubles)")
    return int(traffic), note


def _us(ms):
    """milliseconds -> microseconds rounded, None when not measured"""
    return None if ms != ms else round(ms * 1e3, 2)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def serving_main(args):
    """The serving mix: a seeded Poisson-arrival stream of requests over 6 resident families (2D
    5-point, 3D 7-point, 2D anisotropic 1e-3; n log-uniform in [1e4, 2e6]), 1-8 right-hand sides
    each, GMRES(50) to rtol with FP32 local factors, admitted by device footprint. Passes: (1) every
    request back to back on the plain device solve (the FIFO service times and capacity); (2) every
    request alone through its family's slot solver (the reference bits); (3) the wall-clock server
    with continuous batching at 0.7x and 1.5x the FIFO capacity, and the wall-clock FIFO server at
    0.7x. Latencies are measured on the wall clock from arrival to the solution on the host."""
    from amp_amd import serving as sv
    fams, reqs = sv.draw_mix(args.requests, 6, (1e4, 2e6), 8, seed=args.seed)
    log("families:", ", ".join(f"{f.kind} nx={f.nx} (n={f.n})" for f in fams))
    budget = int(args.serving_budget_gb * 2**30)
    resident = sv.ResidentSet(budget, lambda k: sv.build_family(fams[k], precision="single"))
    t0 = time.perf_counter()
    for k in range(len(fams)):  # a server admits its resident families before taking requests
        M = resident.get(k)
        sv.slot_solver(M, args.slots, 50, args.rtol, 1000)
        resident.fit(k)
        log(f"admitted family {k}: {resident.footprint(M) / 2**20:.1f} MiB on the device (slot solver included)")
    admit_s = time.perf_counter() - t0
    every = max(1, len(reqs) // 10)
    recs = sv.serve(fams, reqs, resident, restart=50, rtol=args.rtol,
                    log=lambda r: log(f"fifo back to back: request {r['id']} {1e3 * r['service']:.2f} ms")
                    if r["id"] % every == 0 else None)
    svc = np.array([r["service"] for r in recs])
    capacity_rps = 1.0 / float(svc.mean())  # requests per second of one FIFO server
    alone = sv.serve_alone(fams, reqs, resident, slots=args.slots, restart=50, rtol=args.rtol,
                           log=lambda r: log(f"alone: request {r['id']}") if r["id"] % every == 0 else None)
    runs = {}
    for mode, load in (("batched", 0.7), ("batched", 1.5), ("fifo", 0.7)):
        scale = 1.0 / (load * capacity_rps)   # seconds per unit of the seeded inter-arrival times
        live = sv.serve_live(fams, reqs, resident, scale=scale, mode=mode, slots=args.slots, restart=50,
                             rtol=args.rtol, log=log)
        summ = {k: round(v, 3) for k, v in sv.live_summary(live).items()}
        summ["offered_requests_per_s"] = round(load * capacity_rps, 2)
        summ["all_converged"] = all(r["converged"] for r in live)
        summ["max_true_rel_residual"] = max(r["true_rel_residual"] for r in live)
        if mode == "batched":
            same = sv.same_as_alone(alone, live)
            summ["results_same_as_alone"] = same["differing"] == 0
            summ["requests_differing_from_alone"] = same["differing"]
            summ["differing_ids"] = same["differing_ids"]
        runs[f"{mode}_load_{load}"] = summ
        log(f"live {mode} at load {load}: {summ}")
    head = runs["batched_load_0.7"]
    out = {"metric": "serving mix: solves/s + p50/p99 request latency on the wall clock (continuous batching, "
                     "70 % of the FIFO capacity offered)",
           "value": head["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": len(reqs), "warmup": 0,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 local factors / f64 Krylov",
           "data": "synthetic (generated matrices, seeded right-hand sides, seeded arrivals)",
           "config": {"workload": f"serving mix: {len(reqs)} requests, 6 resident families "
                                  f"(2D 5-point, 3D 7-point, 2D anisotropic 1e-3; n log-uniform in [1e4, 2e6]), "
                                  f"1-8 right-hand sides each, GMRES(50) to {args.rtol}",
                      "families": [f"{f.kind}:{f.nx}" for f in fams], "slots_per_family": args.slots,
                      "parallelism": "single GPU, one context, stream and host thread per family"},
           "p50_ms": head["p50_ms"], "p99_ms": head["p99_ms"],
           "results_same_as_alone": all(v.get("results_same_as_alone", True) for v in runs.values()),
           "fifo_back_to_back": {"mean_service_ms": round(1e3 * float(svc.mean()), 3),
                                 "capacity_requests_per_s": round(capacity_rps, 2),
                                 "capacity_solves_per_s": round(sum(r["nrhs"] for r in recs) / float(svc.sum()), 2)},
           "live": runs,
           "admission": {"budget_bytes": budget, "builds": resident.builds, "evictions": resident.evictions,
                         "admit_s": round(admit_s, 1), "resident_bytes": int(resident.used())}}
    print(json.dumps(out), flush=True)


def reuse_main(args):
    """The setup-reuse sequence: bubbly pressure matrices (coefficient 1e-3 inside 200 bubbles of radii
    3-8 cells rising one cell per step) at pseudo-times 0..steps-1, 256 subdomains at 128^3, overlap 1,
    nev 8, deflated; the first solve builds the preconditioner, every later step refactorizes (on the
    device: new local factors and A_00 with the same coarse basis) and its second corrector reuses;
    GMRES(50) and BiCGStab, FP32 local factors. Each Krylov method's sequence runs in a child process
    of its own (a fresh process may fork the eigensolver pool of its rebuild)."""
    import subprocess
    kw = dict(BUBBLES_128, nx=args.seq_nx)
    if args.seq_nx != 128:  # the same bubble density on a smaller grid
        kw["nbubbles"] = max(1, round(200 * (args.seq_nx / 128.0) ** 3))
    nparts = max(4, (args.seq_nx ** 3) // 8192)
    if args.reuse_krylov:
        from amp_amd.timeloop import run_time_loop
        krylov = args.reuse_krylov
        recs = run_time_loop("bubbles3d", kw, nparts=nparts, steps=args.seq_steps, correctors=2, reuse=args.reuse_policy,
                             precision="single", restart=50, rtol=args.rtol, nev=8, krylov=krylov,
                             eig_workers=args.setup_workers or 16, device_refactor=not args.host_refactor,
                             eig_backend="host" if args.host_eigensolver else "device",
                             log=lambda r: log(f"{krylov} step {r['step']} corrector {r['corrector']}: {r['level']} "
                                               f"{r['t_update']:.3f} s + {1e3 * r['t_solve']:.1f} ms, "
                                               f"{r['iterations']} its"))
        upd = sum(r["t_update"] for r in recs)
        sol = sum(r["t_solve"] for r in recs)
        refac = [r["t_update"] for r in recs if r["level"] == "refactorize"]
        print(json.dumps({"solves_per_s": round(len(recs) / (upd + sol), 3), "solves": len(recs),
                          "solve_only_per_s": round(len(recs) / sol, 2),
                          "mean_iterations": round(float(np.mean([r["iterations"] for r in recs])), 2),
                          "first_iterations": recs[0]["iterations"], "last_iterations": recs[-1]["iterations"],
                          "rebuild_s": round(recs[0]["t_update"], 2),
                          "refactorize_s_mean": round(float(np.mean(refac)), 3) if refac else None,
                          "refactorize_s_first": round(refac[0], 3) if refac else None,
                          "refactorize_s_median": round(float(np.median(refac)), 3) if refac else None,
                          "refactorize_on": "host" if args.host_refactor else "device",
                          "eigensolver_on": "host" if args.host_eigensolver else "device",
                          "solve_ms_mean": round(1e3 * sol / len(recs), 2),
                          "all_converged": all(r["converged"] for r in recs),
                          "max_true_rel_residual": max(r["true_rel_residual"] for r in recs),
                          "levels": "".join({"rebuild": "2", "refactorize": "1", "reuse": "0"}[r["level"]] for r in recs)}),
              flush=True)
        return
    out = {"metric": "setup-reuse sequence: solves/s over the sequence (updates + solves)", "unit": "solves/s",
           "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 local factors / f64 Krylov", "data": "synthetic (generated matrices, seeded right-hand sides)",
           "config": {"workload": f"bubbly pressure 7-point {args.seq_nx}^3, coefficient 1e-3 inside {kw['nbubbles']} "
                                  f"bubbles of radii 3-8 cells rising 1 cell per step, "
                                  f"{args.seq_steps} steps x 2 correctors, {nparts} subdomains, overlap 1, nev 8, "
                                  f"deflated, reuse policy '{args.reuse_policy}', refactorizations on the "
                                  f"{'host' if args.host_refactor else 'device'}, local eigenproblems of the rebuild "
                                  f"on the {'host' if args.host_eigensolver else 'device'}",
                      "parallelism": "single GPU, all subdomains resident"}}
    for krylov in ("gmres", "bicgstab"):
        cmd = [sys.executable, os.path.abspath(__file__), "--reuse-sequence", "--reuse-krylov", krylov, "--seq-nx",
               str(args.seq_nx), "--seq-steps", str(args.seq_steps), "--reuse-policy", args.reuse_policy, "--rtol",
               str(args.rtol)] + (["--host-refactor"] if args.host_refactor else []) + \
              (["--host-eigensolver"] if args.host_eigensolver else [])
        p = subprocess.run(cmd, stdout=subprocess.PIPE, text=True)
        if p.returncode != 0:
            raise SystemExit(p.returncode)
        out[krylov] = json.loads(p.stdout.strip().splitlines()[-1])
    out["value"] = out["gmres"]["solves_per_s"]
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="poisson3d-64", choices=sorted(WORKLOADS))
    ap.add_argument("--nrhs", type=int, default=0,
                    help="right-hand sides per GMRES solve (0 = the workload's: 16 for highcontrast3d-96, else 1); "
                         "a step solves them together, value counts every right-hand side")
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--graphs", type=int, default=0, choices=(0, 1),
                    help="replay each GMRES iteration's apply + SpMV from a captured HIP graph (s2l_set_graphs)")
    ap.add_argument("--tasks", type=int, default=0,
                    help="tasks per subdomain of the batched local solve (0 = compute units / subdomains)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default local-solve schedule instead of the autotuner's choice (s2l_autotune_local)")
    ap.add_argument("--local-threads", type=int, default=None, choices=(512, 1024),
                    help="threads per local-solve workgroup (s2l_set_local_threads; default: the library's, 1024)")
    ap.add_argument("--dense-top", type=int, default=None,
                    help="widest skeleton replaced by a dense top (s2l_set_dense_top; default: the library's, 0 = none)")
    ap.add_argument("--leaf", type=int, default=0,
                    help="largest nested-dissection leaf of the local factorizations (0 = the setup's default, 16)")
    ap.add_argument("--local-mode", type=int, default=None, choices=(0, 1, 2),
                    help="schedule of the local solves (s2l_set_local_launch_mode; default: the library's)")
    ap.add_argument("--profile-applies", type=int, default=0,
                    help="only run this many FP32 applies after setup (for rocprofv3 runs)")
    ap.add_argument("--no-launch-compare", action="store_true",
                    help="skip the comparison with one launch per subdomain")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) that measure the HBM traffic "
                         "of the local-solve launch")
    ap.add_argument("--profile-precisions", action="store_true",
                    help="with --profile-applies: also time the FP64 and FP32-storage / FP64-arithmetic local solves")
    ap.add_argument("--profile-precision", default="mixed", choices=("single", "mixed", "double"),
                    help="with --profile-applies: precision of the local factors timed (default: the headline's, "
                         "FP32 entries with FP64 arithmetic)")
    ap.add_argument("--profile-one-level", action="store_true",
                    help="with --profile-applies: skip the coarse space (nev 0), for the counter passes")
    ap.add_argument("--setup-workers", type=int, default=0,
                    help="processes of the host eigensolves (0 = min(16, cores); 1 = no fork, needed under "
                         "rocprofv3 --pmc, which initialises the GPU before the program starts)")
    ap.add_argument("--serving-mix", action="store_true",
                    help="run the serving mix (Poisson arrivals over 6 resident families) instead of a workload")
    ap.add_argument("--requests", type=int, default=2000, help="requests of the serving mix")
    ap.add_argument("--serving-budget-gb", type=float, default=64.0,
                    help="device memory the serving mix may keep resident (GiB)")
    ap.add_argument("--slots", type=int, default=16,
                    help="right-hand-side slots of every family's solver in the serving mix")
    ap.add_argument("--reuse-sequence", action="store_true",
                    help="run the setup-reuse sequence (bubbly pressure, refactorized every step) instead")
    ap.add_argument("--seq-steps", type=int, default=20, help="steps of the reuse sequence")
    ap.add_argument("--seq-nx", type=int, default=128, help="grid of the reuse sequence (nx^3)")
    ap.add_argument("--reuse-policy", default="refactor", choices=("off", "refactor", "auto"),
                    help="reuse policy of the sequence")
    ap.add_argument("--reuse-krylov", default="", choices=("", "gmres", "bicgstab"),
                    help="(child run) the sequence of one Krylov method")
    ap.add_argument("--host-eigensolver", action="store_true",
                    help="solve the local eigenproblems of a rebuild on the host (ARPACK) instead of the device")
    ap.add_argument("--host-refactor", action="store_true",
                    help="refactorize on the host (host factorization, A_00 on the host, upload) instead of the device")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the apply path runs on one GPU; --gpus {args.gpus} is not supported (it would "
              f"shard by subdomain: a range of subdomains per GPU, a halo exchange per apply, an "
              f"all-gather of the coarse right-hand side and three all-reduces per GMRES iteration)",
              file=sys.stderr)
        sys.exit(2)
    _load()
    from amp_amd import device, hostsetup as hs

    if args.serving_mix:
        serving_main(args)
        return
    if args.reuse_sequence:
        reuse_main(args)
        return
    gen, kw, N, overlap, nev, restart, desc = WORKLOADS[args.workload]
    nrhs = args.nrhs or DEFAULT_NRHS.get(args.workload, 1)
    if args.profile_applies > 0 and args.profile_one_level:
        nev = 0
    correction = "additive" if args.workload == "poisson2d-256" else "deflated"
    t0 = time.perf_counter()
    A, _, xyz, sym, bs = hs.generate(gen, **kw)
    n = A.shape[0]
    b0 = hs.random_rhs(n, args.seed)
    workers = args.setup_workers or min(16, os.cpu_count() or 1)
    st = hs.build_setup(A, b0, N, xyz=xyz, symmetric=sym, overlap=overlap, nev=nev, eig_workers=workers,
                        eig_tol=1e-6, correction=correction,
                        factorization="lu" if args.workload == "poisson2d-256" else "auto", leaf=args.leaf,
                        eig_backend="device" if (sym and args.workload != "poisson2d-256" and not args.host_eigensolver)
                        else "host")
    t_setup = time.perf_counter() - t0
    log(f"setup {t_setup:.1f} s {st.times}; n0 = {st.coarse.n0}")

    def make(prec):
        return device.HipTwoLevelSchwarz(st, prec, tasks_per_subdomain=args.tasks, dense_top=args.dense_top,
                                         local_threads=args.local_threads)

    if args.profile_applies > 0:
        # only applies of one precision (every local-solve launch set does the same work), event-timed
        Mp = make(args.profile_precision)
        ctx = Mp.ctx
        ctx.set_graphs(False)
        if args.local_mode is not None:
            ctx.set_launch_mode(args.local_mode)
        stats = ctx.stats()
        log(f"factor entries fwd {stats.factor_entries_fwd} bwd {stats.factor_entries_bwd}, levels "
            f"{stats.levels_fwd_max}/{stats.levels_bwd_max}, bytes/local solve {stats.bytes_local_solve / 1e6:.1f} MB")
        r = ctx.alloc(n)
        z = ctx.alloc(n)
        r.upload(st.dec.b)
        avg_ms, _ = time_applies(ctx, r, z, args.profile_applies, warm=PROFILE_WARM)
        # calibration point of the HBM traffic counters: a plain copy of 2^25 doubles (512 MiB moved)
        big_a, big_b = ctx.alloc(COPY_COUNT), ctx.alloc(COPY_COUNT)
        for _ in range(3):
            ctx.copy_device(big_a, big_b)
        ctx.synchronize()
        out = {"mode": "profile", "precision": args.profile_precision, "applies": args.profile_applies,
               "leaf": args.leaf, "tasks": args.tasks,
               "levels": [stats.levels_fwd_max, stats.levels_bwd_max], "supernodes": stats.supernodes,
               "lds_slots": {"cap_per_task": stats.lds_slot_cap, "max_task_slots": stats.max_task_slots,
                             "kept_in_lds": stats.lds_slots},
               "roofline": roofline(stats.bytes_local_solve, avg_ms, args.profile_applies, args.local_threads,
                                    args.profile_precision)}
        if args.profile_precisions:
            for prec in ("single", "double", "mixed"):
                if prec == args.profile_precision:
                    continue
                Mq = make(prec)
                Mq.ctx.set_graphs(False)
                sq = Mq.ctx.stats()
                ms_q, _ = time_applies(Mq.ctx, r, z, args.profile_applies, warm=PROFILE_WARM)
                out[prec] = roofline(sq.bytes_local_solve, ms_q, args.profile_applies, args.local_threads, prec)
                out[prec]["lds_slots"] = {"cap_per_task": sq.lds_slot_cap, "max_task_slots": sq.max_task_slots,
                                          "kept_in_lds": sq.lds_slots}
        print(json.dumps(out), flush=True)
        return

    # the headline: FP32 factor entries applied in FP64 arithmetic (precision "mixed"); beside it the
    # reference's single precision (solveIn<float>: float arithmetic) and the plain FP64 path
    tune_on = (not args.no_autotune and args.tasks == 0 and args.dense_top is None and args.local_threads is None)
    tune = {}
    ctxs = {}
    for key, prec in (("mixed", "mixed"), ("fp32", "single"), ("fp64", "double")):
        Mq = make(prec)
        if tune_on:
            t0 = time.perf_counter()
            cands, best = Mq.autotune(nrhs, 20)
            tune[key] = {"candidates": cands, "kept": best, "s": round(time.perf_counter() - t0, 2)}
            log(f"autotune {key}: kept {cands[best]} of {len(cands)} candidates")
        Mq.ctx.set_graphs(args.graphs != 0)
        if args.local_mode is not None:
            Mq.ctx.set_launch_mode(args.local_mode)
        ctxs[key] = Mq
    if tune_on:
        # the traffic passes measure the schedule kept for the headline
        kept = tune["mixed"]["candidates"][tune["mixed"]["kept"]]
        args.tasks, args.dense_top, args.local_threads = kept["tasks"], kept["dense_top"], kept["threads"]
    args.profile_precision = "mixed"
    ctxm, ctx, ctx64 = ctxs["mixed"].ctx, ctxs["fp32"].ctx, ctxs["fp64"].ctx
    stats = ctxm.stats()
    log(f"factor entries fwd {stats.factor_entries_fwd} bwd {stats.factor_entries_bwd}, levels "
        f"{stats.levels_fwd_max}/{stats.levels_bwd_max}, bytes/local solve {stats.bytes_local_solve / 1e6:.1f} MB")
    nsteps = args.steps + args.warmup
    # right-hand sides in the renumbered order, resident on the device before timing; step k solves
    # the nrhs right-hand sides k * nrhs .. k * nrhs + nrhs - 1 together
    B = np.stack([hs.random_rhs(n, args.seed + k)[st.dec.new_to_old] for k in range(nsteps * nrhs)])
    dB = [ctx.alloc(n * nrhs) for _ in range(nsteps)]
    dX = [ctx.alloc(n * nrhs) for _ in range(nsteps)]
    for k in range(nsteps):
        dB[k].upload(B[k * nrhs:(k + 1) * nrhs].ravel())
    zero = np.zeros(n * nrhs)
    # the initial guess x0 = 0 comes from a resident buffer too: a device copy on the solving
    # context's stream, not an upload from the host, inside the timed region
    dZero = ctx.alloc(n * nrhs)
    dZero.upload(zero)

    class Batch:
        """Results of one step: mean iterations, all converged."""
        def __init__(self, rs):
            self.iterations = float(np.mean([r.iterations for r in rs]))
            self.converged = all(bool(r.converged) for r in rs)

    def run(c, k, flexible=False):
        c.copy_device(dZero, dX[k])
        t = time.perf_counter()
        res = c.gmres_device(dB[k], dX[k], nrhs, restart, args.rtol, 2000, flexible=flexible)
        return time.perf_counter() - t, Batch(res)

    def timed(c, flexible=False):
        """warm-up steps, then the timed steps; (seconds, latencies, iterations, all converged, x of
        the first timed step)"""
        for k in range(args.warmup):
            run(c, k, flexible)
        c.synchronize()
        lat, its, conv = [], [], []
        t0 = time.perf_counter()
        for k in range(args.warmup, nsteps):
            dt, res = run(c, k, flexible)
            lat.append(dt)
            its.append(res.iterations)
            conv.append(bool(res.converged))
        c.synchronize()
        return time.perf_counter() - t0, lat, its, all(conv), dX[args.warmup].download()[:n]

    # ---- the headline: FP32 factor entries, FP64 arithmetic, the reference's GMRES
    tm, latm, itsm, convm, xm = timed(ctxm)
    # ---- plain FP64
    t64, lat64, its64, conv64, x64 = timed(ctx64)
    # ---- the reference's single precision (solveIn<float>), the reference's GMRES
    t32, lat32, its32, conv32, x32 = timed(ctx)
    # ---- FP32 arithmetic with flexible GMRES (M^{-1} v_j kept: no second cycle, no final apply)
    tf, latf, itsf, convf, xf = timed(ctx, True)

    # ---- latency distribution: p50 / p99 over a separate pass of at least LATENCY_SOLVES solves
    # (cycling over the resident right-hand sides)
    nlat = max(LATENCY_SOLVES, args.steps)
    lat_pass = {}
    for name_, c_ in (("mixed", ctxm), ("fp64", ctx64), ("fp32", ctx)):
        lat = [run(c_, args.warmup + q % args.steps)[0] for q in range(nlat)]
        lat_pass[name_] = (1e3 * float(np.percentile(lat, 50)), 1e3 * float(np.percentile(lat, 99)))

    # ---- per-phase times of one headline solve under the reference's timer names (events; a separate
    # pass: the events and the one-by-one launches are not in the timed region)
    ctxm.phase_timing(True)
    ctxm.phase_timing_read(reset=True)
    nph = min(args.steps, 10)
    for q in range(nph):
        run(ctxm, args.warmup + q)
    ph = ctxm.phase_timing_read(reset=True)
    ctxm.phase_timing(False)
    phases = {k: round(v[0] / (nph * nrhs), 4) for k, v in ph.items() if v[1] > 0}

    # ---- solution error against a tight FP64 solve
    xr, rr = ctx64.gmres(B[args.warmup * nrhs], None, restart, 1e-12, 4000)
    err = {k: float(np.linalg.norm(v - xr) / np.linalg.norm(xr)) for k, v in
           (("mixed", xm), ("fp64", x64), ("fp32", x32), ("fgmres", xf))}

    # ---- local-solve schedules (dedicated loops: every timed local-solve launch does the full work,
    # unlike the GMRES loop where a launch after convergence is skipped by its mask): the headline's
    # task queue, and on the FP32-arithmetic context 0 = task queue in one launch per depth group,
    # 2 = level-synchronous (its factors are stored in the arithmetic type), 1 = one launch per
    # subdomain
    r = ctx.alloc(n)
    z = ctx.alloc(n)
    r.upload(B[0])
    nan = (float("nan"), float("nan"))
    mode_ms = {2: nan, 0: nan, 1: nan}
    for mode in ((0, 2) if args.no_launch_compare else (0, 2, 1)):
        ctx.set_launch_mode(mode)
        mode_ms[mode] = time_applies(ctx, r, z, 20, warm=3)
    ctx.set_launch_mode(0)
    mixed_ms = time_applies(ctxm, r, z, 20, warm=3)
    f64_ms = time_applies(ctx64, r, z, 20, warm=3)

    # ---- roofline of the dominant kernel (the headline's batched local solve)
    roof = roofline(stats.bytes_local_solve, mixed_ms[0], 20, args.local_threads, "mixed")
    roof["lds_slots"] = {"cap_per_task": stats.lds_slot_cap, "max_task_slots": stats.max_task_slots,
                         "kept_in_lds": stats.lds_slots}
    if not args.no_traffic:
        roof["traffic"], roof["traffic_note"] = measure_traffic(args)

    # ---- CPU baseline: the oracle (FP64 restatement) on the host cores, one solve
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        threads = args.cpu_threads or min(N, os.cpu_count() or 1, 16)
        O = orc.OracleSchwarz(st, "double", threads=threads)
        t = time.perf_counter()
        xo, ro = orc.gmres(O.spmv, O.apply, B[args.warmup * nrhs], restart=restart, rtol=args.rtol, maxit=2000)
        tc = time.perf_counter() - t
        cpu = {"value": round(1.0 / tc, 4), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": f"1 GMRES({restart}) solve of the same workload with FP64 local factors, "
                         f"{N} subdomains on {threads} threads ({ro.iterations} its, {tc:.1f} s)",
               "cores_rule": "min(subdomains, 16): one subdomain per thread as the reference's one rank per "
                             "subdomain, capped at the 16 host cores one GPU is given",
               "iterations": ro.iterations,
               "x_rel_diff_vs_gpu_fp64": float(np.linalg.norm(xo - x64) / np.linalg.norm(x64))}

    sps = lambda t: round(args.steps * nrhs / t, 3)  # noqa: E731  (solves per second of a timed pass)
    pct = lambda lat, q: round(1e3 * float(np.percentile(lat, q)), 3)  # noqa: E731
    out = {
        "metric": "GMRES solves/s + p50/p99 latency to rel. residual 1e-8; FP32-local vs FP64 plain",
        "value": sps(tm),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * tm / args.steps, 3),
        "rhs_per_step": nrhs,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 local factor entries with f64 arithmetic / f64 Krylov",
        "data": "synthetic (generated matrix, seeded uniform[-1,1] right-hand sides)",
        "config": {"workload": desc, "n": n, "subdomains": N, "overlap": overlap, "nev": nev, "n0": st.coarse.n0,
                   "restart": restart, "rtol": args.rtol, "krylov": "the reference's right-preconditioned GMRES(m)",
                   "parallelism": "single GPU, all subdomains resident"},
        "p50_ms": round(lat_pass["mixed"][0], 3),
        "p99_ms": round(lat_pass["mixed"][1], 3),
        "latency_solves": nlat,
        "phases_ms_per_solve": phases,
        "iterations": float(np.mean(itsm)),
        "converged": convm,
        "solution_rel_err": err["mixed"],
        "iterations_vs_fp64": round(float(np.mean(itsm)) / float(np.mean(its64)), 4),
        "speedup_vs_fp64": round(t64 / tm, 3),
        "fp64_plain": {"value": sps(t64), "unit": "solves/s",
                       "p50_ms": round(lat_pass["fp64"][0], 3), "p99_ms": round(lat_pass["fp64"][1], 3),
                       "iterations": float(np.mean(its64)), "converged": conv64, "solution_rel_err": err["fp64"],
                       "local_solve_us": _us(f64_ms[0])},
        "fp32_arith": {"value": sps(t32), "unit": "solves/s",
                       "p50_ms": round(lat_pass["fp32"][0], 3), "p99_ms": round(lat_pass["fp32"][1], 3),
                       "iterations": float(np.mean(its32)), "converged": conv32, "solution_rel_err": err["fp32"],
                       "speedup_vs_fp64": round(t64 / t32, 3), "local_solve_us": _us(mode_ms[0][0]),
                       "note": "the reference's single precision (solveIn<float>: right-hand side rounded to float, "
                               "both sweeps in float arithmetic), the reference's GMRES: the apply is not linear "
                               "to FP64 rounding, so the true residual at the end of the first cycle misses the "
                               "10 % slack and a second cycle runs"},
        "fgmres_fp32": {"value": sps(tf), "unit": "solves/s", "p50_ms": pct(latf, 50), "p99_ms": pct(latf, 99),
                        "iterations": float(np.mean(itsf)), "converged": convf, "solution_rel_err": err["fgmres"],
                        "speedup_vs_fp64": round(t64 / tf, 3),
                        "note": "flexible GMRES(m) with FP32 arithmetic: the basis M^{-1} v_j is stored (an option "
                                "of s2l_krylov_opts without reference counterpart)"},
        "headline_note": "value: FP32 factor entries (the bytes of the FP32 path) applied with FP64 sums and vector "
                         "(precision 2 of s2l_finalize_local, no reference counterpart), the reference's GMRES; the "
                         "reference's own single precision is fp32_arith",
        "local_solve_us": {"headline": _us(mixed_ms[0]),
                           "fp32_level_synchronous": _us(mode_ms[2][0]),
                           "fp32_task_queue_per_depth_group": _us(mode_ms[0][0]),
                           "fp32_one_launch_per_subdomain": _us(mode_ms[1][0]),
                           "fp32_apply_wall_us_level_synchronous": _us(mode_ms[2][1]),
                           "fp32_apply_wall_us_task_queue": _us(mode_ms[0][1]),
                           "fp32_apply_wall_us_per_subdomain": _us(mode_ms[1][1])},
        "roofline": roof,
        "cpu_baseline": cpu,
        "autotune": tune or None,
        "setup_s": round(t_setup, 2),
        "setup_stages_s": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.times.items()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
