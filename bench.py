"""Benchmark: time to an FP64-checked relative residual of 1e-10 for the
mixed-precision single-kernel CG vs the plain FP64 single-kernel CG, and the
HBM bandwidth of the stand-alone tiled SpMV, on one MI355X.

Workload: 3D 7-point Laplacian on a 128^3 grid (n = 2,097,152,
nnz = 14,581,760, 880,640 16x16 tiles), diagonal 6*10^(3U) with U ~ U[0,1)
from seed 2, b ~ U[-1,1] from seed 2, x0 = 0, tolerance 1e-10.

A step is one complete solve (x reset to 0, one persistent-kernel launch, to
the tolerance) with the matrix and b already resident in HBM; the mixed and
FP64 solves alternate step by step (each timed on its own).  Every solve call
returns only after an event synchronisation on its stream and the copies of
its report, so each timed step is bracketed by device synchronisation.  No
hipDeviceSynchronize is issued between the warm-up and the timed steps: on
this runtime the launch after one takes about 6 ms longer.  `value` is the
mixed-precision solve time in ms (lower is better).  Also reported: the FP64
single-kernel CG time on the same matrix, the speedup, iterations, the FP64
residual recomputed on the CPU, the per-tile precision histogram, the SpMV
roofline (average launch duration of 50 back-to-back launches between two
HIP events, median of 5 such runs, after 10 warm-ups) and a CPU baseline
timed on this host.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--policy learned|errbound|rule|fp64]

The other workloads the library is measured on print one JSON line each
(`config.workload` names them); none replaces the line above:

  --workload convdiff100   BiCGSTAB, 3D 7-point upwind convection-diffusion
                           100^3 (cell Peclet 0.5, velocity (1, 0.5, 0.25)),
                           b = A x*, x* ~ U[-1,1] seed 3: FP64 vs mixed policies
  --workload knn512        batched CG: 512 kNN-graph Laplacians (k = 8, shift
                           1e-2, n ~ U[1k, 64k] seed 4) x 8 right-hand sides in
                           one hipGraph-captured persistent launch
  --workload powerlaw4m    CG on a power-law graph Laplacian (Chung-Lu,
                           exponent 2.5, n = 4M, ~60M nnz, shift 1e-3):
                           learned vs rule vs FP64
  --workload powerlaw1m    the same graph model at n = 1M, 15M nnz

The learned policy uses the MLP shipped with the package (policy_mlp.f32,
trained on the GPU by millefeuille.policy.train_default).
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mille-feuille_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _line(metric, value, unit, args, ms, config, **extra):
    d = {"metric": metric, "value": round(value, 3), "unit": unit, "n_gpus": 1,
         "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
         "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
         "dtype": "f64 (scalars, x, residual, accumulation); f32 products of fp32/fp16/int8 tile payloads",
         "data": "synthetic (seeded generator)", "config": config}
    d.update(extra)
    print(json.dumps(d), flush=True)


# Vector bytes per row and iteration besides the SpMV formula's x read and y
# write: CG's update reads r, p, s, x and writes p, s, x, r, its SpMV epilogue
# reads r; BiCGSTAB's five vector passes (v and r^ in the first SpMV's
# epilogue, s = r - alpha v, s in the second epilogue, x and r, p).  FP32 work
# vectors under the mixed policies, x always FP64.
_VEC_BYTES = {("cg", False): 72, ("cg", True): 44, ("bicgstab", False): 128, ("bicgstab", True): 72}


def _iteration_roofline(tm, res, method, levels=None):
    """Modeled HBM bytes per iteration (the SpMV formula at the tile levels
    the mixed solve ran with -- `levels`, else the handle's after the solve --
    or all FP64, plus the vector passes) and the rate they imply at each
    solve's measured kernel time per iteration."""
    import millefeuille as mf
    spmvs = 2 if method == "bicgstab" else 1
    out = {}
    for pol, r in res.items():
        mixed = pol != "fp64" and not r["fell_back"]
        lv = (tm.get_precision() if levels is None else levels) if mixed else None
        b_spmv = mf.spmv_bytes(tm, levels=lv, x_bytes=4 if mixed else 8)
        b_it = spmvs * b_spmv + _VEC_BYTES[(method, mixed)] * tm.rowA
        us = 1e3 * r["kernel_ms"] / max(r["iterations"], 1)
        gbs = b_it / (us * 1e-6) / 1e9
        out[pol] = {"bytes_per_iteration": int(b_it), "us_per_iteration": round(us, 2),
                    "achieved_gbs": round(gbs, 1), "frac_of_8tbs": round(gbs / HBM_PEAK_GBS, 4),
                    "levels": "per tile, after the solve" if mixed else "all FP64"}
    return out


def _timed_solves(tm, solve, policies, args, n, rp, ci, va, b, dx):
    """Warm-up, then args.steps solves per policy alternated step by step;
    mean wall ms per solve, the last report and the FP64 relative residual of
    its x recomputed on the CPU (oracle CSR product)."""
    import millefeuille as mf
    import oracle
    for _ in range(args.warmup):
        for pol in policies:
            r = solve(pol)
            log(f"warm-up {pol}: {r['iterations']} iterations, {r['time_ms']:.1f} ms")
    mf.synchronize()
    out = {pol: [] for pol in policies}
    reps = {}
    rel = {}
    for _ in range(args.steps):
        for pol in policies:
            t1 = time.perf_counter()
            reps[pol] = solve(pol)
            out[pol].append(time.perf_counter() - t1)
            rel[pol] = oracle.rel_residual(rp, ci, va, b, dx.get()[:n])
            log(f"{pol}: {reps[pol]['iterations']} iterations, {1e3 * out[pol][-1]:.1f} ms")
    res = {}
    for pol in policies:
        r = reps[pol]
        res[pol] = {"ms": round(1e3 * sum(out[pol]) / args.steps, 3),
                    "kernel_ms": round(r["time_ms"], 3), "iterations": r["iterations"],
                    "refreshes": r["refreshes"], "fell_back": r["fell_back"], "status": r["status"],
                    "rel_residual_fp64_cpu": rel[pol], "prec_hist": r["prec_hist"]}
        log(f"{pol}: {res[pol]}")
    return res


def other_workload(args):
    import millefeuille as mf
    from millefeuille import gen, policy as polmod
    t0 = time.time()
    if args.workload == "convdiff100":
        rp, ci, va = gen.convdiff3d(100)
        n = len(rp) - 1
        xstar = gen.uniform_vector(n, 3)
        b = gen.csr_matvec(rp, ci, va, xstar)
        tm = mf.TiledMatrix(rp, ci, va)
        polmod.install(tm, polmod.load_or_train())
        log(f"setup: n={n} nnz={rp[-1]} tiles={tm.tilenum} ({time.time() - t0:.1f}s)")
        db, dx = tm.vector(b), tm.vector()

        def solve(pol):
            dx.zero()
            return tm.bicgstab(db, dx, allow_not_converged=True, tol=1e-10, policy=pol,
                               check_interval=args.check_interval,
                               refresh_ratio=args.refresh_ratio, refresh_every=args.refresh_every)
        res = _timed_solves(tm, solve, ["learned", "errbound", "fp64"], args, n, rp, ci, va, b, dx)
        cpu_baseline = None
        if not args.skip_cpu_baseline:
            import oracle
            tc = time.perf_counter()
            _, it_c, rel_c = oracle.bicgstab(rp, ci, va, b, tol=1e-10, maxiter=20000)
            cpu_s = time.perf_counter() - tc
            cpu, ncpu, env = _cpu_info()
            log(f"cpu ({cpu}): BiCGSTAB {it_c} iterations in {cpu_s:.2f} s, rel {rel_c:.2e}")
            cpu_baseline = {"value": round(cpu_s * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
                            "cpu_model": cpu, "usable_cores": ncpu, "omp_env": env,
                            "sample": f"oracle FP64 BiCGSTAB (textbook, CSR SpMV, 1 thread) on this "
                                      f"workload to 1e-10: {it_c} iterations in {cpu_s:.2f} s, "
                                      f"rel {rel_c:.2e}"}
        cfg = {"workload": "BiCGSTAB, 3D 7-point upwind convection-diffusion 100^3, cell Peclet 0.5, "
                           "velocity (1,0.5,0.25), b = A x*, x* ~ U[-1,1] seed 3, x0 = 0, tol 1e-10",
               "n": n, "nnz": int(rp[-1]), "tiles": tm.tilenum,
               "mixed_refresh": f"FP64 residual after a {1 / args.refresh_ratio:g}x drop or every "
                                f"{args.refresh_every} iterations"}
        _line("time-to-1e-10 residual (FP64-checked): mixed vs FP64 single-kernel BiCGSTAB",
              res["learned"]["ms"], "ms", args, res["learned"]["ms"], cfg, solves=res,
              iteration_roofline=_iteration_roofline(tm, res, "bicgstab"), cpu_baseline=cpu_baseline,
              speedup_vs_fp64=round(res["fp64"]["ms"] / res["learned"]["ms"], 3))
    elif args.workload in ("powerlaw4m", "powerlaw1m"):
        big = args.workload == "powerlaw4m"
        rp, ci, va = gen.powerlaw_laplacian(4_000_000 if big else 1_000_000,
                                            60_000_000 if big else 15_000_000, seed=5)
        n = len(rp) - 1
        log(f"generated n={n} nnz={rp[-1]} ({time.time() - t0:.1f}s)")
        b = gen.uniform_vector(n, 2)
        tm = mf.TiledMatrix(rp, ci, va)
        polmod.install(tm, polmod.load_or_train())
        log(f"setup: tiles={tm.tilenum} ({time.time() - t0:.1f}s)")
        db, dx = tm.vector(b), tm.vector()

        def solve(pol):
            dx.zero()
            return tm.cg(db, dx, allow_not_converged=True, tol=1e-10, policy=pol,
                         check_interval=args.check_interval)
        res = _timed_solves(tm, solve, ["learned", "rule", "fp64"], args, n, rp, ci, va, b, dx)
        cfg = {"workload": "CG, power-law graph Laplacian (Chung-Lu, exponent 2.5, seed 5) + 1e-3 I, "
                           "b ~ U[-1,1] seed 2, x0 = 0, tol 1e-10",
               "n": n, "nnz": int(rp[-1]), "tiles": tm.tilenum}
        _line("time-to-1e-10 residual (FP64-checked): learned vs rule policy, single-kernel CG",
              res["learned"]["ms"], "ms", args, res["learned"]["ms"], cfg, solves=res,
              iteration_roofline=_iteration_roofline(tm, res, "cg"),
              speedup_vs_rule=round(res["rule"]["ms"] / res["learned"]["ms"], 3),
              speedup_vs_fp64=round(res["fp64"]["ms"] / res["learned"]["ms"], 3))
    else:  # knn512
        nsys, nrhs = 512, 8
        sizes = np.random.default_rng(4).integers(1000, 65537, nsys)
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(8) as ex:  # the generator is numpy/scipy work
            mats = list(ex.map(lambda k: gen.knn_laplacian(int(sizes[k]), 8, seed=4 + k, shift=1e-2),
                               range(nsys)))
        log(f"generated {nsys} systems ({time.time() - t0:.1f}s)")
        tms, bs = [], []
        for k, nk in enumerate(sizes):
            tms.append(mf.TiledMatrix(*mats[k]))
            bs.append(np.stack([gen.uniform_vector(int(nk), 1000 * k + j) for j in range(nrhs)]))
        del mats
        log(f"built {nsys} handles ({time.time() - t0:.1f}s)")
        _, rep = mf.batch_cg(tms, bs, nrhs=nrhs, tol=1e-10, policy="fp64", tune_iters=20)
        waves = rep["batch_waves"]
        log(f"batch autotuner: waves per item {waves}, {rep['tune']}")
        res = {}
        for pol in ("errbound", "fp64"):
            for _ in range(args.warmup):
                mf.batch_cg(tms, bs, nrhs=nrhs, tol=1e-10, policy=pol, batch_waves=waves,
                            refresh_ratio=args.refresh_ratio, refresh_every=args.refresh_every)
            tot = []
            for _ in range(args.steps):
                _, rep = mf.batch_cg(tms, bs, nrhs=nrhs, tol=1e-10, policy=pol, batch_waves=waves,
                                     refresh_ratio=args.refresh_ratio,
                                     refresh_every=args.refresh_every)
                tot.append(rep["total_ms"])
            per = np.asarray(rep["per_system_ms"])
            res[pol] = {"total_ms": round(float(np.mean(tot)), 3),
                        "per_system_ms_p50": round(float(np.percentile(per, 50)), 3),
                        "per_system_ms_p99": round(float(np.percentile(per, 99)), 3),
                        "iterations_mean": round(float(np.mean(rep["iterations"])), 1),
                        "iterations_max": int(np.max(rep["iterations"])),
                        "max_rel_residual": float(np.max(rep["rel_residual"]))}
            log(f"{pol}: {res[pol]}")
        cfg = {"workload": "batched CG: 512 kNN-graph Laplacians (k=8, shift 1e-2, n ~ U[1k,64k] seed 4) "
                           "x 8 right-hand sides U[-1,1], one hipGraph-captured persistent launch, tol 1e-10",
               "systems": nsys, "rhs": nrhs, "sum_n": int(sizes.sum()), "waves_per_item": waves}
        _line("batched solve time to 1e-10 (every system and right-hand side)",
              res["errbound"]["total_ms"], "ms", args, res["errbound"]["total_ms"], cfg, solves=res,
              speedup_vs_fp64=round(res["fp64"]["total_ms"] / res["errbound"]["total_ms"], 3))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--policy", default="learned")
    ap.add_argument("--refresh-ratio", type=float, default=0.01,
                    help="mixed CG on the Laplacian and mixed BiCGSTAB: FP64 residual "
                         "recomputation when the recursive residual fell by this factor "
                         "(library default 0.1)")
    ap.add_argument("--refresh-every", type=int, default=128,
                    help="mixed CG on the Laplacian and mixed BiCGSTAB: ... or after this "
                         "many iterations (library default 64)")
    ap.add_argument("--check-interval", type=int, default=32,
                    help="iterations between policy evaluations (learned inference)")
    ap.add_argument("--cpu-iters", type=int, default=0,
                    help="CPU baseline CG iterations (0: run it to the tolerance)")
    ap.add_argument("--spmv-runs", type=int, default=50)
    ap.add_argument("--skip-cpu-baseline", action="store_true",
                    help="profiling runs: leave out the CPU baseline (cpu_baseline is then null)")
    ap.add_argument("--wave-nnz", type=int, default=1024,
                    help="stand-alone SpMV entries per wave (0: pick with the autotuner)")
    ap.add_argument("--workload", default="lap3d128",
                    choices=["lap3d128", "convdiff100", "knn512", "powerlaw4m", "powerlaw1m"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; tile-row sharding over "
                                   "several GPUs is not built; run with --gpus 1"}))
        sys.exit(2)
    if args.workload != "lap3d128":
        other_workload(args)
        return

    import millefeuille as mf
    from millefeuille import gen

    t0 = time.time()
    rp, ci, va = gen.laplace3d_scaled(args.grid, seed=2)
    n = len(rp) - 1
    b = gen.uniform_vector(n, 2)
    tm = mf.TiledMatrix(rp, ci, va)
    log(f"setup: n={n} nnz={rp[-1]} tiles={tm.tilenum} ({time.time() - t0:.1f}s)")
    db, dx = tm.vector(b), tm.vector()

    policy = args.policy
    policy_note = policy
    if policy == "learned":
        from millefeuille import policy as pol
        pol.install(tm, pol.load_or_train())
        policy_note = (f"learned (MLP 8-32-32-4, policy_mlp.f32, in-kernel inference every "
                       f"{args.check_interval} iterations; FP64 residual recomputation when the "
                       f"recursive residual fell {1 / args.refresh_ratio:g}x or every "
                       f"{args.refresh_every} iterations)")

    def solve(pol):
        dx.zero()
        return tm.cg(db, dx, tol=1e-10, policy=pol, check_interval=args.check_interval,
                     refresh_ratio=args.refresh_ratio, refresh_every=args.refresh_every)

    # --- warm-up, then mixed and FP64 solves alternated step by step, so
    # both see the same device state; each solve timed on its own --------
    import oracle
    for _ in range(args.warmup):
        solve(policy)
        solve("fp64")
    # (each solve returns after its kernel and the report copies have
    # completed, so the device is idle here)
    t_mixed, t_fp64 = [], []
    # as timeit does: collect once, then no collector pauses inside the timed steps
    gc.collect()
    gc.disable()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        rep = solve(policy)
        t_mixed.append(time.perf_counter() - t1)
        x = dx.get()[:n]
        lv_mixed = tm.get_precision()  # the levels the mixed solve ended with
        t2 = time.perf_counter()
        rep64 = solve("fp64")
        t_fp64.append(time.perf_counter() - t2)
        log(f"step: mixed {1e3 * t_mixed[-1]:.2f} ms (kernel {rep['time_ms']:.2f}), "
            f"fp64 {1e3 * t_fp64[-1]:.2f} ms (kernel {rep64['time_ms']:.2f})")
    gc.enable()
    x64 = dx.get()[:n]
    mixed_ms = 1e3 * sum(t_mixed) / args.steps
    fp64_ms = 1e3 * sum(t_fp64) / args.steps
    rel_cpu = oracle.rel_residual(rp, ci, va, b, x)
    rel64 = oracle.rel_residual(rp, ci, va, b, x64)
    log(f"mixed[{policy}]: {mixed_ms:.2f} ms/solve (kernel {rep['time_ms']:.2f} ms), it={rep['iterations']} "
        f"refresh={rep['refreshes']} fallback={rep['fell_back']} policy_used={rep['policy_used']} "
        f"rel(cpu)={rel_cpu:.3e} hist={rep['prec_hist']}")
    log(f"fp64: {fp64_ms:.2f} ms/solve (kernel {rep64['time_ms']:.2f} ms), it={rep64['iterations']} "
        f"rel(cpu)={rel64:.3e}")
    # the other policies on the same system, timed the same way (not `value`)
    others = {}
    for pol in ("rule", "errbound"):
        solve(pol)
        t_o = []
        for _ in range(args.steps):
            t1 = time.perf_counter()
            r_o = solve(pol)
            t_o.append(time.perf_counter() - t1)
        ms_o = 1e3 * sum(t_o) / args.steps
        others[pol] = {"ms": round(ms_o, 3), "kernel_ms": round(r_o["time_ms"], 3),
                       "iterations": r_o["iterations"], "refreshes": r_o["refreshes"],
                       "fell_back": r_o["fell_back"], "policy_used": r_o["policy_used"],
                       "rel_residual_fp64_cpu": oracle.rel_residual(rp, ci, va, b, dx.get()[:n]),
                       "prec_hist": r_o["prec_hist"]}
        log(f"{pol}: {others[pol]}")

    # --- stand-alone SpMV roofline (all-FP64 payload) ----------------------
    xs = np.random.default_rng(0).uniform(-1, 1, n)
    dxs, dys = tm.vector(xs), tm.vector()
    tune = None
    if args.wave_nnz > 0:
        tm.set_schedule(args.wave_nnz)
        wave_nnz = args.wave_nnz
    else:
        tune = tm.tune_spmv(0)
        wave_nnz = tune["best"]
        log(f"spmv autotuner: {tune}")
    timer = mf.Timer()
    for _ in range(10):
        tm.spmv(dxs, dys)
    # average launch duration: spmv_runs back-to-back launches between two
    # HIP events on the launch stream (the null stream), median of 5 such runs
    runs = []
    for _ in range(5):
        timer.start()
        for _ in range(args.spmv_runs):
            tm.spmv(dxs, dys)
        runs.append(timer.stop() / args.spmv_runs)
    t_spmv = float(np.median(runs))
    # single launches bracketed by events (includes the launch gap)
    times = []
    for _ in range(20):
        timer.start()
        tm.spmv(dxs, dys)
        times.append(timer.stop())
    t_single = float(np.median(times))
    bytes_spmv = mf.spmv_bytes(tm)
    gbs = bytes_spmv / (t_spmv * 1e-3) / 1e9
    log(f"spmv fp64: {t_spmv * 1e3:.1f} us per launch ({t_single * 1e3:.1f} us single launch), "
        f"{bytes_spmv / 1e6:.1f} MB -> {gbs:.0f} GB/s ({100 * gbs / HBM_PEAK_GBS:.1f}% of 8 TB/s)")
    # --- copy calibration: a plain device copy of known size (its counter
    # readings scale FETCH_SIZE / WRITE_SIZE of the SpMV into HBM bytes) -----
    from millefeuille._lib import lib
    ncopy = 1 << 25  # 32 Mi doubles: 256 MiB read + 256 MiB written
    csrc, cdst = mf.DeviceArray(ncopy, np.float64), mf.DeviceArray(ncopy, np.float64)
    csrc.zero()
    for _ in range(3):
        lib.mf_stream_copy(csrc.ptr, cdst.ptr, ncopy, None)
    ctimes = []
    for _ in range(10):
        timer.start()
        lib.mf_stream_copy(csrc.ptr, cdst.ptr, ncopy, None)
        ctimes.append(timer.stop())
    copy_gbs = 2 * 8 * ncopy / (float(np.median(ctimes)) * 1e-3) / 1e9
    del csrc, cdst
    log(f"copy: {copy_gbs:.0f} GB/s (256 MiB read + 256 MiB written)")
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        if tr.get("grid") == args.grid and tr.get("kernel") == tm.spmv_kernel_name():
            traffic = tr.get("hbm_bytes_per_launch")

    # --- CPU baseline: the oracle's FP64 CG (1 thread) to the same tolerance,
    # or a bounded number of iterations scaled to the GPU's count ----------
    cpu_baseline = None
    if not args.skip_cpu_baseline:
        cpu_baseline = _cpu_baseline(args, rp, ci, va, b, xs, n, rep64)

    # --- the persistent solver against the launch-per-operation one --------
    launches = _single_vs_multikernel(tm, db, dx, rp, ci, va, b,
                                      f"3D 7-point Laplacian {args.grid}^3")
    small = _small_2d_case()


    out = {
        "metric": "time-to-1e-10 residual (FP64-checked): learned/mixed vs FP64 CG; SpMV HBM GB/s",
        "value": round(mixed_ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed_ms, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64 (scalars, x, residual, accumulation); f32 products of fp32/fp16/int8 tile payloads",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"3D 7-point Laplacian {args.grid}^3, diagonal 6*10^(3U) seed 2, "
                               "b~U[-1,1] seed 2, x0=0, tol 1e-10",
                   "n": n, "nnz": int(rp[-1]), "tiles": tm.tilenum, "policy": policy_note},
        "mixed": {"ms": round(mixed_ms, 3), "kernel_ms": round(rep["time_ms"], 3),
                  "iterations": rep["iterations"], "refreshes": rep["refreshes"],
                  "fell_back": rep["fell_back"], "policy_used": rep["policy_used"],
                  "rel_residual_fp64_cpu": rel_cpu, "prec_hist": rep["prec_hist"]},
        "fp64": {"ms": round(fp64_ms, 3), "kernel_ms": round(rep64["time_ms"], 3),
                 "iterations": rep64["iterations"], "rel_residual_fp64_cpu": rel64},
        "speedup_vs_fp64": round(fp64_ms / mixed_ms, 3),
        "rule": others["rule"],
        "errbound": others["errbound"],
        "speedup_vs_rule": round(others["rule"]["ms"] / mixed_ms, 3),
        "spmv": {"us_per_launch": round(t_spmv * 1e3, 2), "us_single_launch": round(t_single * 1e3, 2),
                 "bytes": bytes_spmv, "GBps": round(gbs, 1), "wave_nnz": wave_nnz, "tune": tune,
                 "kernel": tm.spmv_kernel_name(), "copy_GBps_measured": round(copy_gbs, 1)},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu_baseline,
        "single_vs_multikernel_fp64_cg": launches,
        "laplace2d_64": small,
        "iteration_roofline": _iteration_roofline(
            tm, {policy: {"kernel_ms": rep["time_ms"], "iterations": rep["iterations"],
                          "fell_back": rep["fell_back"]},
                 "fp64": {"kernel_ms": rep64["time_ms"], "iterations": rep64["iterations"],
                          "fell_back": 0}}, "cg", levels=lv_mixed),
    }
    if not (rel_cpu <= 1e-10 and rel64 <= 1e-10):
        out["error"] = "FP64-checked residual above 1e-10"
    print(json.dumps(out), flush=True)
    _release(timer, dxs, dys, db, dx, tm)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_info():
    """CPU model, the host cores this process may use, and the OpenMP
    placement variables as set."""
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PLACES", "OMP_PROC_BIND")}
    return _cpu_model(), len(os.sched_getaffinity(0)), env


def _small_2d_case(runs=5):
    """The reference's CPU-runnable case: 2D 5-point Laplacian on a 64 x 64
    grid (n = 4096), b = A * ones, x0 = 0, tol 1e-10. FP64 CG on the GPU
    (persistent kernel, wall ms per solve, mean of `runs` after one warm-up)
    and the oracle's FP64 CG on one host thread, both FP64-checked."""
    import millefeuille as mf
    import oracle
    from millefeuille import gen
    rp, ci, va = gen.laplace2d(64)
    n = len(rp) - 1
    b = gen.csr_matvec(rp, ci, va, np.ones(n))
    t = mf.TiledMatrix(rp, ci, va)
    db, dx = t.vector(b), t.vector()
    t.cg(db, dx, tol=1e-10, policy="fp64")
    wall = []
    for _ in range(runs):
        dx.zero()
        t1 = time.perf_counter()
        r = t.cg(db, dx, tol=1e-10, policy="fp64")
        wall.append(time.perf_counter() - t1)
    rel_g = oracle.rel_residual(rp, ci, va, b, dx.get()[:n])
    _release(db, dx, t)
    tc = time.perf_counter()
    _, it_c, rel_c = oracle.cg(rp, ci, va, b, tol=1e-10, maxiter=20000)
    cpu_ms = (time.perf_counter() - tc) * 1e3
    out = {"n": n, "nnz": int(rp[-1]),
           "gpu_fp64_cg": {"ms": round(1e3 * sum(wall) / runs, 3), "kernel_ms": round(r["time_ms"], 3),
                           "iterations": r["iterations"], "rel_residual_fp64_cpu": rel_g},
           "cpu_oracle_fp64_cg_1thread": {"ms": round(cpu_ms, 3), "iterations": it_c,
                                          "rel_residual": rel_c}}
    log(f"2D 5-point Laplacian 64x64: {out}")
    return out


def _cpu_baseline(args, rp, ci, va, b, xs, n, rep64):
    """The oracle's FP64 CG (textbook, CSR SpMV, 1 thread) to the tolerance on
    this workload (or a bounded number of iterations scaled to the GPU's
    count), and the oracle's tiled SpMV (the blockspmv_cpu tile loop with
    Tile_csr_Col) on 1 thread and on every host core this process may use."""
    import oracle
    tc = time.perf_counter()
    if args.cpu_iters > 0:
        _, it_cpu, rel_c = oracle.cg(rp, ci, va, b, tol=1e-30, maxiter=args.cpu_iters)
    else:
        _, it_cpu, rel_c = oracle.cg(rp, ci, va, b, tol=1e-10, maxiter=20000)
    cpu_s = time.perf_counter() - tc
    per_it_ms = cpu_s * 1e3 / max(it_cpu, 1)
    cpu_ms = cpu_s * 1e3 if args.cpu_iters <= 0 else per_it_ms * rep64["iterations"]
    ot = oracle.tile_create(n, n, rp, ci, va)
    ncpu = len(os.sched_getaffinity(0))
    spmv = {}
    for nt in (1, ncpu):
        oracle.blockspmv_tiled(ot, xs, nthreads=nt)  # first touch
        reps = []
        for _ in range(5):
            tcs = time.perf_counter()
            oracle.blockspmv_tiled(ot, xs, nthreads=nt)
            reps.append((time.perf_counter() - tcs) * 1e3)
        spmv[nt] = float(np.median(reps))
    cpu, _, env = _cpu_info()
    log(f"cpu ({cpu}, {ncpu} usable cores, {env}): CG {it_cpu} iterations in {cpu_s:.2f} s "
        f"({per_it_ms:.1f} ms/iteration, 1 thread), rel {rel_c:.2e}; tiled spmv "
        f"{spmv[1]:.1f} ms (1 thread), {spmv[ncpu]:.1f} ms ({ncpu} threads)")
    return {"value": round(cpu_ms, 1), "unit": "ms", "cores": 1, "kind": "port",
            "cpu_model": cpu, "usable_cores": ncpu, "omp_env": env,
            "sample": (f"oracle FP64 CG (textbook, CSR SpMV, 1 thread) on this workload: "
                       f"{it_cpu} iterations in {cpu_s:.2f} s, rel {rel_c:.2e}"
                       + ("" if args.cpu_iters <= 0 else
                          f", scaled to the GPU FP64 iteration count {rep64['iterations']}")
                       + f"; oracle tiled SpMV (blockspmv_cpu loop with Tile_csr_Col), median of 5: "
                         f"{spmv[1]:.1f} ms on 1 thread, {spmv[ncpu]:.1f} ms on {ncpu} threads"),
            "spmv_ms_1thread": round(spmv[1], 2), "spmv_ms_all_threads": round(spmv[ncpu], 2)}


def _single_vs_multikernel(tm, db, dx, rp, ci, va, b, label, small_grid=32, runs=3):
    """FP64 CG to 1e-10 by the persistent single-kernel solver and by the
    launch-per-operation solver (the reference's multi-kernel comparator,
    cg_solve_reduce, src/main-cg.cu:245-279: one launch per SpMV, dot product
    and vector update, the scalars read back on the host every iteration), on
    the benchmark matrix and on a small grid where the launch gaps dominate.
    Wall ms per solve (mean of `runs` after one warm-up), kernel-region ms
    (HIP events), iterations and the oracle's FP64 residual of each."""
    import millefeuille as mf
    import oracle
    from millefeuille import gen
    rp2, ci2, va2 = gen.laplace3d_scaled(small_grid, seed=2)
    b2 = gen.uniform_vector(len(rp2) - 1, 2)
    tm2 = mf.TiledMatrix(rp2, ci2, va2)
    db2, dx2 = tm2.vector(b2), tm2.vector()
    out = {}
    for name, t, d_b, d_x, (a_rp, a_ci, a_va, a_b) in (
            (label, tm, db, dx, (rp, ci, va, b)),
            (f"3D 7-point Laplacian {small_grid}^3", tm2, db2, dx2, (rp2, ci2, va2, b2))):
        n = len(a_rp) - 1
        row = {"n": n}
        for kind, fn in (("single_kernel", t.cg), ("multi_kernel", t.cg_multikernel)):
            d_x.zero()
            fn(d_b, d_x, tol=1e-10, policy="fp64")
            wall = []
            for _ in range(runs):
                d_x.zero()
                t1 = time.perf_counter()
                r = fn(d_b, d_x, tol=1e-10, policy="fp64")
                wall.append(time.perf_counter() - t1)
            rel = oracle.rel_residual(a_rp, a_ci, a_va, a_b, d_x.get()[:n])
            ms = 1e3 * sum(wall) / runs
            row[kind] = {"ms": round(ms, 3), "kernel_ms": round(r["time_ms"], 3),
                         "iterations": r["iterations"],
                         "us_per_iteration": round(1e3 * ms / max(r["iterations"], 1), 2),
                         "rel_residual_fp64_cpu": rel}
        row["multi_over_single"] = round(row["multi_kernel"]["ms"] / row["single_kernel"]["ms"], 3)
        log(f"single vs multi-kernel FP64 CG, {name}: {row}")
        out[name] = row
    _release(db2, dx2, tm2)
    return out


def _release(*objs):
    """Frees device objects explicitly, in order, while the runtime is fully
    up (nothing is left for interpreter shutdown), then waits for the device."""
    import millefeuille as mf
    for o in objs:
        o.free()
    mf.synchronize()


if __name__ == "__main__":
    main()
