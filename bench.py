#!/usr/bin/env python3
"""Benchmark: block-Jacobi PCG solves on one MI355X.

Default workload (one "step" = one full solve): plain FP32 block-Jacobi PCG
on the 50^3 beam lattice (750,000 unknowns, 30,352,200 nnz, 6x6 blocks,
tol 1e-6, max 5000 iterations), matrix, right-hand side and x0 resident in
HBM before the timed region.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--N 50] [--no-cpu]

Other workloads (extra lines, not the headline): --workload ir | batch | serve.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s + p99 latency at 1e-8 rel. residual; SpMV HBM GB/s vs peak"
HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_bytes(n, nnz, nb):
    """Algorithmic HBM bytes per launch of the PCG kernels: the SpMV in CSR
    accounting (8 bytes per entry, row_ptr, p once, q), the update (x rw, p,
    r_in, q, r_out, z, the 6x6 factors and block starts) and the finish
    (p rw, z; its partials and state are a few KB)."""
    spmv = 8 * nnz + 4 * (n + 1) + 4 * n + 4 * n
    update = 4 * n * 7 + 144 * nb + 4 * (nb + 1)
    finish = 12 * n
    return spmv, update, finish


def kernel_bytes_one_reduction(n, nnz, nb):
    """The same for the one-reduction PCG's two kernels: the SpMV w = A u with
    r read for its dots, and the update (x, p, s_in, s_out, r_in, r_out, u
    rw, w, the factors and block starts)."""
    spmv = 8 * nnz + 4 * (n + 1) + 4 * n + 4 * n + 4 * n
    update = 4 * n * 10 + 144 * nb + 4 * (nb + 1)
    return spmv, update, 0


def traffic_from_profiles(workload_key):
    path = os.path.join(ROOT, "profiles", f"traffic_{workload_key}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def cpu_baseline(L, iters_gpu, sample_iters, sample_iters_1core):
    """The oracle (FP32 restatement of the reference loop) on the host cores,
    on a bounded sample: `sample_iters` PCG iterations of the same system on
    up to 16 threads (and `sample_iters_1core` on one), scaled to solves/s
    with the GPU's iteration count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_lib
    oracle_lib.build()
    st, lu = oracle_lib.build_lu_blocks_6(L["n"], L["row_ptr"], L["col_idx"], L["values"],
                                          L["block_starts"])
    x0 = np.zeros(L["n"], np.float32)

    def sample(threads, iters):
        oracle_lib.set_threads(threads)
        t0 = time.perf_counter()
        _, it, _, _ = oracle_lib.pcg(L["row_ptr"], L["col_idx"], L["values"], lu, L["block_starts"],
                                     L["rhs"], x0, iters, 1e-6, 0.0)
        dt = time.perf_counter() - t0
        per_iter = dt / max(it, 1)
        return 1.0 / (per_iter * (iters_gpu + 1)), it, dt, per_iter

    side = round((L["n"] / 6) ** (1.0 / 3.0))
    threads = min(16, os.cpu_count() or 1)
    v, it, dt, per_iter = sample(threads, sample_iters)
    v1, it1, dt1, per_iter1 = sample(1, sample_iters_1core)
    if it >= iters_gpu:  # the whole solve ran: no scaling
        v = 1.0 / dt
        what = f"the whole solve of the same {side}^3 system ({it} iterations) in {dt:.1f} s"
    else:
        what = (f"{it} PCG iterations of the same {side}^3 system in {dt:.1f} s "
                f"({per_iter * 1e3:.2f} ms/iteration), scaled to {iters_gpu} iterations")
    return {"value": v, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": what,
            "single_core": {"value": v1, "unit": "solves/s", "cores": 1,
                            "sample": f"{it1} iterations in {dt1:.1f} s ({per_iter1 * 1e3:.2f} ms/iteration)"}}


def bench_plain(args):
    import wgpu_solver_backend_amd as H
    N = args.N
    log(f"generating {N}^3 lattice")
    L = H.lattice(N, seed=0)
    n, nnz, nb = L["n"], L["nnz"], L["n"] // 6
    ctx = H.GpuContext.create(0)
    log(ctx.describe())
    system = H.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    one_red = args.alg == "one-reduction"
    if one_red:
        system.set_algorithm(H.ALG_PCG_ONE_REDUCTION)
    system.set_rhs(L["rhs"], np.zeros(n, np.float32))
    max_iter, tol = 5000, 1e-6
    best, tune_us = system.autotune(40)
    info = system.info()
    log(f"autotune: SpMV variant {best}; us/iteration per variant {np.round(tune_us, 2).tolist()}; "
        f"{info['local_tiles']}/{info['n_tiles']} tiles gather from LDS")
    copy_bytes = 256 << 20
    copy_us = ctx.stream_copy_us(copy_bytes, 10)
    copy_gbps = 2 * copy_bytes / (copy_us * 1e-6) / 1e9
    log(f"stream copy {copy_bytes >> 20} MiB: {copy_us:.1f} us, {copy_gbps:.0f} GB/s")
    if args.profile_iters:
        # short counter-collection run: a few full-size iterations, no JSON headline
        for _ in range(max(args.steps, 1)):
            r = system.solve_resident(args.profile_iters, tol, 0.0, raise_on_error=False)
            log(f"profile run: status {r.status}, {r.iterations} its")
        print(json.dumps({"profile_run": True, "iterations_per_solve": args.profile_iters,
                          "copy_bytes": copy_bytes, "copy_us": copy_us}), flush=True)
        return None
    for i in range(args.warmup):
        res = system.solve_resident(max_iter, tol, 0.0)
        log(f"warmup {i}: {res.iterations} its, {res.solve_ms:.1f} ms")
    system.set_kernel_timing(args.sample_every)
    lat = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        res = system.solve_resident(max_iter, tol, 0.0)  # returns after the device finished
        lat.append((time.perf_counter() - ts) * 1e3)
    elapsed = time.perf_counter() - t0
    avg_us, cnt = system.kernel_timing()
    system.set_kernel_timing(0)
    log(f"timed: {args.steps} solves in {elapsed:.3f} s; {res.iterations} its; "
        f"kernel avg us {avg_us.round(2).tolist()} ({cnt.tolist()} samples)")
    b1, b2, b3 = (kernel_bytes_one_reduction if one_red else kernel_bytes)(n, nnz, nb)
    gbps = lambda b, us: b / (us * 1e-6) / 1e9 if us > 0 else None
    spmv_gbps = b1 / (avg_us[0] * 1e-6) / 1e9
    layout_bytes = info["spmv_sliced_bytes"] if best in (4, 5) and info["spmv_sliced_bytes"] else \
        info["spmv_layout_bytes"]
    layout_gbps = layout_bytes / (avg_us[0] * 1e-6) / 1e9
    iter_us = float(avg_us.sum())
    iter_bytes = b1 + b2 + b3
    key = f"plain_N{N}"
    traffic = traffic_from_profiles(key)
    out = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"plain FP32 block-Jacobi PCG, {N}^3 beam lattice ({n:,} unknowns, "
                                f"{nnz:,} nnz, 6x6 blocks), tol {tol:g}, max {max_iter}, x0 = 0"
                                + (", one-reduction (Chronopoulos-Gear) PCG: two launches per "
                                   "iteration, not bit-identical to the reference" if one_red else "")),
                   "n": n, "nnz": nnz, "iterations": int(res.iterations), "seed": 0,
                   "algorithm": "one-reduction" if one_red else "pcg"},
        "latency_ms": {"p50": statistics.median(lat), "p99": float(np.percentile(lat, 99)),
                       "max": max(lat)},
        "true_rel_residual_fp64": res.rel_res_true_fp64,
        "spmv_gbps": spmv_gbps,
        "stream_copy_gbps": copy_gbps,
        "iteration": {"us": iter_us, "wall_us": elapsed / args.steps / int(res.iterations + 1) * 1e6,
                      "bytes": iter_bytes, "gbps": iter_bytes / (iter_us * 1e-6) / 1e9,
                      "kernel_us": {"spmv": avg_us[0], "update": avg_us[1],
                                    "finish": avg_us[2]},
                      "kernel_gbps": {"spmv": spmv_gbps, "update": gbps(b2, avg_us[1]),
                                      "finish": gbps(b3, avg_us[2])}},
        "spmv": {"variant": best, "autotune_us_per_iter": [float(x) for x in tune_us],
                 "local_tiles": info["local_tiles"], "n_tiles": info["n_tiles"],
                 "layout_bytes": layout_bytes, "layout_gbps": layout_gbps},
        "roofline": {"bound": "hbm", "kernel": "k_ccg_spmv" if one_red else "k_pcg_spmv",
                     "achieved": spmv_gbps,
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": spmv_gbps / HBM_PEAK_GBPS,
                     "traffic": None if one_red else
                     ((traffic or {}).get("k_pcg_spmv_bytes_by_variant") or {}).get(str(best)),
                     "algorithmic_bytes": b1,
                     "note": "achieved = CSR algorithmic bytes (8 nnz + 12 n + 4) / kernel time. "
                             "The kernel's own layout streams fewer bytes than CSR accounting "
                             "(spmv.layout_bytes: the sliced layout reads about 4.6 bytes per "
                             "entry, the tile-local one 6), so frac is not the bandwidth drawn: "
                             "that is spmv.layout_gbps, and traffic / time from the counters; "
                             "variant 5 reads the layout through the caches, and a layout that fits "
                             "the 256 MB last-level cache is partly served from there"},
    }
    if not args.no_cpu:
        log("cpu baseline (oracle on host cores)")
        out["cpu_baseline"] = cpu_baseline(L, int(res.iterations), args.cpu_iters, args.cpu_iters_1core)
    else:
        out["cpu_baseline"] = None
    system.close()
    ctx.close()
    return out


def _lattice_sides(rng, count, lo, hi):
    return [int(round(np.exp(rng.uniform(np.log(lo), np.log(hi))))) for _ in range(count)]


def _host_rel_res(L, x):
    import scipy.sparse as sp
    n = L["n"]
    A = sp.csr_matrix((L["values"].astype(np.float64), L["col_idx"], L["row_ptr"]), shape=(n, n))
    b = L["rhs"].astype(np.float64)
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))


def bench_ir(args):
    """Mixed precision at 50^3 with E varied over 3 decades: FP32 PCG inside FP64
    refinement to a 1e-10 relative residual (verified on the host in FP64)."""
    import wgpu_solver_backend_amd as H
    N = args.N
    L = H.lattice(N, seed=0, e_log10_half_range=1.5)
    n, nnz = L["n"], L["nnz"]
    ctx = H.GpuContext.create(0)
    system = H.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    # the SpMV variant, timed on this system's FP32 iterations
    system.set_rhs(L["rhs"], np.zeros(n, np.float32))
    best, tune_us = system.autotune(40)
    log(f"autotune: SpMV variant {best}; us/iteration per variant {np.round(tune_us, 2).tolist()}")
    b64 = L["rhs"].astype(np.float64)
    system.set_rhs64(b64)
    params = H.IrParams(max_outer=40, max_inner=20000, inner_rel_tol=args.inner_tol, target_rel_res=1e-10)
    for _ in range(args.warmup):
        x, res = system.solve_ir(b64, None, params)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        x, res = system.solve_ir(b64, None, params)
        lat.append((time.perf_counter() - ts) * 1e3)
    elapsed = time.perf_counter() - t0
    host_res = _host_rel_res(L, x)
    # plain FP32 PCG on the same system for comparison (its FP64 true residual)
    system.set_rhs(L["rhs"])
    r32 = system.solve_resident(20000, 1e-6, 0.0, raise_on_error=False)
    return {"metric": METRIC, "value": args.steps / elapsed, "unit": "solves/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
            "data": "synthetic",
            "config": {"workload": f"mixed precision: {N}^3 beam lattice, E over 3 decades, FP32 PCG "
                                   f"(inner rel tol {args.inner_tol:g}) in FP64 refinement to 1e-10",
                       "n": n, "nnz": nnz, "outer": int(res.outer_iterations),
                       "inner_iterations": int(res.iterations)},
            "latency_ms": {"p50": statistics.median(lat), "p99": float(np.percentile(lat, 99))},
            "spmv": {"variant": best, "autotune_us_per_iter": [float(x) for x in tune_us]},
            "true_rel_residual_fp64_device": res.rel_res_true_fp64,
            "true_rel_residual_fp64_host": host_res,
            "plain_fp32_true_rel_residual_fp64": r32.rel_res_true_fp64,
            "plain_fp32_iterations": int(r32.iterations)}


def bench_batch(args):
    """256 independent lattices (sides log-uniform in [4,24]) solved together,
    FP32 inner + FP64 refinement to 1e-8, against the same systems solved one
    at a time through the per-system path."""
    import wgpu_solver_backend_amd as H
    rng = np.random.default_rng(0)
    sides = _lattice_sides(rng, args.batch, 4, 24)
    ctx = H.GpuContext.create(0)
    Ls = [H.lattice(N, seed=1000 + i) for i, N in enumerate(sides)]
    systems = [H.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"]) for L in Ls]
    for L, s in zip(Ls, systems):
        s.set_rhs64(L["rhs"].astype(np.float64))
    params = H.IrParams(max_outer=30, max_inner=20000, inner_rel_tol=args.inner_tol, target_rel_res=1e-8)
    batch = H.Batch(ctx, systems)
    best, tune_us = batch.autotune(20, ir=True)
    log(f"batch autotune: SpMV variant {best}; us/iteration per variant {np.round(tune_us, 1).tolist()}")
    for _ in range(args.warmup):
        batch.solve_ir(params)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st, results = batch.solve_ir(params)
    elapsed = time.perf_counter() - t0
    batched_sps = args.steps * len(systems) / elapsed
    # per-system path on the same systems (one solve each)
    t1 = time.perf_counter()
    alone = []
    for L, s in zip(Ls, systems):
        x, r = s.solve_ir(L["rhs"].astype(np.float64), None, params)
        alone.append(r)
    t_alone = time.perf_counter() - t1
    alone_sps = len(systems) / t_alone
    worst = max(r.rel_res_true_fp64 for r in results)
    same = all(a.iterations == r.iterations for a, r in zip(alone, results))
    ns = [L["n"] for L in Ls]
    return {"metric": METRIC, "value": batched_sps, "unit": "solves/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
            "data": "synthetic",
            "config": {"workload": f"batched: {len(systems)} beam lattices, sides log-uniform [4,24] "
                                   f"({min(ns):,}..{max(ns):,} unknowns, sum {sum(ns):,}), FP32 inner "
                                   f"(rel tol {args.inner_tol:g}) + FP64 refinement to 1e-8, one tile "
                                   f"table, graph-replayed chunks",
                       "systems": len(systems)},
            "spmv": {"variant": best, "autotune_us_per_iter": [float(u) for u in tune_us]},
            "per_system_path_solves_per_s": alone_sps, "speedup_vs_per_system": batched_sps / alone_sps,
            "worst_true_rel_residual_fp64": worst, "iterations_identical_alone_vs_batched": same,
            "status_all_ok": st == 0}


def bench_serve(args):
    """Poisson request stream (seeded, target 200 req/s) of lattices with sides
    log-uniform in [6,40], IR to 1e-8, continuous batching under an HBM budget."""
    import wgpu_solver_backend_amd as H
    import threading
    rng = np.random.default_rng(0)
    count = args.requests
    sides = _lattice_sides(rng, count, 6, 40)
    gaps = rng.exponential(1.0 / args.rate, count)
    log(f"serve: generating {count} requests")
    reqs = [H.lattice(N, seed=2000 + i) for i, N in enumerate(sides)]
    ctx = H.GpuContext.create(0)
    params = H.IrParams(max_outer=30, max_inner=20000, inner_rel_tol=args.inner_tol, target_rel_res=1e-8)
    srv = H.Server(ctx, int(args.budget_gb * (1 << 30)), max_active=args.max_active, chunk_iters=8)
    tickets = [None] * count
    lat = [None] * count
    status = [None] * count
    iters = [(0, 0)] * count

    def waiter():
        for i in range(count):
            while tickets[i] is None:
                time.sleep(0.0005)
            st, x, res, ms = srv.wait(tickets[i])
            lat[i], status[i] = ms, st
            iters[i] = (res.iterations, res.outer_iterations)

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    t_next = t0
    for i, L in enumerate(reqs):
        t_next += gaps[i]
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        tickets[i] = srv.submit(L["row_ptr"], L["col_idx"], L["values"], L["block_starts"],
                                L["rhs"].astype(np.float64), params)
    th.join()
    elapsed = time.perf_counter() - t0
    stats = srv.stats()
    srv.close()
    ok = sum(1 for s in status if s == 0)
    # bytes the requests' iterations stream in CSR accounting: per inner PCG
    # iteration 8 nnz + 76.7 n (SpMV, dots, updates, block solves fused), per
    # FP64 residual step 8 nnz + 28 n + 4
    moved = sum(it * (8 * L["nnz"] + 76.7 * L["n"]) + (outer + 1) * (8 * L["nnz"] + 28 * L["n"] + 4)
                for (it, outer), L in zip(iters, reqs))
    return {"metric": METRIC, "value": ok / elapsed, "unit": "solves/s", "n_gpus": 1,
            "steps": count, "warmup": 0, "ms_per_step": elapsed / count * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
            "data": "synthetic",
            "config": {"workload": f"serving: Poisson arrivals at {args.rate:g} req/s (seed 0), {count} "
                                   f"requests, lattice sides log-uniform [6,40], FP64-verified 1e-8 (FP32 inner "
                                   f"rel tol {args.inner_tol:g}), "
                                   f"HBM budget {args.budget_gb:g} GiB, <= {args.max_active} resident"},
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "max": float(max(lat))},
            "offered_rate": count / (t_next - t0), "completed_ok": ok,
            "inner_iterations_total": int(sum(i for i, _ in iters)),
            "streamed_gbps": moved / elapsed / 1e9, "server": stats}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=50)
    ap.add_argument("--sample-every", type=int, default=32,
                    help="HIP-event timing of every k-th PCG iteration's kernels (inside the "
                         "timed solves; each sample costs a few microseconds)")
    ap.add_argument("--cpu-iters", type=int, default=5000,
                    help="most oracle PCG iterations timed on up to 16 host threads (the default "
                         "lets the 50^3 solve run to convergence: the whole solve is timed)")
    ap.add_argument("--cpu-iters-1core", type=int, default=800,
                    help="oracle PCG iterations timed on one host thread (scaled)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="plain", choices=["plain", "ir", "batch", "serve"])
    ap.add_argument("--alg", default="pcg", choices=["pcg", "one-reduction"],
                    help="plain workload: the reference's PCG (default, bit-identical to the "
                         "reference) or the one-reduction PCG (two launches per iteration)")
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="inner FP32 PCG relative tolerance of the refinement loop (default: 1e-4 "
                         "for --workload ir, 1e-3 for batch and serve; the fastest in a sweep over "
                         "1e-2..1e-5 on one MI355X)")
    ap.add_argument("--profile-iters", type=int, default=0,
                    help="counter runs: solves capped at this many iterations, no headline line")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--requests", type=int, default=400)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--max-active", type=int, default=256)
    args = ap.parse_args()
    ws = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or ws != 1:
        print(json.dumps({"error": "this build measures one GPU: independent systems shard by "
                                   "system across GPUs with no exchange, so N>1 is not run here",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.inner_tol is None:
        args.inner_tol = 1e-4 if args.workload == "ir" else 1e-3
    out = {"plain": bench_plain, "ir": bench_ir, "batch": bench_batch, "serve": bench_serve}[args.workload](args)
    if out is not None:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
