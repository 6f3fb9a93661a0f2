#!/usr/bin/env python3
"""Benchmark: block-Jacobi PCG solves on one MI355X.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Default workload (the headline): 256 independent beam-lattice systems, sides
log-uniform in [4, 24] nodes (384 .. 82,944 unknowns each), FP32 block-Jacobi
PCG inside FP64 iterative refinement to a 1e-8 relative residual, all
systems advanced together.  One step = one solve of all 256 systems from
x0 = 0, matrices and right-hand sides resident in HBM.  `value` is systems
solved per second; latency_ms holds each system's time to its FP64-verified
solution within its step.  The same line carries:
  * serving: a seeded Poisson stream of lattices with sides log-uniform in
    [6, 40], FP64 refinement to 1e-8, at the configured 200 requests/s and at
    0.7x the completed rate measured there (p50 / p99 latency, offered and
    completed rates);
  * the plain FP32 PCG on the 50^3 lattice (750,000 unknowns, 30,352,200
    nnz, tol 1e-6): solves/s, and the SpMV roofline -- the kernel's mean
    duration (HIP events, sampled in an untimed pass after the timed solves)
    against the bytes its layout streams, and a 64^3 point whose layout does
    not fit the 256 MB last-level cache (HBM alone);
  * cpu_baseline: the same batch workload on the host cores (FP32 PCG
    restatement of the reference inside FP64 refinement), on a bounded sample.

Other workloads (one line each): --workload plain | ir | batch | serve.
"""
import argparse
import json
import os
import statistics
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s + p99 latency at 1e-8 rel. residual; SpMV HBM GB/s vs peak"
HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads():
    """Host threads for the CPU baselines: the cores this process may run on,
    capped by OMP_NUM_THREADS (a machine that grants a job a share of its
    cores says so there), and the machine's core count for the record."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(affinity, cap) if cap > 0 else affinity
    return max(threads, 1), {"nproc": os.cpu_count(), "affinity": affinity,
                             "omp_num_threads": cap or None}


def kernel_bytes(n, nnz, nb):
    """Algorithmic HBM bytes per launch of the PCG kernels in CSR accounting:
    the SpMV (8 bytes per entry, row_ptr, p once, q), the update (x rw, p,
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


def spmv_stream_bytes(info, variant):
    """Bytes one launch of the chosen SpMV variant streams: the sliced layout
    (variants 4 and 5 on a system that has it), the tile-local layout
    (variants 2, 3, and 4/5 without a sliced layout) or CSR (0, 1)."""
    if variant in (4, 5) and info["spmv_sliced_bytes"]:
        return info["spmv_sliced_bytes"], "sliced"
    if variant >= 2:
        return info["spmv_layout_bytes"], "tile-local"
    return info["spmv_csr_bytes"], "csr"


def traffic_from_profiles(workload_key):
    path = os.path.join(ROOT, "profiles", f"traffic_{workload_key}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def lattice_sides(seed, count, lo, hi):
    rng = np.random.default_rng(seed)
    return [int(round(np.exp(rng.uniform(np.log(lo), np.log(hi))))) for _ in range(count)]


def host_rel_res(L, x):
    import scipy.sparse as sp
    n = L["n"]
    A = sp.csr_matrix((L["values"].astype(np.float64), L["col_idx"], L["row_ptr"]), shape=(n, n))
    b = L["rhs"].astype(np.float64)
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))


def pct(v, q):
    return float(np.percentile(np.asarray(v, np.float64), q))


# ----------------------------------------------------------------- batch

def run_batch(H, ctx, args, steps, warmup):
    """256 lattices (sides log-uniform [4,24], seeds 1000+i) solved together to
    an FP64 1e-8 relative residual; per-system time to solution; host FP64
    verification of every result of the last step; the per-system path on the
    same systems."""
    sides = lattice_sides(0, args.batch, 4, 24)
    log(f"batch: generating {len(sides)} lattices")
    Ls = [H.lattice(N, seed=1000 + i) for i, N in enumerate(sides)]
    systems = [H.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"]) for L in Ls]
    for L, s in zip(Ls, systems):
        s.set_rhs64(L["rhs"].astype(np.float64))
    params = H.IrParams(max_outer=30, max_inner=20000, inner_rel_tol=args.inner_tol, target_rel_res=1e-8)
    batch = H.Batch(ctx, systems)
    best, tune_us = batch.autotune(20, ir=True)
    log(f"batch autotune: SpMV variant {best}; us/iteration per variant {np.round(tune_us, 1).tolist()}")
    for i in range(warmup):
        st, _ = batch.solve_ir(params)
        if st != 0:
            raise RuntimeError(f"batch warmup status {st}: {H._lib.last_error()}")
    lat, step_ms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        st, results = batch.solve_ir(params)
        step_ms.append((time.perf_counter() - ts) * 1e3)
        if st != 0:
            raise RuntimeError(f"batch status {st}: {H._lib.last_error()}")
        lat.extend(r.solve_ms for r in results)
    elapsed = time.perf_counter() - t0
    value = steps * len(systems) / elapsed
    log(f"batch: {steps} steps of {len(systems)} systems in {elapsed:.3f} s = {value:.1f} solves/s")
    # per-kernel durations from an untimed solve (the timed ones replay the
    # graphs without events)
    kern = batch_kernel_rates(batch, systems, best, params)
    # FP64 verification on the host of every system's last result
    dev_res = [r.rel_res_true_fp64 for r in results]
    host_res = [host_rel_res(L, s.get_x64()) for L, s in zip(Ls, systems)]
    iters = [r.iterations for r in results]
    outer = [r.outer_iterations for r in results]
    # the per-system path: each system solved alone, one after another
    same, alone_sps = None, None
    if not args.skip_per_system:
        t1 = time.perf_counter()
        same = True
        for L, s, r in zip(Ls, systems, results):
            _, ra = s.solve_ir(L["rhs"].astype(np.float64), None, params)
            same &= ra.iterations == r.iterations
        alone_sps = len(systems) / (time.perf_counter() - t1)
    # all three kernels' algorithmic bytes of one solve over the timed step time
    # (both tables running side by side)
    step_s = elapsed / steps
    kern["effective"] = {"bytes_per_solve": kern["solve_bytes"], "step_ms": step_s * 1e3,
                         "gbps": kern["solve_bytes"] / step_s / 1e9,
                         "frac": kern["solve_bytes"] / step_s / 1e9 / HBM_PEAK_GBPS}
    log(f"batch: {kern['solve_bytes'] / 1e9:.3f} GB of kernel bytes per solve, "
        f"{kern['effective']['gbps']:.0f} GB/s over the timed steps")
    # the same systems refined to 1e-10: batched, and one by one
    r10 = run_batch_1e10(H, batch, systems, Ls, args)
    ns = [L["n"] for L in Ls]
    out = {
        "value": value, "elapsed": elapsed, "step_ms": step_ms, "latency_ms": lat,
        "best": best, "tune_us": tune_us, "Ls": Ls, "params": params, "sides": sides,
        "summary": {
            "systems": len(systems), "unknowns_min": min(ns), "unknowns_max": max(ns),
            "unknowns_sum": int(sum(ns)), "nnz_sum": int(sum(L["nnz"] for L in Ls)),
            "inner_iterations_mean": float(np.mean(iters)), "outer_iterations_max": int(max(outer)),
            "worst_true_rel_residual_fp64_device": float(max(dev_res)),
            "worst_true_rel_residual_fp64_host": float(max(host_res)),
            "all_verified_1e-8": bool(max(host_res) <= 1e-8 * 1.001),
            "per_system_path_solves_per_s": alone_sps,
            "speedup_vs_per_system": value / alone_sps if alone_sps else None,
            "iterations_identical_alone_vs_batched": same,
            "spmv_variant": best, "autotune_us_per_iter": [float(u) for u in tune_us],
            "refined_to_1e-10": r10,
        },
        "kernels": kern,
    }
    batch.close()
    for s in systems:
        s.close()
    return out


def run_batch_1e10(H, batch, systems, Ls, args, steps=2):
    """The batch refined to a 1e-10 relative residual (FP32 inner PCG,
    FP64 residuals and dots): solves/s batched (one warmup, `steps` timed),
    every result verified on the host, and the same refinement one system
    at a time."""
    p10 = H.IrParams(max_outer=40, max_inner=20000, inner_rel_tol=args.inner_tol, target_rel_res=1e-10)
    st, _ = batch.solve_ir(p10)
    if st != 0:
        raise RuntimeError(f"batch 1e-10 warmup status {st}: {H._lib.last_error()}")
    t0 = time.perf_counter()
    for _ in range(steps):
        st, res = batch.solve_ir(p10)
        if st != 0:
            raise RuntimeError(f"batch 1e-10 status {st}: {H._lib.last_error()}")
    el = time.perf_counter() - t0
    value = steps * len(systems) / el
    host = max(host_rel_res(L, s.get_x64()) for L, s in zip(Ls, systems))
    out = {"solves_per_s": value, "worst_true_rel_residual_fp64_host": host,
           "all_verified_1e-10": bool(host <= 1e-10 * 1.001),
           "inner_iterations_mean": float(np.mean([r.iterations for r in res])),
           "outer_iterations_max": int(max(r.outer_iterations for r in res)),
           "latency_ms": {"p50": pct([r.solve_ms for r in res], 50),
                          "p99": pct([r.solve_ms for r in res], 99)}}
    if not args.skip_per_system:
        t1 = time.perf_counter()
        for L, s in zip(Ls, systems):
            _, ra = s.solve_ir(L["rhs"].astype(np.float64), None, p10)
        alone = len(systems) / (time.perf_counter() - t1)
        out["per_system_path_solves_per_s"] = alone
        out["speedup_vs_per_system"] = value / alone
    log(f"batch to 1e-10: {value:.1f} solves/s, worst host residual {host:.2e}"
        + (f", per-system path {out['per_system_path_solves_per_s']:.1f}" if "per_system_path_solves_per_s" in out else ""))
    return out


def batch_kernel_rates(batch, systems, variant, params, every=4):
    """HIP-event durations of the batch's three kernels over one untimed
    solve (every `every`-th launch set) and the bytes each sampled launch
    streams: a launch set works only on the systems still running (a finished
    system's workgroups exit before any load), so its algorithmic bytes are
    those systems' -- the SpMV's layout with the rows' own p and q, the
    update's 28 n + 144 nb (x rw, p, q, r in and out, z; the 6x6 factors),
    the finish's 12 n (p rw, z).  Sets that run a system's FP64 refinement
    residual (3-4 of its ~1,100) are counted as iterations."""
    per = []
    for s in systems:
        info = s.info()
        sb, _ = spmv_stream_bytes(info, variant)
        per.append((sb, 28 * info["n"] + 144 * info["n_blocks"], 12 * info["n"]))
    per = np.array(per, np.float64)
    batch.set_kernel_timing(every)
    st, _ = batch.solve_ir(params)
    batch.set_kernel_timing(0)
    if st != 0:
        raise RuntimeError(f"batch timing solve status {st}")
    rec, sets = batch.kernel_samples()
    names = ("spmv", "update", "finish")
    out = {"samples": int(len(rec)), "every": every, "kernel": {}}
    full_sets = int(sets.min())
    for k, nm in enumerate(names):
        t = rec[:, 1 + k]
        ok = t >= 0
        if not ok.any():
            continue
        live_bytes = np.array([per[sets > int(i), k].sum() for i in rec[ok, 0]])
        tt = t[ok]
        full = rec[ok, 0] < full_sets  # every system still running
        e = {"us_mean": float(tt.mean()), "bytes_mean": float(live_bytes.mean()),
             "gbps": float(live_bytes.sum() / (tt.sum() * 1e-6) / 1e9)}
        if full.any():
            e["all_systems"] = {"samples": int(full.sum()), "us_mean": float(tt[full].mean()),
                                "bytes": float(per[:, k].sum()),
                                "gbps": float(per[:, k].sum() / (tt[full].mean() * 1e-6) / 1e9)}
        out["kernel"][nm] = e
    # every launch set of the solve: system i took part in sets [0, sets[i])
    out["solve_bytes"] = float((per.sum(axis=1) * sets.astype(np.float64)).sum())
    sp = out["kernel"]["spmv"]
    log(f"batch kernels: spmv {sp['us_mean']:.1f} us, {sp['gbps']:.0f} GB/s over the solve; "
        f"all systems running: {sp.get('all_systems', {}).get('us_mean', 0):.1f} us, "
        f"{sp.get('all_systems', {}).get('gbps', 0):.0f} GB/s")
    return out


# ----------------------------------------------------------------- serving

def serve_stream(H, ctx, reqs, rate, seed, params, budget_gb, max_active, submitters=4):
    """Poisson arrivals (exponential gaps, mean 1/rate, default_rng(seed)): at
    each scheduled arrival time the request goes to one of `submitters`
    threads, which submit in parallel (a submission copies the request, tens
    of MB for the large ones, so one thread alone falls behind the schedule);
    another thread waits for the results.  A request's latency is counted from
    its scheduled arrival: the library's submit-to-completion time plus the
    time between the arrival and the end of its submit call."""
    from concurrent.futures import ThreadPoolExecutor
    count = len(reqs)
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, count)
    b64s = [L["rhs"].astype(np.float64) for L in reqs]
    # 64 iterations between status polls: p50 / p99 at 50 req/s 83 / 276 ms, against 81 / 296 with
    # 32 and 87 / 314 with 8; 76 req/s completed at 200 req/s offered (75.5 with 8)
    srv = H.Server(ctx, int(budget_gb * (1 << 30)), max_active=max_active, chunk_iters=64)
    tickets = [None] * count
    t_arr = [0.0] * count
    t_ret = [0.0] * count
    lat = [None] * count
    status = [None] * count
    dev_res = [None] * count
    iters = [(0, 0)] * count
    t_done = [0.0]

    def submit(i):
        L = reqs[i]
        tk = srv.submit(L["row_ptr"], L["col_idx"], L["values"], L["block_starts"], b64s[i], params)
        t_ret[i] = time.perf_counter()
        tickets[i] = tk

    def waiter():
        for i in range(count):
            while tickets[i] is None:
                time.sleep(0.0005)
            st, _, res, ms = srv.wait(tickets[i])
            lat[i] = ms + (t_ret[i] - t_arr[i]) * 1e3
            status[i], dev_res[i] = st, res.rel_res_true_fp64
            iters[i] = (res.iterations, res.outer_iterations)
        t_done[0] = time.perf_counter()

    th = threading.Thread(target=waiter)
    th.start()
    pool = ThreadPoolExecutor(max_workers=submitters)
    t0 = time.perf_counter()
    t_next = t0
    for i in range(count):
        t_next += gaps[i]
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        t_arr[i] = t_next
        pool.submit(submit, i)
    pool.shutdown(wait=True)
    submit_lag = max(t_ret[i] - t_arr[i] for i in range(count))
    th.join()
    elapsed = t_done[0] - t0
    stats = srv.stats()
    srv.close()
    ok = sum(1 for s in status if s == 0)
    # bytes the requests' iterations stream in CSR accounting: per inner PCG
    # iteration 8 nnz + 76.7 n (SpMV, dots, updates, block solves fused), per
    # FP64 residual step 8 nnz + 28 n + 4
    moved = sum(it * (8 * L["nnz"] + 76.7 * L["n"]) + (outer + 1) * (8 * L["nnz"] + 28 * L["n"] + 4)
                for (it, outer), L in zip(iters, reqs))
    return {"offered_rate": count / (t_next - t0), "target_rate": rate,
            "completed_rate": ok / elapsed, "requests": count, "completed_ok": ok,
            "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": float(max(lat))},
            "max_submit_lag_ms": submit_lag * 1e3,
            "worst_true_rel_residual_fp64_device": float(max(r for r in dev_res if r is not None)),
            "inner_iterations_total": int(sum(i for i, _ in iters)),
            "streamed_gbps_csr_accounting": moved / elapsed / 1e9, "elapsed_s": elapsed,
            "server": dict({k: stats[k] for k in ("peak_active", "peak_device_bytes", "chunks",
                                                  "chunks_cached_spmv", "table_changes",
                                                  "arena_waits", "admit_ms", "table_ms",
                                                  "upload_bytes")},
                           admit_ms_per_request=stats["admit_ms"] / max(count, 1))}


def run_serving(H, ctx, args):
    sides = lattice_sides(0, args.requests, 6, 40)
    log(f"serve: generating {len(sides)} requests")
    reqs = [H.lattice(N, seed=2000 + i) for i, N in enumerate(sides)]
    params = H.IrParams(max_outer=30, max_inner=20000, inner_rel_tol=args.inner_tol, target_rel_res=1e-8)
    at_target = serve_stream(H, ctx, reqs, args.rate, 1, params, args.budget_gb, args.max_active)
    log(f"serve at {args.rate:g} req/s: completed {at_target['completed_rate']:.1f}/s, "
        f"p50 {at_target['latency_ms']['p50']:.0f} ms, p99 {at_target['latency_ms']['p99']:.0f} ms")
    capacity = min(at_target["completed_rate"], args.rate)
    sustained_rate = 0.7 * capacity
    sustained = serve_stream(H, ctx, reqs, sustained_rate, 2, params, args.budget_gb, args.max_active)
    log(f"serve at {sustained_rate:.1f} req/s: completed {sustained['completed_rate']:.1f}/s, "
        f"p50 {sustained['latency_ms']['p50']:.0f} ms, p99 {sustained['latency_ms']['p99']:.0f} ms")
    ns = [L["n"] for L in reqs]
    return {"workload": (f"serving: Poisson arrivals (seeded), {len(reqs)} requests, lattice sides "
                         f"log-uniform [6,40] ({min(ns):,}..{max(ns):,} unknowns), FP64 refinement to "
                         f"1e-8 (FP32 inner rel tol {args.inner_tol:g}), HBM budget {args.budget_gb:g} "
                         f"GiB, <= {args.max_active} resident"),
            "at_target_rate": at_target, "at_0.7x_capacity": sustained}


# ----------------------------------------------------------------- plain 50^3

def run_plain(H, ctx, N, steps, warmup, alg="pcg", sample_every=4):
    """Plain FP32 PCG on the N^3 lattice: timed solves on the production path
    (graph replay, no kernel timing), then one untimed solve with HIP-event
    kernel timing for the per-kernel durations."""
    log(f"plain: generating {N}^3 lattice")
    L = H.lattice(N, seed=0)
    n, nnz, nb = L["n"], L["nnz"], L["n"] // 6
    system = H.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    one_red = alg == "one-reduction"
    if one_red:
        system.set_algorithm(H.ALG_PCG_ONE_REDUCTION)
    system.set_rhs(L["rhs"], np.zeros(n, np.float32))
    max_iter, tol = 5000, 1e-6
    best, tune_us = system.autotune(40)
    info = system.info()
    log(f"plain {N}^3 autotune: SpMV variant {best}; us/iteration {np.round(tune_us, 2).tolist()}")
    for _ in range(warmup):
        system.solve_resident(max_iter, tol, 0.0)
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        res = system.solve_resident(max_iter, tol, 0.0)
        lat.append((time.perf_counter() - ts) * 1e3)
    elapsed = time.perf_counter() - t0
    # untimed pass: kernel durations from HIP events (every k-th iteration)
    system.set_kernel_timing(sample_every)
    system.solve_resident(max_iter, tol, 0.0)
    avg_us, cnt = system.kernel_timing()
    system.set_kernel_timing(0)
    b1, b2, b3 = (kernel_bytes_one_reduction if one_red else kernel_bytes)(n, nnz, nb)
    sb, layout = spmv_stream_bytes(info, best)
    two = info["launches_per_iteration"] == 2 and not one_red
    if two:
        # the SpMV launch also finishes the previous iteration: it reads z and
        # writes the new p for its own rows (p_prev replaces the own p it reads)
        b1, sb, b3 = b1 + 8 * n, sb + 8 * n, 0
    gbps = lambda b, us: b / (us * 1e-6) / 1e9 if us > 0 else None
    out = {"L": L, "value": steps / elapsed, "elapsed": elapsed, "lat": lat, "res": res,
           "best": best, "tune_us": tune_us, "info": info, "avg_us": avg_us, "cnt": cnt,
           "n": n, "nnz": nnz,
           "spmv": {"variant": best, "layout": layout, "stream_bytes": sb,
                    "kernel": "k_ccg_spmv" if one_red else ("k_pcg_fspmv" if two else "k_pcg_spmv"),
                    "launches_per_iteration": info["launches_per_iteration"],
                    "us": float(avg_us[0]), "gbps": gbps(sb, avg_us[0]),
                    "csr_bytes": b1, "csr_gbps": gbps(b1, avg_us[0]),
                    "autotune_us_per_iter": [float(x) for x in tune_us],
                    "local_tiles": info["local_tiles"], "n_tiles": info["n_tiles"],
                    "nonzero_only_waves": info["sliced_nz_waves"], "waves": 4 * info["n_tiles"]},
           "iteration": {"kernel_us": {"spmv": float(avg_us[0]), "update": float(avg_us[1]),
                                       "finish": float(avg_us[2])},
                         "kernel_gbps_csr_accounting": {"spmv": gbps(b1, avg_us[0]),
                                                        "update": gbps(b2, avg_us[1]),
                                                        "finish": gbps(b3, avg_us[2])},
                         "us": float(avg_us.sum()),
                         "wall_us": elapsed / steps / int(res.iterations + 1) * 1e6,
                         "samples": [int(c) for c in cnt]}}
    system.close()
    return out


def run_hbm_only(H, ctx, N=100, iters=200):
    """The SpMV on a lattice whose layout exceeds the 256 MB last-level cache
    (100^3: about 595 MB with its nonzero-only waves, 6,000,000 unknowns):
    the autotuner's variant, HIP-event durations over `iters` iterations."""
    log(f"hbm-only: generating {N}^3 lattice")
    L = H.lattice(N, seed=0)
    s = H.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    s.set_rhs(L["rhs"], np.zeros(L["n"], np.float32))
    best, tune_us = s.autotune(20)
    info = s.info()
    s.set_kernel_timing(2)
    s.solve_resident(iters, 0.0, 0.0, raise_on_error=False)  # runs `iters` iterations
    avg_us, cnt = s.kernel_timing()
    s.close()
    sb, layout = spmv_stream_bytes(info, best)
    two = info["launches_per_iteration"] == 2
    kernel = "k_pcg_fspmv" if two else "k_pcg_spmv"
    if two:
        sb += 8 * L["n"]  # the finish folded in: the rows' z read and new p written
    g = sb / (avg_us[0] * 1e-6) / 1e9
    traffic = traffic_from_profiles(f"plain_N{N}")
    t_bytes = ((traffic or {}).get(f"{kernel}_bytes_by_variant") or {}).get(str(best))
    log(f"hbm-only {N}^3: variant {best}, {kernel} {avg_us[0]:.2f} us, {layout} {sb / 1e6:.1f} MB, {g:.0f} GB/s")
    return {"lattice": f"{N}^3", "n": L["n"], "nnz": L["nnz"], "kernel": kernel, "variant": best,
            "layout": layout, "stream_bytes": sb, "us": float(avg_us[0]), "achieved": g,
            "frac": g / HBM_PEAK_GBPS, "traffic": t_bytes, "csr_bytes": info["spmv_csr_bytes"],
            "nonzero_only_waves": info["sliced_nz_waves"], "waves": 4 * info["n_tiles"],
            "iteration_kernel_us": {"spmv": float(avg_us[0]), "update": float(avg_us[1]),
                                    "finish": float(avg_us[2])},
            "samples": int(cnt[0]), "autotune_us_per_iter": [float(x) for x in tune_us]}


# ----------------------------------------------------------------- CPU baselines

def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_lib
    oracle_lib.build()
    return oracle_lib


def cpu_baseline_batch(Ls, inner_tol, budget_s):
    """The batch workload on the host cores: for each system in order (FP32 LU
    build excluded, as on the GPU), FP64 refinement to 1e-8 around the
    oracle's FP32 restatement of the reference PCG (compute.rs:189-495) on the
    scaled residual; stops after the system during which `budget_s` elapsed."""
    O = _oracle()
    threads, cores = host_threads()
    O.set_threads(threads)
    done, t_solve, t_start = 0, 0.0, time.perf_counter()
    for L in Ls:
        st, lu = O.build_lu_blocks_6(L["n"], L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
        b = L["rhs"].astype(np.float64)
        t0 = time.perf_counter()
        x = np.zeros(L["n"])
        bn = np.linalg.norm(b)
        for _ in range(30):
            r = b - O_spmv64(O, L, x)
            rn = np.linalg.norm(r)
            if rn <= 1e-8 * bn:
                break
            b32 = (r / rn).astype(np.float32)
            _, _, d, _ = O.pcg(L["row_ptr"], L["col_idx"], L["values"], lu, L["block_starts"], b32,
                               np.zeros(L["n"], np.float32), 20000, inner_tol, 0.0)
            x += d.astype(np.float64) * rn
        t_solve += time.perf_counter() - t0
        done += 1
        if time.perf_counter() - t_start > budget_s:
            break
    return {"value": done / t_solve, "unit": "solves/s", "cores": threads, "kind": "port",
            "host": cores,
            "sample": (f"the first {done} of the {len(Ls)} systems of the same batch, each refined in "
                       f"FP64 to 1e-8 around the oracle's FP32 PCG (inner rel tol {inner_tol:g}), one "
                       f"after another, {t_solve:.1f} s of solves on {threads} threads")}


def O_spmv64(O, L, x):
    y = np.zeros(max(L["n"], 1))
    O.lib().or_spmv64(L["n"], L["row_ptr"], L["col_idx"], L["values"], np.ascontiguousarray(x), y)
    return y[: L["n"]]


def cpu_baseline_plain(L, iters_gpu, sample_iters, sample_iters_1core):
    """The oracle (FP32 restatement of the reference loop) on the host cores:
    `sample_iters` PCG iterations of the same system on all granted threads
    (the whole solve when it converges within them) and `sample_iters_1core`
    on one, scaled to solves/s with the GPU's iteration count."""
    O = _oracle()
    st, lu = O.build_lu_blocks_6(L["n"], L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    x0 = np.zeros(L["n"], np.float32)

    def sample(threads, iters):
        O.set_threads(threads)
        t0 = time.perf_counter()
        _, it, _, _ = O.pcg(L["row_ptr"], L["col_idx"], L["values"], lu, L["block_starts"],
                            L["rhs"], x0, iters, 1e-6, 0.0)
        dt = time.perf_counter() - t0
        per_iter = dt / max(it, 1)
        return 1.0 / (per_iter * (iters_gpu + 1)), it, dt, per_iter

    threads, cores = host_threads()
    side = round((L["n"] / 6) ** (1.0 / 3.0))
    v, it, dt, per_iter = sample(threads, sample_iters)
    v1, it1, dt1, per_iter1 = sample(1, sample_iters_1core)
    if it >= iters_gpu:
        v = 1.0 / dt
        what = f"the whole solve of the same {side}^3 system ({it} iterations) in {dt:.1f} s"
    else:
        what = (f"{it} PCG iterations of the same {side}^3 system in {dt:.1f} s "
                f"({per_iter * 1e3:.2f} ms/iteration), scaled to {iters_gpu} iterations")
    return {"value": v, "unit": "solves/s", "cores": threads, "kind": "port", "host": cores,
            "sample": what,
            "single_core": {"value": v1, "unit": "solves/s", "cores": 1,
                            "sample": f"{it1} iterations in {dt1:.1f} s ({per_iter1 * 1e3:.2f} ms/iteration)"}}


# ----------------------------------------------------------------- workloads

def roofline_from_plain(P, N, hbm_only=None):
    sp_ = P["spmv"]
    traffic = traffic_from_profiles(f"plain_N{N}")
    t_bytes = ((traffic or {}).get(f"{sp_['kernel']}_bytes_by_variant") or {}).get(str(P["best"]))
    r = {"bound": "hbm", "kernel": sp_["kernel"], "workload": f"plain FP32 PCG SpMV, {N}^3 lattice",
         "achieved": sp_["gbps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
         "frac": sp_["gbps"] / HBM_PEAK_GBPS, "traffic": t_bytes,
         "algorithmic_bytes": sp_["stream_bytes"], "kernel_us": sp_["us"],
         "csr_accounting": {"bytes": sp_["csr_bytes"], "gbps": sp_["csr_gbps"],
                            "frac": sp_["csr_gbps"] / HBM_PEAK_GBPS},
         "note": (f"achieved = the bytes SpMV variant {P['best']} streams per launch ({sp_['layout']} "
                  "layout: values -- in waves of rows with explicit zeros only the nonzero values "
                  "and each row's 8-byte mask of stored entries --, runs, column lists, headers, the "
                  "rows' own p and q; with two "
                  "launches per iteration also the rows' z and new p) / its mean "
                  "duration from HIP events on its own stream; csr_accounting divides the "
                  "reference's CSR bytes (8 nnz + 12 n + 4) by the same time; traffic = FETCH_SIZE + "
                  "WRITE_SIZE per launch from rocprofv3 (profiles/), which cannot separate "
                  "last-level-cache hits from HBM reads")}
    if hbm_only:
        r["hbm_only"] = hbm_only
    return r


def roofline_from_batch(B, P, HB):
    """The headline workload's dominant kernel: the table SpMV of the
    256-system batch (its layouts, about 0.4 GB, exceed the 256 MB last-level
    cache; the autotuner's non-temporal variant 4).  achieved = the bytes the
    sampled launches stream (each launch's running systems: their sliced
    layouts with the rows' own p and q) / the launches' HIP-event time, over
    an untimed solve of the same batch; the plain 50^3 point (and its 100^3
    HBM-only point) as before."""
    k = B["kernels"]["kernel"]["spmv"]
    traffic = traffic_from_profiles("batch256")
    t_bytes = (traffic or {}).get("k_pcg_spmv_bytes")
    r = {"bound": "hbm", "kernel": "k_pcg_spmv (table of systems)",
         "workload": "256-system batch, FP64 refinement to 1e-8: the SpMV launches of one solve",
         "achieved": k["gbps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": k["gbps"] / HBM_PEAK_GBPS,
         "traffic": t_bytes, "algorithmic_bytes": k["bytes_mean"], "kernel_us": k["us_mean"],
         "all_systems_running": k.get("all_systems"),
         "note": ("achieved = mean algorithmic bytes per launch (the running systems' sliced layouts: "
                  "values -- only the nonzero ones and an 8-byte mask per row in waves of rows with "
                  "explicit zeros --, runs, column lists, headers, the rows' own p and q) / mean "
                  "HIP-event duration, over the sampled launches of an untimed solve (which runs the "
                  "batch as one table, so each launch has the GPU to itself; the timed solves run "
                  "two tables side by side); traffic = "
                  "FETCH_SIZE + WRITE_SIZE per launch with all 256 systems running (rocprofv3, "
                  "profiles/traffic_batch256.json), to compare with all_systems_running.bytes"),
         "plain_fp32_50": roofline_from_plain(P, 50, HB)}
    return r


def bench_headline(args):
    import wgpu_solver_backend_amd as H
    ctx = H.GpuContext.create(0)
    log(ctx.describe())
    copy_bytes = 256 << 20
    copy_us = ctx.stream_copy_us(copy_bytes, 10)
    copy_gbps = 2 * copy_bytes / (copy_us * 1e-6) / 1e9
    B = run_batch(H, ctx, args, args.steps, args.warmup)
    S = run_serving(H, ctx, args) if not args.no_serve else None
    P = run_plain(H, ctx, 50, args.plain_steps, 1)
    HB = run_hbm_only(H, ctx, args.hbm_N) if not args.no_hbm_only else None
    bsum = B["summary"]
    out = {
        "metric": METRIC,
        "value": B["value"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": B["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"batched: {bsum['systems']} beam lattices, sides log-uniform [4,24] "
                                f"({bsum['unknowns_min']:,}..{bsum['unknowns_max']:,} unknowns, sum "
                                f"{bsum['unknowns_sum']:,}), FP32 block-Jacobi PCG (inner rel tol "
                                f"{args.inner_tol:g}) in FP64 refinement to a 1e-8 relative residual, "
                                f"all systems advanced together; one step = one solve of all "
                                f"{bsum['systems']} from x0 = 0"),
                   "systems": bsum["systems"], "target_rel_residual": 1e-8, "seed": 0},
        "latency_ms": {"p50": pct(B["latency_ms"], 50), "p99": pct(B["latency_ms"], 99),
                       "max": float(max(B["latency_ms"])),
                       "what": "each system's time from its step's start to the status poll that "
                               "showed its FP64-verified solution"},
        "verification": {k: bsum[k] for k in ("worst_true_rel_residual_fp64_device",
                                              "worst_true_rel_residual_fp64_host", "all_verified_1e-8")},
        "batch": dict(bsum, kernels=B["kernels"]),
        "serving": S,
        "plain_fp32_50": {"workload": "plain FP32 block-Jacobi PCG, 50^3 beam lattice (750,000 unknowns, "
                                      "30,352,200 nnz, 6x6 blocks), tol 1e-6 (recurrence), max 5000, x0 = 0",
                          "solves_per_s": P["value"], "iterations": int(P["res"].iterations),
                          "latency_ms": {"p50": statistics.median(P["lat"]), "max": max(P["lat"])},
                          "true_rel_residual_fp64": P["res"].rel_res_true_fp64,
                          "spmv": P["spmv"], "iteration": P["iteration"]},
        "stream_copy_gbps": copy_gbps,
        "roofline": roofline_from_batch(B, P, HB),
    }
    if not args.no_cpu:
        log("cpu baseline: the batch workload on the host cores")
        out["cpu_baseline"] = cpu_baseline_batch(B["Ls"], args.inner_tol, args.cpu_budget_s)
        log("cpu baseline: plain 50^3")
        out["plain_fp32_50"]["cpu_baseline"] = cpu_baseline_plain(
            P["L"], int(P["res"].iterations), args.cpu_iters, args.cpu_iters_1core)
    else:
        out["cpu_baseline"] = None
    ctx.close()
    return out


def bench_plain(args):
    import wgpu_solver_backend_amd as H
    ctx = H.GpuContext.create(0)
    P = run_plain(H, ctx, args.N, args.steps, args.warmup, args.alg)
    one_red = args.alg == "one-reduction"
    out = {"metric": f"solves/s (plain FP32 PCG, recurrence rel tol 1e-6); SpMV HBM GB/s vs peak",
           "value": P["value"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": P["elapsed"] / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": (f"plain FP32 block-Jacobi PCG, {args.N}^3 beam lattice ({P['n']:,} "
                                   f"unknowns, {P['nnz']:,} nnz, 6x6 blocks), tol 1e-6, max 5000, x0 = 0"
                                   + (", one-reduction (Chronopoulos-Gear) PCG" if one_red else "")),
                      "iterations": int(P["res"].iterations), "algorithm": args.alg},
           "latency_ms": {"p50": statistics.median(P["lat"]), "p99": pct(P["lat"], 99)},
           "true_rel_residual_fp64": P["res"].rel_res_true_fp64,
           "spmv": P["spmv"], "iteration": P["iteration"],
           "roofline": roofline_from_plain(P, args.N)}
    if one_red:
        out["roofline"]["traffic"] = None
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_plain(P["L"], int(P["res"].iterations), args.cpu_iters,
                                                 args.cpu_iters_1core)
    ctx.close()
    return out


def bench_profile(args):
    """Counter runs: `--profile-iters` PCG iterations per solve of the plain
    N^3 system with SpMV variant `--variant` (or the autotuner's), no JSON headline."""
    import wgpu_solver_backend_amd as H
    ctx = H.GpuContext.create(0)
    L = H.lattice(args.N, seed=0)
    s = H.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    s.set_rhs(L["rhs"], np.zeros(L["n"], np.float32))
    if args.variant is None:
        best, _ = s.autotune(20)
    else:
        best = args.variant
        s.set_spmv_variant(best)
    copy_bytes = 256 << 20
    copy_us = ctx.stream_copy_us(copy_bytes, 10)
    for _ in range(max(args.steps, 1)):
        r = s.solve_resident(args.profile_iters, 1e-6, 0.0, raise_on_error=False)
        log(f"profile run: status {r.status}, {r.iterations} its")
    info = s.info()
    sb, layout = spmv_stream_bytes(info, best)
    print(json.dumps({"profile_run": True, "N": args.N, "variant": best, "layout": layout,
                      "stream_bytes": sb, "csr_bytes": info["spmv_csr_bytes"],
                      "iterations_per_solve": args.profile_iters, "copy_bytes": copy_bytes,
                      "copy_us": copy_us}), flush=True)
    s.close()
    ctx.close()
    return None


def bench_profile_batch(args):
    """Counter runs of the table kernels: the 256-system batch in one
    refinement step of `--profile-iters` inner iterations per system (tolerances
    that cannot be met), so every launch works on all 256 systems (the two
    FP64 residual launches aside); the engine's SpMV variant (4, or
    HBJ_SPMV_VARIANT)."""
    import wgpu_solver_backend_amd as H
    ctx = H.GpuContext.create(0)
    sides = lattice_sides(0, args.batch, 4, 24)
    Ls = [H.lattice(N, seed=1000 + i) for i, N in enumerate(sides)]
    systems = [H.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"]) for L in Ls]
    for L, s in zip(Ls, systems):
        s.set_rhs64(L["rhs"].astype(np.float64))
    batch = H.Batch(ctx, systems)
    copy_bytes = 256 << 20
    copy_us = ctx.stream_copy_us(copy_bytes, 10)
    p = H.IrParams(max_outer=1, max_inner=args.profile_iters, inner_rel_tol=1e-30, target_rel_res=1e-300)
    for _ in range(max(args.steps, 1)):
        st, res = batch.solve_ir(p)
        log(f"profile batch run: status {st}, {res[0].iterations} inner iterations per system")
    variant = int(os.environ.get("HBJ_SPMV_VARIANT", "4"))
    per = np.zeros(3)
    for s in systems:
        info = s.info()
        per += (spmv_stream_bytes(info, variant)[0], 28 * info["n"] + 144 * info["n_blocks"], 12 * info["n"])
    print(json.dumps({"profile_run": True, "workload": "batch256", "variant": variant,
                      "iterations_per_solve": args.profile_iters,
                      "algorithmic_bytes": {"k_pcg_spmv": per[0], "k_pcg_update_grp": per[1],
                                            "k_pcg_finish": per[2]},
                      "copy_bytes": copy_bytes, "copy_us": copy_us}), flush=True)
    batch.close()
    for s in systems:
        s.close()
    ctx.close()
    return None


def bench_ir(args):
    """Mixed precision at 50^3 with E varied over 3 decades: FP32 PCG inside FP64
    refinement to a 1e-10 relative residual (verified on the host in FP64)."""
    import wgpu_solver_backend_amd as H
    N = args.N
    L = H.lattice(N, seed=0, e_log10_half_range=1.5)
    n, nnz = L["n"], L["nnz"]
    ctx = H.GpuContext.create(0)
    system = H.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
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
    host_res = host_rel_res(L, x)
    system.set_rhs(L["rhs"])
    r32 = system.solve_resident(20000, 1e-6, 0.0, raise_on_error=False)
    return {"metric": "solves/s at a 1e-10 FP64-verified relative residual (mixed precision)",
            "value": args.steps / elapsed, "unit": "solves/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
            "data": "synthetic",
            "config": {"workload": f"mixed precision: {N}^3 beam lattice, E over 3 decades, FP32 PCG "
                                   f"(inner rel tol {args.inner_tol:g}) in FP64 refinement to 1e-10",
                       "n": n, "nnz": nnz, "outer": int(res.outer_iterations),
                       "inner_iterations": int(res.iterations)},
            "latency_ms": {"p50": statistics.median(lat), "p99": pct(lat, 99)},
            "spmv": {"variant": best, "autotune_us_per_iter": [float(x) for x in tune_us]},
            "true_rel_residual_fp64_device": res.rel_res_true_fp64,
            "true_rel_residual_fp64_host": host_res,
            "plain_fp32_true_rel_residual_fp64": r32.rel_res_true_fp64,
            "plain_fp32_iterations": int(r32.iterations)}


def bench_batch(args):
    import wgpu_solver_backend_amd as H
    ctx = H.GpuContext.create(0)
    B = run_batch(H, ctx, args, args.steps, args.warmup)
    bsum = B["summary"]
    out = {"metric": METRIC, "value": B["value"], "unit": "solves/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": B["elapsed"] / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
           "data": "synthetic",
           "config": {"workload": f"batched: {bsum['systems']} beam lattices, sides log-uniform [4,24], "
                                  f"FP64 refinement to 1e-8 (inner rel tol {args.inner_tol:g})"},
           "latency_ms": {"p50": pct(B["latency_ms"], 50), "p99": pct(B["latency_ms"], 99)},
           "batch": bsum, "kernels": B["kernels"]}
    ctx.close()
    return out


def bench_serve(args):
    import wgpu_solver_backend_amd as H
    ctx = H.GpuContext.create(0)
    S = run_serving(H, ctx, args)
    a = S["at_target_rate"]
    ctx.close()
    return {"metric": METRIC, "value": a["completed_rate"], "unit": "solves/s", "n_gpus": 1,
            "steps": a["requests"], "warmup": 0, "ms_per_step": a["elapsed_s"] / a["requests"] * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
            "data": "synthetic", "config": {"workload": S["workload"]},
            "latency_ms": a["latency_ms"], "serving": S}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=50)
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "plain", "ir", "batch", "serve", "profile", "profile-batch"])
    ap.add_argument("--alg", default="pcg", choices=["pcg", "one-reduction"],
                    help="plain workload: the reference's PCG (default) or the one-reduction PCG")
    ap.add_argument("--plain-steps", type=int, default=5,
                    help="headline: timed plain 50^3 solves")
    ap.add_argument("--cpu-iters", type=int, default=5000,
                    help="most oracle PCG iterations of the plain system timed on the granted host "
                         "threads (the default lets the 50^3 solve converge: the whole solve is timed)")
    ap.add_argument("--cpu-iters-1core", type=int, default=400,
                    help="oracle PCG iterations of the plain system timed on one host thread (scaled)")
    ap.add_argument("--cpu-budget-s", type=float, default=12.0,
                    help="host seconds of the batch workload's CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-hbm-only", action="store_true")
    ap.add_argument("--hbm-N", type=int, default=100,
                    help="lattice side of the HBM-only SpMV point (layout past the last-level cache)")
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="inner FP32 PCG relative tolerance of the refinement loop (default: 1e-4 "
                         "for --workload ir, 1e-3 for the 1e-8 workloads)")
    ap.add_argument("--profile-iters", type=int, default=0,
                    help="counter runs: plain solves capped at this many iterations, no headline line")
    ap.add_argument("--variant", type=int, default=None, help="profile: SpMV variant (default: autotuned)")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--skip-per-system", action="store_true",
                    help="batch: do not also solve each system alone (profiled runs)")
    ap.add_argument("--requests", type=int, default=200)
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
    if args.profile_iters and args.workload != "profile-batch":
        args.workload = "profile"
    fn = {"headline": bench_headline, "plain": bench_plain, "ir": bench_ir, "batch": bench_batch,
          "serve": bench_serve, "profile": bench_profile,
          "profile-batch": bench_profile_batch}[args.workload]
    out = fn(args)
    if out is not None:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
