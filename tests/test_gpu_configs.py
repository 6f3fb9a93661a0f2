"""The benchmark configurations at their configured workloads on the GPU.

* mixed precision: the 50^3 lattice with E over 3 decades, FP32 PCG inside
  FP64 refinement to a 1e-10 relative residual;
* batched: 256 lattices with sides log-uniform in [4, 24], FP64 refinement
  to 1e-8, all systems advanced together;
* serving: a seeded Poisson stream at 200 requests/s of lattices with sides
  log-uniform in [6, 40], FP64 refinement to 1e-8, continuous batching.

Every result is checked three ways:
  1. bit-identical to the same system solved alone through the per-system
     path (batching and serving never mix systems' arithmetic);
  2. the relative residual ||b - A x||_2 / ||b||_2 recomputed on the host in
     FP64 from the returned x reaches the target (the device stops on its own
     FP64 evaluation, <= target; the two FP64 evaluations differ only in
     summation order, so the host value is allowed 0.1% above it);
  3. the error against an FP64 CPU reference: the oracle's FP64 block-Jacobi
     PCG run to a 1e-12 relative residual, compared with the metric of
     compare_x_vectors (cli main.rs:603-662: FP64-accumulated relative L2
     norm of x - x_ref), bound 1e-6.  At a 1e-8 residual these lattices give
     errors around 1e-8 (measured with the FP64 oracle at both tolerances);
     the bound leaves two orders of margin.
The FP64 reference costs seconds per large system on the host, so it runs on
a stated subset of the results (every 8th system of the batch; every served
request of side <= 24 and the largest one); checks 1 and 2 run on all."""
import threading
import time

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu
ERR_BOUND = 1e-6
RES_SLACK = 1.001


def bits64(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def sides_logu(seed, count, lo, hi):
    """Side lengths round(exp(U[ln lo, ln hi])) from numpy's default_rng(seed),
    as the benchmark draws them."""
    rng = np.random.default_rng(seed)
    return [int(round(np.exp(rng.uniform(np.log(lo), np.log(hi))))) for _ in range(count)]


def host_rel_res(L, x):
    n = L["n"]
    A = sp.csr_matrix((L["values"].astype(np.float64), L["col_idx"], L["row_ptr"]), shape=(n, n))
    b = L["rhs"].astype(np.float64)
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))


def fp64_reference(oracle, L, tol=1e-12):
    """The oracle's FP64 PCG from x0 = 0 to a `tol` relative residual."""
    b = L["rhs"].astype(np.float64)
    st, it, xr = oracle.pcg64(L["row_ptr"], L["col_idx"], L["values"], L["block_starts"], b,
                              np.zeros(L["n"]), 60000, tol)
    assert st == 0, f"FP64 reference did not reach {tol:g} (n={L['n']})"
    return xr


def rel_l2(x, xr):
    """compare_x_vectors' relative L2 error (FP64 accumulation)."""
    d = np.asarray(x, np.float64) - xr
    return float(np.sqrt(np.dot(d, d)) / np.sqrt(np.dot(xr, xr)))


@pytest.mark.slow
def test_config3_ir_50_cubed_to_1e10(ctx, hbj):
    """50^3 lattice (750,000 unknowns), E scaled per element by 10^U[-1.5,1.5]:
    FP64 refinement around FP32 PCG reaches 1e-10, verified on the host."""
    print("config 3: 50^3, E over 3 decades, IR to 1e-10", flush=True)
    L = hbj.lattice(50, seed=0, e_log10_half_range=1.5)
    s = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    s.set_spmv_variant(5)
    p = hbj.IrParams(max_outer=40, max_inner=20000, inner_rel_tol=1e-4, target_rel_res=1e-10)
    x, res = s.solve_ir(L["rhs"].astype(np.float64), None, p)
    hr = host_rel_res(L, x)
    print(f"outer={res.outer_iterations} inner={res.iterations} dev={res.rel_res_true_fp64:.3e} "
          f"host={hr:.3e} ms={res.solve_ms:.0f}", flush=True)
    assert res.status == 0 and res.rel_res_true_fp64 <= 1e-10
    assert hr <= 1e-10 * RES_SLACK
    # repeated solve (another variant): the same bits
    s.set_spmv_variant(4)
    x2, res2 = s.solve_ir(L["rhs"].astype(np.float64), None, p)
    assert res2.iterations == res.iterations and np.array_equal(bits64(x2), bits64(x))
    s.close()


def test_ir_1e10_error_vs_fp64_reference(ctx, hbj, oracle):
    """The refined solution against the FP64 reference at a side where the
    reference is cheap (20^3, E over 3 decades).  With E spread over 3
    decades the FP64 PCG stagnates near a 1e-12 residual, so the reference is
    taken at 1e-11."""
    L = hbj.lattice(20, seed=3, e_log10_half_range=1.5)
    s = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    p = hbj.IrParams(max_outer=40, max_inner=20000, inner_rel_tol=1e-4, target_rel_res=1e-10)
    x, res = s.solve_ir(L["rhs"].astype(np.float64), None, p)
    assert res.status == 0 and host_rel_res(L, x) <= 1e-10 * RES_SLACK
    err = rel_l2(x, fp64_reference(oracle, L, 1e-11))
    print(f"20^3 IR 1e-10: rel L2 error vs FP64 reference {err:.3e}", flush=True)
    assert err <= 1e-8
    s.close()


def test_ir_residual_paths_bitwise(ctx, hbj):
    """The FP64 refinement residual reads the CSR arrays under SpMV variants
    0-3 and the sliced layout under 4 and 5: the same products in the same
    order, so every variant refines to the same bits."""
    L = hbj.lattice(12, seed=9, e_log10_half_range=1.5)
    s = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    assert s.info()["spmv_sliced_bytes"] > 0
    p = hbj.IrParams(max_outer=30, max_inner=20000, inner_rel_tol=1e-4, target_rel_res=1e-10)
    ref = None
    for v in range(hbj.SPMV_VARIANTS):
        s.set_spmv_variant(v)
        x, res = s.solve_ir(L["rhs"].astype(np.float64), None, p)
        assert res.status == 0
        if ref is None:
            ref = (x, res.iterations, res.outer_iterations)
        assert (res.iterations, res.outer_iterations) == ref[1:], v
        assert np.array_equal(bits64(x), bits64(ref[0])), v
    s.close()


@pytest.mark.slow
def test_config4_batch_256(ctx, hbj, oracle):
    """256 lattices, sides log-uniform in [4, 24] (384 .. 82,944 unknowns),
    FP64 refinement to 1e-8 with FP32 inner PCG (rel tol 1e-3), the batch's
    SpMV variant chosen by the autotuner."""
    sides = sides_logu(0, 256, 4, 24)
    print(f"config 4: 256 systems, sides {min(sides)}..{max(sides)}", flush=True)
    Ls = [hbj.lattice(N, seed=1000 + i) for i, N in enumerate(sides)]
    systems = [hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"]) for L in Ls]
    for L, s in zip(Ls, systems):
        s.set_rhs64(L["rhs"].astype(np.float64))
    p = hbj.IrParams(max_outer=30, max_inner=20000, inner_rel_tol=1e-3, target_rel_res=1e-8)
    batch = hbj.Batch(ctx, systems)
    best, us = batch.autotune(20, ir=True)
    t0 = time.perf_counter()
    st, results = batch.solve_ir(p)
    dt = time.perf_counter() - t0
    print(f"batch: variant {best}, {dt * 1e3:.0f} ms, {len(systems) / dt:.0f} solves/s", flush=True)
    assert st == 0
    xb = [s.get_x64() for s in systems]
    worst_res = 0.0
    for i, (L, r, x) in enumerate(zip(Ls, results, xb)):
        assert r.status == 0 and r.rel_res_true_fp64 <= 1e-8, i
        hr = host_rel_res(L, x)
        worst_res = max(worst_res, hr)
        assert hr <= 1e-8 * RES_SLACK, (i, hr)
    # alone, through the per-system path: the same bits and counts
    for i, (L, s, r, x) in enumerate(zip(Ls, systems, results, xb)):
        xa, ra = s.solve_ir(L["rhs"].astype(np.float64), None, p)
        assert ra.iterations == r.iterations and ra.outer_iterations == r.outer_iterations, i
        assert np.array_equal(bits64(xa), bits64(x)), i
    errs = []
    for i in range(0, len(Ls), 8):
        errs.append(rel_l2(xb[i], fp64_reference(oracle, Ls[i])))
    print(f"worst host residual {worst_res:.3e}; error vs FP64 reference on {len(errs)} systems: "
          f"max {max(errs):.3e}, median {np.median(errs):.3e}", flush=True)
    assert max(errs) <= ERR_BOUND
    batch.close()
    for s in systems:
        s.close()


def poisson_serve(hbj, ctx, Ls, rate, seed, params, budget, max_active=256):
    """Submits the requests with exponential gaps (mean 1/rate, default_rng(seed))
    from one thread while another waits for them in order; returns per request
    (status, x, result, latency_ms) and the server's stats."""
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, len(Ls))
    srv = hbj.Server(ctx, budget, max_active=max_active, chunk_iters=64)
    tickets = [None] * len(Ls)
    out = [None] * len(Ls)

    def waiter():
        for i in range(len(Ls)):
            while tickets[i] is None:
                time.sleep(0.0005)
            out[i] = srv.wait(tickets[i], n=Ls[i]["n"])

    th = threading.Thread(target=waiter)
    th.start()
    t_next = time.perf_counter()
    for i, L in enumerate(Ls):
        t_next += gaps[i]
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        tickets[i] = srv.submit(L["row_ptr"], L["col_idx"], L["values"], L["block_starts"],
                                L["rhs"].astype(np.float64), params)
    th.join()
    stats = srv.stats()
    srv.close()
    return out, stats


@pytest.mark.slow
def test_config5_serving_poisson_200(ctx, hbj, oracle):
    """48 requests of a seeded Poisson stream at 200 requests/s, sides
    log-uniform in [6, 40] (1,296 .. 384,000 unknowns), FP64 refinement to
    1e-8 (FP32 inner rel tol 1e-3), 16 GiB admission budget."""
    sides = sides_logu(0, 48, 6, 40)
    print(f"config 5: 48 requests at 200 req/s, sides {min(sides)}..{max(sides)}", flush=True)
    Ls = [hbj.lattice(N, seed=2000 + i) for i, N in enumerate(sides)]
    p = hbj.IrParams(max_outer=30, max_inner=20000, inner_rel_tol=1e-3, target_rel_res=1e-8)
    out, stats = poisson_serve(hbj, ctx, Ls, 200.0, 1, p, 16 << 30)
    lat = [o[3] for o in out]
    print(f"served: p50 {np.percentile(lat, 50):.0f} ms, p99 {np.percentile(lat, 99):.0f} ms, "
          f"peak active {stats['peak_active']}", flush=True)
    assert stats["completed"] == len(Ls) and stats["rejected"] == 0
    assert stats["peak_active"] > 1  # requests overlapped on the device
    largest = int(np.argmax(sides))
    errs = []
    for i, (L, (st, x, r, ms)) in enumerate(zip(Ls, out)):
        assert st == 0 and r.status == 0 and r.rel_res_true_fp64 <= 1e-8, i
        assert host_rel_res(L, x) <= 1e-8 * RES_SLACK, i
        s = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
        xa, ra = s.solve_ir(L["rhs"].astype(np.float64), None, p)
        s.close()
        assert ra.iterations == r.iterations and np.array_equal(bits64(xa), bits64(x)), i
        if sides[i] <= 24 or i == largest:
            errs.append(rel_l2(x, fp64_reference(oracle, L)))
    print(f"error vs FP64 reference on {len(errs)} requests: max {max(errs):.3e}", flush=True)
    assert max(errs) <= ERR_BOUND
