"""Mixed-precision refinement, batching and serving on the GPU.

Tolerances: the refined FP64 solution must reach the requested relative
residual ||b - A x||_2 / ||b||_2 (checked here on the host in FP64 from the
returned x); batched and served results must be bit-identical to the same
system solved alone (per-system reductions never mix systems)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu
F = np.float32


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if np.asarray(a).dtype == np.float32 else np.uint64)


def host_rel_res(L, x):
    n = L["n"]
    A = sp.csr_matrix((L["values"].astype(np.float64), L["col_idx"], L["row_ptr"]), shape=(n, n))
    b = L["rhs"].astype(np.float64)
    return np.linalg.norm(b - A @ x) / np.linalg.norm(b), A, b


@pytest.mark.parametrize("N,h", [(10, 1.5), (8, 0.0), (14, 1.5)])
def test_ir_reaches_1e10(ctx, hbj, N, h):
    print(f"IR N={N} h={h}", flush=True)
    L = hbj.lattice(N, seed=2, e_log10_half_range=h)
    s = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    p = hbj.IrParams(max_outer=30, max_inner=5000, inner_rel_tol=1e-4, target_rel_res=1e-10)
    x, res = s.solve_ir(L["rhs"].astype(np.float64), None, p)
    rr, A, b = host_rel_res(L, x)
    print(f"outer={res.outer_iterations} inner={res.iterations} dev_res={res.rel_res_true_fp64:.3e} "
          f"host_res={rr:.3e}", flush=True)
    assert res.status == 0 and res.rel_res_true_fp64 <= 1e-10
    assert rr <= 1.5e-10
    assert abs(rr - res.rel_res_true_fp64) <= 1e-3 * rr + 1e-14
    # error against an FP64 direct solve of the same FP32-stored matrix
    xd = spla.spsolve(A.tocsc(), b)
    err = np.linalg.norm(x - xd) / np.linalg.norm(xd)
    print(f"rel error vs FP64 direct: {err:.3e}", flush=True)
    assert err < 1e-6
    # plain FP32 PCG cannot get there on its own
    x32, r32 = s.solve(L["rhs"], np.zeros(L["n"], F), 5000, 1e-6, 0.0)
    assert host_rel_res(L, x32.astype(np.float64))[0] > 1e-8


def test_ir_warm_start_and_edges(ctx, hbj):
    L = hbj.lattice(6, seed=4)
    s = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    b = L["rhs"].astype(np.float64)
    p = hbj.IrParams(target_rel_res=1e-10)
    x, res = s.solve_ir(b, None, p)
    assert res.status == 0
    x2, res2 = s.solve_ir(b, x, p)  # already converged: no refinement step
    assert res2.status == 0 and res2.outer_iterations == 0 and np.array_equal(x2, x)
    xz, resz = s.solve_ir(np.zeros_like(b), None, p)  # b = 0 -> Ok, x = x0 = 0
    assert resz.status == 0 and np.all(xz == 0)
    with pytest.raises(hbj.SolverError) as e:
        s.solve_ir(b, None, hbj.IrParams(max_outer=1, target_rel_res=1e-15))
    assert e.value.status_name == "NOT_CONVERGED"
    # repeated solves are bit-identical
    x3, _ = s.solve_ir(b, None, p)
    assert np.array_equal(bits(x3), bits(x))


def _systems(ctx, hbj, sides, seed0=10, h=0.0):
    out = []
    for i, N in enumerate(sides):
        L = hbj.lattice(N, seed=seed0 + i, e_log10_half_range=h)
        out.append((L, hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])))
    return out


@pytest.mark.parametrize("fused", ["", "1"])
def test_batch_fp32_identical_to_alone(ctx, hbj, oracle, monkeypatch, fused):
    """A batch's table runs the three-launch iteration by default and the
    two-launch one with HBJ_FUSED=1: both give each system's alone bits."""
    monkeypatch.setenv("HBJ_FUSED", fused)
    print(f"batch of 7, HBJ_FUSED={fused!r}", flush=True)
    sides = [4, 9, 5, 12, 7, 4, 10]
    sysl = _systems(ctx, hbj, sides)
    alone = []
    for L, s in sysl:
        x, res = s.solve(L["rhs"], np.zeros(L["n"], F), 5000, 1e-6, 0.0)
        alone.append((x, res.iterations))
        s.set_rhs(L["rhs"])
    batch = hbj.Batch(ctx, [s for _, s in sysl])
    best, us = batch.autotune(10)  # any variant: the same bits
    assert 0 <= best < hbj.SPMV_VARIANTS and np.all(us > 0) and us[best] == us.min()
    st, results = batch.solve(5000, 1e-6, 0.0)
    assert st == 0
    for (L, s), (xa, ita), r in zip(sysl, alone, results):
        assert r.status == 0 and r.iterations == ita
        assert np.array_equal(bits(s.get_x()), bits(xa))
    # and against the oracle for one of them
    L, s = sysl[3]
    stl, lu = oracle.build_lu_blocks_6(L["n"], L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    sto, ito, xo, _ = oracle.pcg(L["row_ptr"], L["col_idx"], L["values"], lu, L["block_starts"],
                                 L["rhs"], np.zeros(L["n"], F), 5000, 1e-6, 0.0)
    assert results[3].iterations == ito and np.array_equal(bits(s.get_x()), bits(xo))
    batch.close()


def test_batch_kernel_samples(ctx, hbj):
    """hbj_batch_kernel_samples: with every launch set sampled, one record
    per set in which some system ran (sets numbered from 0 in order), three
    positive durations each (the three-launch table iteration), and per system
    the launch sets that found it running: its init set plus one per
    iteration.  The sampled solve gives the unsampled solve's bits."""
    sides = [4, 7, 5]
    sysl = _systems(ctx, hbj, sides)
    for L, s in sysl:
        s.set_rhs(L["rhs"])
    batch = hbj.Batch(ctx, [s for _, s in sysl])
    st, ref = batch.solve(5000, 1e-6, 0.0)
    assert st == 0
    xr = [s.get_x().copy() for _, s in sysl]
    batch.set_kernel_timing(1)
    st, res = batch.solve(5000, 1e-6, 0.0)
    batch.set_kernel_timing(0)
    assert st == 0
    rec, sets = batch.kernel_samples()
    its = [r.iterations for r in res]
    print(f"samples {len(rec)}, sets {sets.tolist()}, iterations {its}", flush=True)
    assert [r.iterations for r in ref] == its
    assert sets.tolist() == [i + 1 for i in its]
    assert len(rec) == max(sets)
    assert np.array_equal(rec[:, 0], np.arange(len(rec)))
    assert np.all(rec[:, 1:] > 0.0)
    for (L, s), x in zip(sysl, xr):
        assert np.array_equal(bits(s.get_x()), bits(x))
    batch.close()


def test_two_batches_on_two_threads(ctx, hbj):
    """Two batches of one context solved at once from two host threads (each
    batch also runs its own two tables on a helper thread): every engine
    captures and replays its graphs while the others do, and compacts its
    table with copies on its own stream -- a synchronous copy on the legacy
    stream is refused while another thread captures.  Results equal the
    sequential solves bit for bit."""
    import threading
    groups = [[4, 9, 6, 11, 5], [10, 4, 8, 7, 12]]
    batches, sysls = [], []
    for gi, sides in enumerate(groups):
        sysl = _systems(ctx, hbj, sides, seed0=70 + 10 * gi)
        for L, s in sysl:
            s.set_rhs64(L["rhs"].astype(np.float64))
        sysls.append(sysl)
        batches.append(hbj.Batch(ctx, [s for _, s in sysl]))
    p = hbj.IrParams(max_outer=30, max_inner=20000, inner_rel_tol=1e-3, target_rel_res=1e-8)
    seq = []
    for b, sysl in zip(batches, sysls):
        st, res = b.solve_ir(p)
        assert st == 0
        seq.append(([r.iterations for r in res], [s.get_x64().copy() for _, s in sysl]))
    out = [None, None]

    def work(k):
        for _ in range(3):
            out[k] = batches[k].solve_ir(p)
            if out[k][0] != 0:
                return

    th = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for k in range(2):
        st, res = out[k]
        assert st == 0, hbj._lib.last_error()
        assert [r.iterations for r in res] == seq[k][0]
        for (L, s), x in zip(sysls[k], seq[k][1]):
            assert np.array_equal(bits(s.get_x64()), bits(x))
    for b in batches:
        b.close()


def test_batch_isolates_failures(ctx, hbj):
    sysl = _systems(ctx, hbj, [5, 6])
    n = 12
    ident = hbj.System(ctx, np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32),
                       np.ones(n, F), np.arange(0, n + 1, 6, dtype=np.uint32))
    b = np.arange(1, n + 1, dtype=F)
    ident.set_rhs(b, b)  # x0 = b: p.Ap = 0 breakdown
    for L, s in sysl:
        s.set_rhs(L["rhs"])
    batch = hbj.Batch(ctx, [sysl[0][1], ident, sysl[1][1]])
    st, results = batch.solve(5000, 1e-6, 0.0)
    assert st == 3
    assert results[0].status == 0 and results[2].status == 0 and results[1].status == 3
    for (L, s), r in zip(sysl, (results[0], results[2])):
        x, ra = s.solve(L["rhs"], np.zeros(L["n"], F), 5000, 1e-6, 0.0)
        assert ra.iterations == r.iterations


@pytest.mark.parametrize("fused", ["", "1"])
def test_batch_ir_identical_to_alone(ctx, hbj, monkeypatch, fused):
    monkeypatch.setenv("HBJ_FUSED", fused)
    print(f"refined batch of 5, HBJ_FUSED={fused!r}", flush=True)
    sides = [4, 8, 5, 11, 6]
    sysl = _systems(ctx, hbj, sides, seed0=30, h=1.5)
    p = hbj.IrParams(target_rel_res=1e-10)
    alone = []
    for L, s in sysl:
        x, res = s.solve_ir(L["rhs"].astype(np.float64), None, p)
        alone.append((x, res))
        s.set_rhs64(L["rhs"].astype(np.float64))
    batch = hbj.Batch(ctx, [s for _, s in sysl])
    best, us = batch.autotune(10, ir=True)
    assert 0 <= best < hbj.SPMV_VARIANTS and np.all(us > 0)
    st, results = batch.solve_ir(p)
    assert st == 0
    for (L, s), (xa, ra), r in zip(sysl, alone, results):
        assert r.status == 0 and r.iterations == ra.iterations and r.outer_iterations == ra.outer_iterations
        xb = s.get_x64()
        assert np.array_equal(bits(xb), bits(xa))
        assert host_rel_res(L, xb)[0] <= 1.5e-10
    batch.close()


def test_batch_ir_partials_paths(ctx, hbj):
    """A refined batch whose systems take each of the table kernels' ways to
    their partials: one tile (no reduction), one chunk of at most 256 tiles
    and two chunks (the prefetch before the state), more than 512 tiles (the
    reduction after the state) -- each system's bits equal its alone solve's,
    which never takes the table kernels."""
    sides = [3, 7, 24, 28]  # 1, 8, 324 and 515 tiles
    sysl = _systems(ctx, hbj, sides, seed0=50)
    tiles = [s.info()["n_tiles"] for _, s in sysl]
    print(f"refined batch, tiles {tiles}", flush=True)
    assert tiles[0] == 1 and tiles[1] <= 256 < tiles[2] <= 512 < tiles[3]
    p = hbj.IrParams(max_outer=30, max_inner=20000, inner_rel_tol=1e-3, target_rel_res=1e-8)
    alone = []
    for L, s in sysl:
        x, res = s.solve_ir(L["rhs"].astype(np.float64), None, p)
        alone.append((x, res))
        s.set_rhs64(L["rhs"].astype(np.float64))
    batch = hbj.Batch(ctx, [s for _, s in sysl])
    st, results = batch.solve_ir(p)
    assert st == 0
    for (L, s), (xa, ra), r in zip(sysl, alone, results):
        assert r.status == 0 and r.iterations == ra.iterations and r.outer_iterations == ra.outer_iterations
        xb = s.get_x64()
        assert np.array_equal(bits(xb), bits(xa))
        assert host_rel_res(L, xb)[0] <= 1e-8 * 1.001
    batch.close()


def test_server_mixed_stream_identical_to_alone(ctx, hbj):
    rng = np.random.default_rng(0)
    sides = [int(round(np.exp(rng.uniform(np.log(4), np.log(12))))) for _ in range(24)]
    p = hbj.IrParams(target_rel_res=1e-8)
    systems = [hbj.lattice(N, seed=100 + i) for i, N in enumerate(sides)]
    # reference: each alone
    alone = []
    for L in systems:
        s = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
        x, res = s.solve_ir(L["rhs"].astype(np.float64), None, p)
        alone.append((x, res))
        s.close()
    budget = 64 << 20  # forces queueing: not all requests fit at once
    srv = hbj.Server(ctx, budget, max_active=6, chunk_iters=8)
    tickets = [srv.submit(L["row_ptr"], L["col_idx"], L["values"], L["block_starts"],
                          L["rhs"].astype(np.float64), p) for L in systems]
    lat = []
    for t, L, (xa, ra) in zip(tickets, systems, alone):
        st, x, res, ms = srv.wait(t, n=L["n"])
        assert st == 0 and res.status == 0
        assert res.iterations == ra.iterations
        assert np.array_equal(bits(x), bits(xa))
        assert host_rel_res(L, x)[0] <= 1.5e-8
        lat.append(ms)
    stats = srv.stats()
    assert stats["completed"] == len(systems) and stats["submitted"] == len(systems)
    assert stats["peak_device_bytes"] <= budget
    assert 1 < stats["peak_active"] <= 6
    assert min(lat) > 0
    srv.close()


def test_server_rejects_and_reports_failures(ctx, hbj):
    """A request whose device-byte lower bound (its nonzero values, LU blocks
    and vectors: 9.5 MB for the 22^3 lattice) exceeds the 8 MiB budget is
    rejected at submit; a zero pivot fails its own request only."""
    srv = hbj.Server(ctx, 8 << 20, max_active=4)
    big = hbj.lattice(22)
    with pytest.raises(hbj.SolverError) as e:
        srv.submit(big["row_ptr"], big["col_idx"], big["values"], big["block_starts"],
                   big["rhs"].astype(np.float64))
    assert e.value.status_name == "REJECTED"
    ok = hbj.lattice(4)
    t1 = srv.submit(ok["row_ptr"], ok["col_idx"], ok["values"], ok["block_starts"],
                    ok["rhs"].astype(np.float64))
    rp = np.array([0, 1, 2], np.uint32)
    t2 = srv.submit(rp, np.array([0, 1], np.uint32), np.array([0, 1], F), np.array([0, 2], np.uint32),
                    np.ones(2))
    st2, _, res2, _ = srv.wait(t2, n=2)
    assert st2 == 2 and "LU zero pivot at k = 0" in hbj._lib.last_error()
    st1, x1, res1, _ = srv.wait(t1, n=ok["n"])
    assert st1 == 0
    s = srv.stats()
    assert s["rejected"] == 1 and s["completed"] == 2
    srv.close()


def test_server_waits_for_a_free_range(ctx, hbj):
    """The arena holds exactly the budget.  A (short solve) takes the front,
    B (long solve) the next block; C fits the budget once A is done but not
    into either free range (A's old block, the tail), so it waits until B
    gives its block back -- and every result is still the one solved alone."""
    print("server arena: sides 10 / 12 / 11, E over 3 decades", flush=True)
    p = hbj.IrParams(target_rel_res=1e-8)
    Ls = [hbj.lattice(N, seed=200 + i, e_log10_half_range=1.5) for i, N in enumerate((10, 12, 11))]
    alone = []
    for L in Ls:
        s = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
        alone.append(s.solve_ir(L["rhs"].astype(np.float64), None, p))
        s.close()
    assert alone[0][1].iterations < alone[1][1].iterations  # A finishes first
    # admission reserves each request's exact block, and the arena places it
    a, b, c = (hbj.Server.request_bytes(L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
               for L in Ls)
    assert a < c
    budget = a + b + c - a // 2  # free ranges once A is done: A's block and a tail of c - a/2
    assert a + b <= budget < a + b + c  # A and B admitted at once, C only after A
    assert b + c <= budget  # ... and then admitted, but placed only once B is done
    srv = hbj.Server(ctx, budget, max_active=4, chunk_iters=8)
    tickets = [srv.submit(L["row_ptr"], L["col_idx"], L["values"], L["block_starts"],
                          L["rhs"].astype(np.float64), p) for L in Ls]
    for t, L, (xa, ra) in zip(tickets, Ls, alone):
        st, x, res, ms = srv.wait(t, n=L["n"])
        assert st == 0 and res.iterations == ra.iterations
        assert np.array_equal(bits(x), bits(xa))
    stats = srv.stats()
    assert stats["arena_waits"] >= 1
    assert stats["peak_device_bytes"] <= budget
    assert stats["upload_bytes"] > 0
    srv.close()


def test_server_keeps_only_the_sliced_layout(ctx, hbj):
    """A served lattice system keeps its sliced SpMV layout and no CSR arrays
    on the device: its block is under 60% of the same system created for
    standalone solves (CSR, entry indices, both layouts), and a request whose
    exact block exceeds the budget completes with REJECTED."""
    L = hbj.lattice(10, seed=7)
    s = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    s.set_rhs64(L["rhs"].astype(np.float64))  # the refinement block too
    standalone = s.info()["device_bytes"]
    s.close()
    served = hbj.Server.request_bytes(L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    print(f"10^3 lattice: served block {served} B, standalone {standalone} B", flush=True)
    assert served < 0.6 * standalone
    # above the budget only by its exact size: admitted as far as the bound, rejected later
    srv = hbj.Server(ctx, served - 4096, max_active=2)
    t = srv.submit(L["row_ptr"], L["col_idx"], L["values"], L["block_starts"],
                   L["rhs"].astype(np.float64))
    st, _, _, _ = srv.wait(t, n=L["n"])
    assert st == 9 and "above the admission budget" in hbj._lib.last_error()
    assert srv.stats()["rejected"] == 1
    srv.close()
