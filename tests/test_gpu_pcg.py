"""The drop-in solve on the GPU against the oracle: same status, same
iteration count, bit-identical x (the FP32 path follows the reference's
operation order), the reference's golden x_ref tolerances, error paths."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu
F = np.float32


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def oracle_solve(oracle, rp, ci, va, bs, b, x0, max_iter, rel, ab):
    n = len(rp) - 1
    st, lu = oracle.build_lu_blocks_6(n, rp, ci, va, bs)
    assert st == 0
    return oracle.pcg(rp, ci, va, lu, bs, b, x0, max_iter, rel, ab)


DATASETS = [(1, 1e-2, 42), (2, 1e0, 370), (3, 1e1, 1273)]


@pytest.mark.parametrize("k,atol,its", DATASETS)
def test_datasets_bitwise_and_x_ref(ctx, hbj, oracle, k, atol, its):
    print(f"dataset {k}", flush=True)
    d = oracle.load_dataset(k)
    sys_ = hbj.System(ctx, d["row_ptr"], d["col_idx"], d["values"], d["block_starts"])
    x, res = sys_.solve(d["b"], d["x0"], 2000, 1e-4, 1e-7)
    st, it, xo, _ = oracle_solve(oracle, d["row_ptr"], d["col_idx"], d["values"], d["block_starts"],
                                 d["b"], d["x0"], 2000, 1e-4, 1e-7)
    assert res.status == st == 0
    assert res.iterations == it == its
    assert np.array_equal(bits(x), bits(xo))
    xr = d["x_ref"].astype(np.float64)
    err = x.astype(np.float64) - xr
    rel_l2 = np.linalg.norm(err) / np.linalg.norm(xr)
    assert rel_l2 <= 1e-3 or np.sqrt(np.mean(err ** 2)) <= atol  # commands.txt:55-78
    # the FP64 residual the library reports for x
    n = int(d["n"])
    A = sp.csr_matrix((d["values"].astype(np.float64), d["col_idx"], d["row_ptr"]), shape=(n, n))
    b64 = d["b"].astype(np.float64)
    tr = np.linalg.norm(b64 - A @ x.astype(np.float64)) / np.linalg.norm(b64)
    assert abs(res.rel_res_true_fp64 - tr) <= 1e-9 * max(tr, 1e-12)
    sys_.close()


def test_reference_named_entry_point(ctx, hbj, oracle):
    d = oracle.load_dataset(2)
    n = int(d["n"])
    lu = hbj.build_lu_blocks_from_csr_block_starts_6(n, d["row_ptr"], d["col_idx"], d["values"],
                                                     d["block_starts"])
    spmv = hbj.SpmvExecutor.create(ctx, n, d["row_ptr"], d["col_idx"], d["values"])
    vec = hbj.VecOpsExecutor.create(ctx)
    dot = hbj.DotScalarExecutor.create(ctx, n, 7)
    bj = hbj.BlockJacobiExecutor.create(ctx, n, lu, d["block_starts"])
    upd = hbj.PcgUpdateScalarsExecutor.create(ctx)
    x = d["x0"].copy()
    it = hbj.pcg_block_jacobi_csr_wgpu(n, d["b"], x, 2000, 1e-4, 1e-7, ctx, spmv, vec, dot, bj, upd)
    st, ito, xo, _ = oracle_solve(oracle, d["row_ptr"], d["col_idx"], d["values"], d["block_starts"],
                                  d["b"], d["x0"], 2000, 1e-4, 1e-7)
    assert it == ito and np.array_equal(bits(x), bits(xo))
    with pytest.raises(hbj.SolverError, match="dimension mismatch"):
        hbj.pcg_block_jacobi_csr_wgpu(n, d["b"][:-1], x, 10, 1e-4, 0, ctx, spmv, vec, dot, bj, upd)


@pytest.mark.parametrize("N,h,seed", [(10, 0.0, 0), (6, 1.5, 4), (13, 0.0, 2)])
def test_lattice_bitwise(ctx, hbj, oracle, N, h, seed):
    print(f"lattice N={N} h={h}", flush=True)
    L = hbj.lattice(N, seed=seed, e_log10_half_range=h)
    sys_ = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    x0 = np.zeros(L["n"], F)
    x, res = sys_.solve(L["rhs"], x0, 5000, 1e-6, 0.0)
    st, it, xo, _ = oracle_solve(oracle, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"],
                                 L["rhs"], x0, 5000, 1e-6, 0.0)
    assert res.status == st == 0 and res.iterations == it
    assert np.array_equal(bits(x), bits(xo))
    # a second solve of the same resident system is bit-identical (no state leaks)
    x2, res2 = sys_.solve(L["rhs"], x0, 5000, 1e-6, 0.0)
    assert res2.iterations == it and np.array_equal(bits(x2), bits(x))
    sys_.close()


def test_resident_api_and_warm_start(ctx, hbj, oracle):
    L = hbj.lattice(8, seed=3)
    n = L["n"]
    sys_ = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    x0 = (np.random.default_rng(1).standard_normal(n) * 1e-4).astype(F)
    x0[: 6 * 64] = 0
    sys_.set_rhs(L["rhs"], x0)
    res = sys_.solve_resident(5000, 1e-6, 0.0)
    x = sys_.get_x()
    st, it, xo, _ = oracle_solve(oracle, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"],
                                 L["rhs"], x0, 5000, 1e-6, 0.0)
    assert res.iterations == it and np.array_equal(bits(x), bits(xo))
    # kernel timing mode changes nothing numerically
    sys_.set_kernel_timing(4)
    res_t = sys_.solve_resident(5000, 1e-6, 0.0)
    avg, cnt = sys_.kernel_timing()
    assert res_t.iterations == it and np.array_equal(bits(sys_.get_x()), bits(xo))
    k = sys_.info()["launches_per_iteration"]
    assert k in (2, 3)
    assert np.all(cnt[:k] > 0) and np.all(avg[:k] > 0) and np.all(cnt[k:] == 0)
    # samples of launch sets 0, 4, 8, ... that found the system running (the
    # chunk's trailing launches drop out): with three launches the init set
    # and one per iteration; with two, also the set whose first kernel
    # finishes the last iteration
    sets = it + 1 if k == 3 else it + 2
    assert np.all(cnt[:k] == -(-sets // 4)), (cnt, it)
    sys_.set_kernel_timing(0)


def test_zero_rhs_ok0_x_untouched(ctx, hbj, oracle):
    d = oracle.load_dataset(1)
    sys_ = hbj.System(ctx, d["row_ptr"], d["col_idx"], d["values"], d["block_starts"])
    x0 = np.arange(int(d["n"]), dtype=F)
    x, res = sys_.solve(np.zeros(int(d["n"]), F), x0, 100, 1e-4, 0.0)
    assert res.status == 0 and res.iterations == 0 and np.array_equal(x, x0)


def test_breakdown_pap(ctx, hbj):
    n = 12
    rp = np.arange(n + 1, dtype=np.uint32)
    ci = np.arange(n, dtype=np.uint32)
    va = np.ones(n, F)
    sys_ = hbj.System(ctx, rp, ci, va, np.arange(0, n + 1, 6, dtype=np.uint32))
    b = np.arange(1, n + 1, dtype=F)
    with pytest.raises(hbj.SolverError) as e:
        sys_.solve(b, b, 100, 1e-6, 0.0)
    assert str(e.value) == "PCG(BlockJacobiGpu): dot(p,Ap) is zero (breakdown)"
    assert e.value.result.iterations == 1


def test_breakdown_rz(ctx, hbj):
    rp = np.array([0, 2, 4, 5], np.uint32)
    ci = np.array([0, 1, 0, 1, 2], np.uint32)
    va = np.array([1, 0.5, 0.5, -1, 1], F)
    sys_ = hbj.System(ctx, rp, ci, va, np.array([0, 1, 2, 3], np.uint32))
    x0 = np.zeros(3, F)
    x, res = sys_.solve(np.array([1, 1, 0], F), x0, 10, 1e-6, 0.0, raise_on_error=False)
    assert res.status == 4 and res.iterations == 1 and np.array_equal(x, x0)
    assert hbj._lib.last_error() == "PCG(BlockJacobiGpu): rz_old is zero (breakdown)"


@pytest.mark.parametrize("max_iter", [0, 1, 7])
def test_not_converged_keeps_x0(ctx, hbj, max_iter):
    L = hbj.lattice(5, seed=1)
    sys_ = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    x0 = np.full(L["n"], 0.5, F)
    x, res = sys_.solve(L["rhs"], x0, max_iter, 1e-6, 0.0, raise_on_error=False)
    assert res.status == 5 and np.array_equal(x, x0)
    assert hbj._lib.last_error() == f"PCG(BlockJacobiGpu): did not converge in {max_iter} iterations"


def test_invalid_inputs_rejected_on_host(ctx, hbj):
    rp = np.array([0, 1, 2], np.uint32)
    with pytest.raises(hbj.SolverError, match="block_starts must be strictly increasing"):
        hbj.System(ctx, rp, np.array([0, 1], np.uint32), np.ones(2, F), np.array([0, 1, 1, 2], np.uint32))
    with pytest.raises(hbj.SolverError, match="out-of-range column 5"):
        hbj.System(ctx, rp, np.array([0, 5], np.uint32), np.ones(2, F), np.array([0, 2], np.uint32))
    with pytest.raises(hbj.SolverError, match="LU zero pivot at k = 0"):
        hbj.System(ctx, rp, np.array([0, 1], np.uint32), np.array([0, 1], F), np.array([0, 2], np.uint32))


@pytest.mark.slow
def test_full_size_lattice_50_bitwise(ctx, hbj, oracle):
    """The 50^3 configuration (750,000 unknowns, 30,352,200 nnz) against the
    oracle on the host cores: same iteration count and bit-identical x."""
    print("lattice N=50 full size", flush=True)
    L = hbj.lattice(50, seed=0)
    sys_ = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    x0 = np.zeros(L["n"], F)
    x, res = sys_.solve(L["rhs"], x0, 5000, 1e-6, 0.0)
    print(f"gpu: {res.iterations} its, {res.solve_ms:.1f} ms, true res {res.rel_res_true_fp64:.3e}",
          flush=True)
    st, it, xo, _ = oracle_solve(oracle, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"],
                                 L["rhs"], x0, 5000, 1e-6, 0.0)
    assert res.status == st == 0 and res.iterations == it
    assert np.array_equal(bits(x), bits(xo))
    # variant 5 (the sliced layout read through the caches: what the autotuner
    # picks for this system) on the same oracle x
    sys_.set_spmv_variant(5)
    x5, res5 = sys_.solve(L["rhs"], x0, 5000, 1e-6, 0.0)
    print(f"variant 5: {res5.iterations} its, {res5.solve_ms:.1f} ms", flush=True)
    assert res5.iterations == it and np.array_equal(bits(x5), bits(xo))
    sys_.close()


@pytest.mark.gpu
def test_spmv_variants_bitwise_and_autotune(ctx, hbj, oracle):
    """Every SpMV streaming variant gives the same x bits and iteration count
    (they differ only in how entries are staged); the autotuner picks one of
    them and the solve after it is unchanged."""
    L = hbj.lattice(10, seed=4)
    n = L["n"]
    st, it, xo, _ = oracle_solve(oracle, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"],
                                 L["rhs"], np.zeros(n, F), 2000, 1e-6, 0.0)
    sys_ = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    sys_.set_rhs(L["rhs"])
    info = sys_.info()
    assert info["n"] == n and info["nnz"] == L["nnz"] and info["local_tiles"] == info["n_tiles"]
    assert info["spmv_csr_bytes"] == 8 * L["nnz"] + 12 * n + 4
    assert 0 < info["spmv_layout_bytes"] < info["spmv_csr_bytes"]
    for v in range(hbj.SPMV_VARIANTS):
        sys_.set_spmv_variant(v)
        assert sys_.info()["spmv_variant"] == v
        res = sys_.solve_resident(2000, 1e-6, 0.0)
        assert res.iterations == it and np.array_equal(bits(sys_.get_x()), bits(xo)), v
    best, us = sys_.autotune(20)
    assert 0 <= best < hbj.SPMV_VARIANTS and np.all(us > 0) and us[best] == us.min()
    res = sys_.solve_resident(2000, 1e-6, 0.0)
    assert res.iterations == it and np.array_equal(bits(sys_.get_x()), bits(xo))
    with pytest.raises(hbj.SolverError):
        sys_.set_spmv_variant(hbj.SPMV_VARIANTS)
    sys_.close()


def _csr_from_dense(A):
    """CSR of A's nonzeros plus its whole diagonal (explicit zeros kept)."""
    rp, ci, va = [0], [], []
    for i in range(A.shape[0]):
        keep = A[i] != 0
        keep[i] = True
        nz = np.nonzero(keep)[0]
        ci.extend(nz.tolist())
        va.extend(A[i, nz].tolist())
        rp.append(len(ci))
    return np.array(rp, np.uint32), np.array(ci, np.uint32), np.array(va, F)


def test_device_lu_blocks_bitwise_vs_host_builder(ctx, hbj, oracle):
    """The LU blocks a system builds on the device from its CSR are the host
    builder's (build_lu_blocks_from_csr_block_starts_6 restated) bit for bit:
    the reference's datasets (including 3-row blocks), lattices, and a matrix
    with an oversize block and duplicate CSR entries (last write wins)."""
    cases = [oracle.load_dataset(k) for k in (1, 2, 3)]
    for N in (10, 50):
        L = hbj.lattice(N, seed=3)
        cases.append({"n": L["n"], "row_ptr": L["row_ptr"], "col_idx": L["col_idx"],
                      "values": L["values"], "block_starts": L["block_starts"]})
    rng = np.random.default_rng(5)
    A = (rng.standard_normal((14, 14)) * 0.1).astype(F) + 4 * np.eye(14, dtype=F)
    rp, ci, va = _csr_from_dense(A)
    # duplicate (row 2, col 3) entry after the original: the later value wins
    k = int(rp[2]) + int(np.searchsorted(ci[rp[2]:rp[3]], 3)) + 1
    ci = np.insert(ci, k, 3).astype(np.uint32)
    va = np.insert(va, k, np.float32(7.25)).astype(F)
    rp = rp.copy()
    rp[3:] += 1
    cases.append({"n": 14, "row_ptr": rp, "col_idx": ci, "values": va,
                  "block_starts": np.array([0, 8, 11, 14], np.uint32)})
    for d in cases:
        n = int(d["n"])
        st, ref = oracle.build_lu_blocks_6(n, d["row_ptr"], d["col_idx"], d["values"], d["block_starts"])
        assert st == 0
        s = hbj.System(ctx, d["row_ptr"], d["col_idx"], d["values"], d["block_starts"])
        got = s.lu()
        assert got.shape[0] == len(d["block_starts"]) - 1
        assert np.array_equal(got.reshape(-1).view(np.uint32), np.asarray(ref, F).reshape(-1).view(np.uint32))


def test_device_lu_reports_first_failing_block(ctx, hbj):
    """Two blocks with zero pivots (block 0 at k = 3, block 1 at k = 0): the
    error names the first block's pivot, as the host loop does."""
    A = np.eye(12, dtype=F)
    A[3, 3] = 0.0
    A[6, 6] = 0.0
    rp, ci, va = _csr_from_dense(A)
    with pytest.raises(hbj.SolverError, match="LU zero pivot at k = 3") as e:
        hbj.System(ctx, rp, ci, va, np.array([0, 6, 12], np.uint32))
    assert e.value.status_name == "ZERO_PIVOT"


def _arrow_spd(n=3000, dense=(5, 1700, 2999), seed=3):
    """SPD CSR with a few dense rows/columns (rows of ~n entries: several SpMV
    passes per row, and tiles referencing more than 2,048 columns, which
    gather from HBM in every variant) around a random near-band part;
    diagonally dominant.  Blocks of 1 to 6 rows."""
    rng = np.random.default_rng(seed)
    rows = [dict() for _ in range(n)]
    for i in range(n):
        for k in rng.integers(-30, 31, 6):
            j = i + int(k)
            if 0 <= j < n and j != i:
                v = float(rng.uniform(-1.0, 1.0))
                rows[i][j] = v
                rows[j][i] = v
    for dr in dense:
        for j in range(n):
            if j != dr:
                v = float(rng.uniform(-0.01, 0.01))
                rows[dr][j] = v
                rows[j][dr] = v
    for i in range(n):
        rows[i][i] = sum(abs(v) for j, v in rows[i].items() if j != i) + 1.0
    rp, ci, va = [0], [], []
    for i in range(n):
        for j in sorted(rows[i]):
            ci.append(j)
            va.append(rows[i][j])
        rp.append(len(ci))
    sizes, starts, s = [6, 6, 3, 1, 4, 6, 2, 5, 6, 6], [0], 0
    while s < n:
        s = min(n, s + sizes[len(starts) % len(sizes)])
        starts.append(s)
    b = rng.uniform(-1.0, 1.0, n).astype(F)
    return (np.array(rp, np.uint32), np.array(ci, np.uint32), np.array(va, F),
            np.array(starts, np.uint32), b)


def test_long_rows_and_tiles_without_column_lists_bitwise(ctx, hbj, oracle):
    print("arrow SPD n=3000: long rows, HBM-gathering tiles, irregular blocks", flush=True)
    rp, ci, va, bs, b = _arrow_spd()
    n = len(rp) - 1
    assert np.diff(rp.astype(np.int64)).max() > 2048  # rows longer than one pass of any variant
    off, cnt, _, _ = hbj.spmv_tile_tables(rp, ci)
    assert 0 < int(np.sum(cnt == 0xFFFFFFFF)) < cnt.size  # both kinds of tiles
    x0 = np.zeros(n, F)
    st, it, xo, _ = oracle_solve(oracle, rp, ci, va, bs, b, x0, 2000, 1e-6, 0.0)
    assert st == 0 and it > 5
    st_lu, lu = oracle.build_lu_blocks_6(n, rp, ci, va, bs)
    st_c, it_c, xc = oracle.ccg(rp, ci, va, lu, bs, b, x0, 2000, 1e-6, 0.0)
    assert st_c == 0
    sys_ = hbj.System(ctx, rp, ci, va, bs)
    for alg, its, xref in ((hbj.ALG_PCG, it, xo), (hbj.ALG_PCG_ONE_REDUCTION, it_c, xc)):
        sys_.set_algorithm(alg)
        for v in range(hbj.SPMV_VARIANTS):
            print(f"  algorithm {alg}, SpMV variant {v}", flush=True)
            sys_.set_spmv_variant(v)
            x, res = sys_.solve(b, x0, 2000, 1e-6, 0.0)
            assert res.status == 0 and res.iterations == its, (alg, v)
            assert np.array_equal(bits(x), bits(xref)), (alg, v)
    sys_.close()


def test_one_tile_update_kernel_bitwise(ctx, hbj, oracle, monkeypatch):
    """HBJ_UPDATE_GROUPS=0 selects the update kernel with one tile per
    workgroup (the default covers two, or four in a table): same bits."""
    monkeypatch.setenv("HBJ_UPDATE_GROUPS", "0")
    L = hbj.lattice(9, seed=6)
    x0 = np.zeros(L["n"], F)
    st, it, xo, _ = oracle_solve(oracle, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"],
                                 L["rhs"], x0, 5000, 1e-6, 0.0)
    sys_ = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    x, res = sys_.solve(L["rhs"], x0, 5000, 1e-6, 0.0)
    assert res.iterations == it and np.array_equal(bits(x), bits(xo))
    sys_.close()


def test_final_residual_matches_the_history_fixture(ctx, hbj, oracle):
    """The GPU's iteration count and final r.r equal the committed oracle
    histories (tests/golden/histories.npz) without running the oracle."""
    import os
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "histories.npz"))
    cases = []
    for k in (1, 2, 3):
        d = oracle.load_dataset(k)
        cases.append((f"ds{k}", d["row_ptr"], d["col_idx"], d["values"], d["block_starts"], d["b"],
                      d["x0"], 2000, 1e-4, 1e-7))
    L = hbj.lattice(10, seed=0)
    cases.append(("lattice10", L["row_ptr"], L["col_idx"], L["values"], L["block_starts"], L["rhs"],
                  np.zeros(L["n"], F), 5000, 1e-6, 0.0))
    for name, rp, ci, va, bs, b, x0, mi, rel, ab in cases:
        print(f"history fixture {name}", flush=True)
        sys_ = hbj.System(ctx, rp, ci, va, bs)
        x, res = sys_.solve(b, x0, mi, rel, ab)
        st, it = fx[f"{name}_status_iterations"].tolist()
        assert res.status == st and res.iterations == it, name
        assert np.float32(res.rr).view(np.uint32) == fx[f"{name}_rr"][-1].view(np.uint32), name
        sys_.close()


def _banded_spd(n, seed):
    """Symmetric, diagonally dominant, entries within 40 of the diagonal."""
    rng = np.random.default_rng(seed)
    rows = [dict() for _ in range(n)]
    for i in range(n):
        for k in rng.integers(1, 41, 4):
            j = i + int(k)
            if j < n:
                v = float(rng.uniform(-1.0, 1.0))
                rows[i][j] = v
                rows[j][i] = v
    for i in range(n):
        rows[i][i] = sum(abs(v) for v in rows[i].values()) + 1.0
    rp, ci, va = [0], [], []
    for i in range(n):
        for j in sorted(rows[i]):
            ci.append(j)
            va.append(rows[i][j])
        rp.append(len(ci))
    bs = np.array(list(range(0, n, 6)) + [n], np.uint32)  # 6-row blocks, the last one partial
    return (np.array(rp, np.uint32), np.array(ci, np.uint32), np.array(va, F), bs,
            rng.uniform(-1.0, 1.0, n).astype(F))


@pytest.mark.parametrize("n", [1, 5, 7, 255, 256, 257, 511, 513, 1000])
def test_tile_and_block_edges_bitwise(ctx, hbj, oracle, n):
    """Systems ending inside a tile and inside a 6-row block (the arithmetic
    block path with a partial last block), both algorithms, every variant."""
    print(f"edges n={n}", flush=True)
    rp, ci, va, bs, b = _banded_spd(n, seed=n)
    x0 = np.zeros(n, F)
    st, it, xo, _ = oracle_solve(oracle, rp, ci, va, bs, b, x0, 500, 1e-6, 0.0)
    st_lu, lu = oracle.build_lu_blocks_6(n, rp, ci, va, bs)
    st_c, it_c, xc = oracle.ccg(rp, ci, va, lu, bs, b, x0, 500, 1e-6, 0.0)
    sys_ = hbj.System(ctx, rp, ci, va, bs)
    for alg, sts, its, xref in ((hbj.ALG_PCG, st, it, xo), (hbj.ALG_PCG_ONE_REDUCTION, st_c, it_c, xc)):
        sys_.set_algorithm(alg)
        for v in range(hbj.SPMV_VARIANTS):
            sys_.set_spmv_variant(v)
            x, res = sys_.solve(b, x0, 500, 1e-6, 0.0, raise_on_error=False)
            assert res.status == sts and res.iterations == its, (alg, v)
            if sts == 0:
                assert np.array_equal(bits(x), bits(xref)), (alg, v)
    sys_.close()


def test_small_blocks_stage_factors_in_groups(ctx, hbj, oracle):
    """Tiles meeting more blocks than one LDS staging group holds (48): 300
    one-row blocks (tile 0 meets 256 blocks, six groups), then sizes 1-6 at
    random (the full tiles after it meet 72-105 blocks).  Both algorithms, every variant, and
    the system inside a batch table beside a lattice: bit-identical to the
    oracle."""
    n = 1400
    rp, ci, va, _, b = _banded_spd(n, seed=77)
    rng = np.random.default_rng(78)
    starts = list(range(0, 300))
    while starts[-1] < n:
        starts.append(min(n, starts[-1] + int(rng.integers(1, 7))))
    bs = np.array(starts, np.uint32)
    per_tile = [int(np.searchsorted(bs, min(n, 256 * (k + 1)), "left") - np.searchsorted(bs, 256 * k, "right") + 1)
                for k in range((n + 255) // 256)]
    print(f"small blocks: n={n}, {len(bs) - 1} blocks, blocks per tile {per_tile}", flush=True)
    assert per_tile[0] > 5 * 48 and min(per_tile[1:-1]) > 48
    x0 = np.zeros(n, F)
    st, it, xo, _ = oracle_solve(oracle, rp, ci, va, bs, b, x0, 500, 1e-6, 0.0)
    st_lu, lu = oracle.build_lu_blocks_6(n, rp, ci, va, bs)
    st_c, it_c, xc = oracle.ccg(rp, ci, va, lu, bs, b, x0, 500, 1e-6, 0.0)
    assert st == st_c == 0
    sys_ = hbj.System(ctx, rp, ci, va, bs)
    for alg, its, xref in ((hbj.ALG_PCG, it, xo), (hbj.ALG_PCG_ONE_REDUCTION, it_c, xc)):
        sys_.set_algorithm(alg)
        for v in range(hbj.SPMV_VARIANTS):
            sys_.set_spmv_variant(v)
            x, res = sys_.solve(b, x0, 500, 1e-6, 0.0)
            assert res.iterations == its and np.array_equal(bits(x), bits(xref)), (alg, v)
    sys_.set_algorithm(hbj.ALG_PCG)
    sys_.set_rhs(b)
    L = hbj.lattice(5, seed=5)
    other = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    xl, rl = other.solve(L["rhs"], np.zeros(L["n"], F), 5000, 1e-6, 0.0)
    other.set_rhs(L["rhs"])
    batch = hbj.Batch(ctx, [sys_, other])
    stb, results = batch.solve(500, 1e-6, 0.0)
    assert stb == 0 and results[0].iterations == it and results[1].iterations == rl.iterations
    assert np.array_equal(bits(sys_.get_x()), bits(xo))
    assert np.array_equal(bits(other.get_x()), bits(xl))
    batch.close()
    other.close()
    sys_.close()


def _insert_zero(rp, ci, va, r, c):
    """CSR with an explicit zero entry added at (r, c) (columns kept sorted)."""
    e0, e1 = int(rp[r]), int(rp[r + 1])
    pos = e0 + int(np.searchsorted(ci[e0:e1], c))
    assert pos == e1 or ci[pos] != c
    rp2 = rp.copy()
    rp2[r + 1:] += 1
    return rp2, np.insert(ci, pos, np.uint32(c)), np.insert(va, pos, F(0.0))


def test_sliced_layout_mixed_groups_bitwise(ctx, hbj, oracle):
    """Sliced layout with both storage forms in one system: a lattice's node rows
    share their runs and its tiles' column lists are whole groups of 6 columns,
    stored once per group; an explicit zero at (3000, 0) gives one row runs of
    its own and its tile a list that is not whole groups, one at (6000, 6012)
    changes a row's runs only.  A is unchanged, so the oracle's x is the
    lattice's; every variant and both algorithms stay bit-identical to it."""
    print("lattice N=12 with two explicit zeros: mixed sliced storage", flush=True)
    L = hbj.lattice(12, seed=3)
    rp, ci, va = L["row_ptr"], L["col_idx"], L["values"]
    for r, c in ((3000, 0), (6000, 6012)):
        rp, ci, va = _insert_zero(rp, ci, va, r, c)
    bs, b = L["block_starts"], L["rhs"]
    n = L["n"]
    x0 = np.zeros(n, F)
    base = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], bs)
    base_bytes = base.info()["spmv_sliced_bytes"]
    base.close()
    sys_ = hbj.System(ctx, rp, ci, va, bs)
    mixed_bytes = sys_.info()["spmv_sliced_bytes"]
    assert 0 < base_bytes < mixed_bytes  # both systems sliced; the mixed one partly per row
    st, it, xo, _ = oracle_solve(oracle, rp, ci, va, bs, b, x0, 5000, 1e-6, 0.0)
    st_lu, lu = oracle.build_lu_blocks_6(n, rp, ci, va, bs)
    st_c, it_c, xc = oracle.ccg(rp, ci, va, lu, bs, b, x0, 5000, 1e-6, 0.0)
    assert st == 0 and st_c == 0
    for alg, its, xref in ((hbj.ALG_PCG, it, xo), (hbj.ALG_PCG_ONE_REDUCTION, it_c, xc)):
        sys_.set_algorithm(alg)
        for v in range(hbj.SPMV_VARIANTS):
            print(f"  algorithm {alg}, SpMV variant {v}", flush=True)
            sys_.set_spmv_variant(v)
            x, res = sys_.solve(b, x0, 5000, 1e-6, 0.0)
            assert res.status == 0 and res.iterations == its, (alg, v)
            assert np.array_equal(bits(x), bits(xref)), (alg, v)
    sys_.close()


def test_three_launch_form_above_prefetch_limit(ctx, hbj, oracle, monkeypatch):
    """A system of more than 4,096 tiles runs the three-launch iteration (its
    r.z partials exceed what a workgroup prefetches) with the first-level
    sums of its partials computed once per launch (k_level1); a smaller one
    the two-launch iteration.  Both give the oracle's bits, and the
    refinement of the large one (FP64 partials) gives the bits of the same
    solve with every workgroup reducing every partial (HBJ_L1=0).  The large
    system is a chain of 6x6 node blocks, diagonally dominant, so the oracle's
    solve is short: 183,000 nodes, 1,098,000 rows (4,290 tiles)."""
    nodes = 183_000
    n = 6 * nodes
    rng = np.random.default_rng(21)
    # per node: diagonal block (symmetric, dominant) and the coupling block to
    # the next node (its transpose couples back)
    D = rng.uniform(-0.1, 0.1, (nodes, 6, 6)).astype(np.float32)
    D = (D + D.transpose(0, 2, 1)) * np.float32(0.5) + np.eye(6, dtype=np.float32) * np.float32(4.0)
    C = rng.uniform(-0.2, 0.2, (nodes - 1, 6, 6)).astype(np.float32)
    rows = []
    # CSR row by row: [C_{g-1}^T (row d)] [D_g (row d)] [C_g (row d)], columns ascending
    counts = np.full(nodes, 18, np.int64)
    counts[0] -= 6
    counts[-1] -= 6
    row_len = np.repeat(counts, 6)
    row_ptr = np.zeros(n + 1, np.uint32)
    row_ptr[1:] = np.cumsum(row_len)
    nnz = int(row_ptr[-1])
    col = np.empty(nnz, np.uint32)
    val = np.empty(nnz, np.float32)
    g = np.arange(nodes)
    for d in range(6):
        r = 6 * g + d
        pos = row_ptr[r].astype(np.int64)
        has_prev = g > 0
        k = pos.copy()
        for j in range(6):  # lower block: (C_{g-1}^T)[d, j] = C_{g-1}[j, d]
            sel = has_prev
            col[k[sel]] = 6 * (g[sel] - 1) + j
            val[k[sel]] = C[g[sel] - 1, j, d]
            k[sel] += 1
        for j in range(6):
            col[k] = 6 * g + j
            val[k] = D[g, d, j]
            k += 1
        has_next = g < nodes - 1
        for j in range(6):
            sel = has_next
            col[k[sel]] = 6 * (g[sel] + 1) + j
            val[k[sel]] = C[g[sel], d, j]
            k[sel] += 1
    bs = np.arange(0, n + 1, 6, dtype=np.uint32)
    b = rng.uniform(-1, 1, n).astype(F)
    x0 = np.zeros(n, F)
    s = hbj.System(ctx, row_ptr, col, val, bs)
    info = s.info()
    assert info["n_tiles"] > 4096 and info["launches_per_iteration"] == 3
    x, res = s.solve(b, x0, 500, 1e-6, 0.0)
    st, it, xo, _ = oracle_solve(oracle, row_ptr, col, val, bs, b, x0, 500, 1e-6, 0.0)
    print(f"chain n={n}: {res.iterations} iterations", flush=True)
    assert res.status == st == 0 and res.iterations == it
    assert np.array_equal(bits(x), bits(xo))
    b64 = b.astype(np.float64)
    x1, r1 = s.solve_ir(b64)
    s.close()
    monkeypatch.setenv("HBJ_L1", "0")
    s3 = hbj.System(ctx, row_ptr, col, val, bs)
    x3, r3 = s3.solve_ir(b64)
    print(f"chain refinement: {r1.iterations} inner iterations, {r1.outer_iterations} outer", flush=True)
    assert r1.status == r3.status == 0
    assert (r1.iterations, r1.outer_iterations) == (r3.iterations, r3.outer_iterations)
    assert np.array_equal(x1.view(np.uint64), x3.view(np.uint64))
    s3.close()
    small = hbj.lattice(8, seed=1)
    s2 = hbj.System(ctx, small["row_ptr"], small["col_idx"], small["values"], small["block_starts"])
    assert s2.info()["launches_per_iteration"] == 2
    s2.close()


@pytest.mark.parametrize("fused", ["1", "0"])
def test_repeated_solves_multi_level_reductions(ctx, hbj, oracle, monkeypatch, fused):
    """A system whose dot partials take two levels of the reduction tree
    (24^3 lattice: 324 tiles): repeated solves on the same system, under
    several SpMV variants, in the two-launch (HBJ_FUSED=1) and the
    three-launch (HBJ_FUSED=0) iteration, give the oracle's bits every time
    (nothing a solve leaves behind -- state copies, partials, p of either
    parity -- reaches the next one); so do two refinement solves in a row."""
    monkeypatch.setenv("HBJ_FUSED", fused)
    print(f"24^3 lattice, repeated solves, HBJ_FUSED={fused}", flush=True)
    L = hbj.lattice(24, seed=2)
    n = L["n"]
    x0 = np.zeros(n, F)
    st, it, xo, _ = oracle_solve(oracle, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"],
                                 L["rhs"], x0, 5000, 1e-6, 0.0)
    assert st == 0
    s = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    assert s.info()["n_tiles"] > 256
    assert s.info()["launches_per_iteration"] == (2 if fused == "1" else 3)
    for v in (5, 5, 2, 4):
        s.set_spmv_variant(v)
        x, res = s.solve(L["rhs"], x0, 5000, 1e-6, 0.0)
        assert res.status == 0 and res.iterations == it, v
        assert np.array_equal(bits(x), bits(xo)), v
    p = hbj.IrParams(max_outer=30, max_inner=20000, inner_rel_tol=1e-4, target_rel_res=1e-9)
    x1, r1 = s.solve_ir(L["rhs"].astype(np.float64), None, p)
    x2, r2 = s.solve_ir(L["rhs"].astype(np.float64), None, p)
    assert r1.status == r2.status == 0
    assert (r1.iterations, r1.outer_iterations) == (r2.iterations, r2.outer_iterations)
    assert np.array_equal(x1.view(np.uint64), x2.view(np.uint64))
    x, res = s.solve(L["rhs"], x0, 5000, 1e-6, 0.0)
    assert res.iterations == it and np.array_equal(bits(x), bits(xo))
    s.close()
