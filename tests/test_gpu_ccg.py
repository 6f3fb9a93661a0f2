"""The optional one-reduction PCG (hbj_system_set_algorithm) on the GPU against
the oracle's restatement of it (or_ccg): same status and iteration count,
bit-identical x; every SpMV variant alike; error paths; and the drop-in
algorithm unchanged after switching back."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F = np.float32


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _oracle_ccg(oracle, rp, ci, va, bs, b, x0, max_iter, rel, ab):
    st, lu = oracle.build_lu_blocks_6(len(rp) - 1, rp, ci, va, bs)
    assert st == 0
    return oracle.ccg(rp, ci, va, lu, bs, b, x0, max_iter, rel, ab)


def _ccg_system(ctx, hbj, rp, ci, va, bs):
    s = hbj.System(ctx, rp, ci, va, bs)
    s.set_algorithm(hbj.ALG_PCG_ONE_REDUCTION)
    return s


@pytest.mark.parametrize("k", [1, 2, 3])
def test_ccg_datasets_bitwise_vs_oracle(ctx, hbj, oracle, k):
    print(f"one-reduction PCG, dataset {k}", flush=True)
    d = oracle.load_dataset(k)
    s = _ccg_system(ctx, hbj, d["row_ptr"], d["col_idx"], d["values"], d["block_starts"])
    x, res = s.solve(d["b"], d["x0"], 2000, 1e-4, 1e-7)
    st, it, xo = _oracle_ccg(oracle, d["row_ptr"], d["col_idx"], d["values"], d["block_starts"],
                             d["b"], d["x0"], 2000, 1e-4, 1e-7)
    assert res.status == st == 0 and res.iterations == it
    assert np.array_equal(bits(x), bits(xo))
    assert res.rel_res_true_fp64 > 0
    s.close()


@pytest.mark.parametrize("N,h,seed", [(10, 0.0, 0), (6, 1.5, 4), (13, 0.0, 2)])
def test_ccg_lattice_bitwise_vs_oracle(ctx, hbj, oracle, N, h, seed):
    print(f"one-reduction PCG, lattice N={N} h={h}", flush=True)
    L = hbj.lattice(N, seed=seed, e_log10_half_range=h)
    s = _ccg_system(ctx, hbj, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    x0 = np.zeros(L["n"], F)
    x, res = s.solve(L["rhs"], x0, 5000, 1e-6, 0.0)
    st, it, xo = _oracle_ccg(oracle, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"],
                             L["rhs"], x0, 5000, 1e-6, 0.0)
    assert res.status == st == 0 and res.iterations == it
    assert np.array_equal(bits(x), bits(xo))
    x2, res2 = s.solve(L["rhs"], x0, 5000, 1e-6, 0.0)  # a second solve: no state leaks
    assert res2.iterations == it and np.array_equal(bits(x2), bits(xo))
    s.close()


def test_ccg_spmv_variants_and_kernel_timing(ctx, hbj, oracle):
    L = hbj.lattice(12, seed=7)
    s = _ccg_system(ctx, hbj, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    x0 = np.zeros(L["n"], F)
    st, it, xo = _oracle_ccg(oracle, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"],
                             L["rhs"], x0, 5000, 1e-6, 0.0)
    for v in range(hbj.SPMV_VARIANTS):
        print(f"one-reduction PCG, SpMV variant {v}", flush=True)
        s.set_spmv_variant(v)
        x, res = s.solve(L["rhs"], x0, 5000, 1e-6, 0.0)
        assert res.iterations == it and np.array_equal(bits(x), bits(xo))
    s.set_kernel_timing(4)
    x, res = s.solve(L["rhs"], x0, 5000, 1e-6, 0.0)
    avg, cnt = s.kernel_timing()
    assert cnt[0] > 0 and cnt[1] > 0 and avg[0] > 0 and avg[1] > 0  # SpMV, update
    assert cnt[2] == 0  # no third kernel
    s.close()


def test_ccg_error_paths(ctx, hbj, oracle):
    d = oracle.load_dataset(1)
    n = int(d["n"])
    s = _ccg_system(ctx, hbj, d["row_ptr"], d["col_idx"], d["values"], d["block_starts"])
    x0 = np.arange(n, dtype=F)
    x, res = s.solve(np.zeros(n, F), x0, 100, 1e-4, 0.0)
    assert res.status == 0 and res.iterations == 0 and np.array_equal(x, x0)
    for mi in (0, 5):
        x, res = s.solve(d["b"], x0, mi, 1e-4, 1e-7, raise_on_error=False)
        assert res.status == 5 and res.iterations == mi and np.array_equal(x, x0)
        assert hbj._lib.last_error() == f"PCG(BlockJacobiGpu): did not converge in {mi} iterations"
    s.close()
    m = 12
    rp = np.arange(m + 1, dtype=np.uint32)
    ident = _ccg_system(ctx, hbj, rp, np.arange(m, dtype=np.uint32), np.ones(m, F),
                        np.arange(0, m + 1, 6, dtype=np.uint32))
    b = np.arange(1, m + 1, dtype=F)
    x, res = ident.solve(b, b.copy(), 100, 1e-6, 0.0, raise_on_error=False)
    assert res.status == 3 and res.iterations == 0
    assert hbj._lib.last_error() == "PCG(BlockJacobiGpu): dot(p,Ap) is zero (breakdown)"
    ident.close()


def test_switching_back_restores_the_reference_algorithm(ctx, hbj, oracle):
    d = oracle.load_dataset(2)
    s = _ccg_system(ctx, hbj, d["row_ptr"], d["col_idx"], d["values"], d["block_starts"])
    s.solve(d["b"], d["x0"], 2000, 1e-4, 1e-7)
    s.set_algorithm(hbj.ALG_PCG)
    x, res = s.solve(d["b"], d["x0"], 2000, 1e-4, 1e-7)
    st, lu = oracle.build_lu_blocks_6(int(d["n"]), d["row_ptr"], d["col_idx"], d["values"],
                                      d["block_starts"])
    st, it, xo, _ = oracle.pcg(d["row_ptr"], d["col_idx"], d["values"], lu, d["block_starts"], d["b"],
                               d["x0"], 2000, 1e-4, 1e-7)
    assert res.iterations == it == 370 and np.array_equal(bits(x), bits(xo))
    with pytest.raises(hbj.SolverError):
        s.set_algorithm(2)
    s.close()


def test_ccg_full_size_lattice_50(ctx, hbj):
    print("one-reduction PCG, lattice N=50", flush=True)
    L = hbj.lattice(50, seed=0)
    s = hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    x0 = np.zeros(L["n"], F)
    xp, rp_ = s.solve(L["rhs"], x0, 5000, 1e-6, 0.0)
    s.set_algorithm(hbj.ALG_PCG_ONE_REDUCTION)
    xc, rc = s.solve(L["rhs"], x0, 5000, 1e-6, 0.0)
    assert rc.status == 0 and abs(rc.iterations - rp_.iterations) <= rp_.iterations // 100
    assert rc.rel_res_true_fp64 <= 3.0 * rp_.rel_res_true_fp64
    s.close()
