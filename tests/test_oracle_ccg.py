"""Pins the oracle's one-reduction (Chronopoulos-Gear) PCG, or_ccg -- the CPU
restatement of the HIP library's optional algorithm, which is not a reference
function.  Checked against an independent numpy FP64 restatement of the same
recurrences (which must take the textbook FP64 PCG's iteration count), against
or_pcg on the beam lattices (same iterates in exact arithmetic), the reference
datasets' x_ref with the compare-x rule, and the error paths."""
import numpy as np
import pytest
import scipy.sparse as sp

F = np.float32
DATASETS = [(1, 1e-2), (2, 1e0), (3, 1e1)]


def _lu(oracle, n, rp, ci, va, bs):
    st, lu = oracle.build_lu_blocks_6(n, rp, ci, va, bs)
    assert st == 0
    return lu


def _csr(rows):
    rp, ci, va = [0], [], []
    for r in rows:
        for c, v in r:
            ci.append(c)
            va.append(v)
        rp.append(len(ci))
    return np.array(rp, np.uint32), np.array(ci, np.uint32), np.array(va, F)


def _spd(n=600, seed=9):
    rng = np.random.default_rng(seed)
    B = sp.random(n, n, density=0.01, random_state=seed + 1)
    A = (B @ B.T + sp.eye(n) * 2.0).tocsr().astype(F)
    A.sort_indices()
    return A, np.arange(0, n + 1, 6, dtype=np.uint32), rng.standard_normal(n).astype(F)


def _fp64_counts(A, b, tol2, maxit=2000):
    """textbook FP64 block-Jacobi PCG and FP64 Chronopoulos-Gear: iteration counts"""
    n = A.shape[0]
    A64 = A.astype(np.float64)
    Minv = sp.block_diag([np.linalg.inv(A64[s:s + 6, s:s + 6].toarray()) for s in range(0, n, 6)])
    b64 = b.astype(np.float64)
    bb = b64 @ b64
    # textbook
    r = b64.copy()
    z = Minv @ r
    p = z.copy()
    rz = r @ z
    it_pcg = None
    for k in range(maxit):
        q = A64 @ p
        a = rz / (p @ q)
        r = r - a * q
        if r @ r <= tol2 * bb:
            it_pcg = k + 1
            break
        z = Minv @ r
        rz, rz_old = r @ z, rz
        p = z + (rz / rz_old) * p
    # Chronopoulos-Gear
    r = b64.copy()
    u = Minv @ r
    p = np.zeros(n)
    s = np.zeros(n)
    g_prev = a_prev = None
    it_ccg = None
    for k in range(maxit + 1):
        w = A64 @ u
        g, dl = r @ u, w @ u
        if k > 0 and r @ r <= tol2 * bb:
            it_ccg = k
            break
        beta = 0.0 if k == 0 else g / g_prev
        a = g / dl if k == 0 else g / (dl - beta * g / a_prev)
        p = u + beta * p
        s = w + beta * s
        r = r - a * s
        u = Minv @ r
        g_prev, a_prev = g, a
    return it_pcg, it_ccg


def test_ccg_recurrences_take_the_pcg_iteration_count(oracle):
    A, bs, b = _spd()
    it_pcg, it_ccg = _fp64_counts(A, b, 1e-12)
    assert it_pcg is not None and it_ccg == it_pcg
    n = A.shape[0]
    rp, ci = A.indptr.astype(np.uint32), A.indices.astype(np.uint32)
    lu = _lu(oracle, n, rp, ci, A.data, bs)
    st, it, x = oracle.ccg(rp, ci, A.data, lu, bs, b, np.zeros(n, F), 500, 1e-6, 0.0)
    assert st == 0 and abs(it - it_pcg) <= 1
    xd = np.linalg.solve(A.astype(np.float64).toarray(), b.astype(np.float64))
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-5


@pytest.mark.parametrize("N,h", [(6, 0.0), (8, 1.5)])
def test_ccg_matches_pcg_on_lattices(oracle, hbj, N, h):
    L = hbj.lattice(N, seed=5, e_log10_half_range=h)
    n = L["n"]
    lu = _lu(oracle, n, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])
    args = (L["row_ptr"], L["col_idx"], L["values"], lu, L["block_starts"], L["rhs"], np.zeros(n, F),
            5000, 1e-6, 0.0)
    st_p, it_p, x_p, _ = oracle.pcg(*args)
    st_c, it_c, x_c = oracle.ccg(*args)
    assert st_p == st_c == 0
    assert abs(it_c - it_p) <= max(2, it_p // 100)
    b64 = L["rhs"].astype(np.float64)
    tr_p = oracle.true_rel_residual64(L["row_ptr"], L["col_idx"], L["values"], b64, x_p.astype(np.float64))
    tr_c = oracle.true_rel_residual64(L["row_ptr"], L["col_idx"], L["values"], b64, x_c.astype(np.float64))
    assert tr_c <= 3.0 * tr_p + 1e-6


@pytest.mark.parametrize("k,atol", DATASETS)
def test_ccg_datasets_meet_x_ref(oracle, k, atol):
    d = oracle.load_dataset(k)
    n = int(d["n"])
    lu = _lu(oracle, n, d["row_ptr"], d["col_idx"], d["values"], d["block_starts"])
    st, it, x = oracle.ccg(d["row_ptr"], d["col_idx"], d["values"], lu, d["block_starts"], d["b"],
                           d["x0"], 2000, 1e-4, 1e-7)
    assert st == 0 and 0 < it < 2000
    xr = d["x_ref"].astype(np.float64)
    err = x.astype(np.float64) - xr
    rel_l2 = np.linalg.norm(err) / np.linalg.norm(xr)
    assert rel_l2 <= 1e-3 or np.sqrt(np.mean(err ** 2)) <= atol  # compare-x PASS rule


def test_ccg_error_paths(oracle):
    d = oracle.load_dataset(1)
    n = int(d["n"])
    lu = _lu(oracle, n, d["row_ptr"], d["col_idx"], d["values"], d["block_starts"])
    x0 = np.arange(n, dtype=F)
    st, it, x = oracle.ccg(d["row_ptr"], d["col_idx"], d["values"], lu, d["block_starts"],
                           np.zeros(n, F), x0, 100, 1e-4, 0.0)
    assert st == 0 and it == 0 and np.array_equal(x, x0)  # Ok(0), x untouched
    for mi in (0, 1, 5):
        st, it, x = oracle.ccg(d["row_ptr"], d["col_idx"], d["values"], lu, d["block_starts"],
                               d["b"], x0, mi, 1e-4, 1e-7)
        assert st == 5 and it == mi and np.array_equal(x, x0)
        assert oracle.last_error() == f"PCG(BlockJacobiGpu): did not converge in {mi} iterations"
    # identity A with x0 = b: r0 = u0 = w0 = 0, so (p, A p) = (w, u) = 0
    m = 12
    rp, ci, va = _csr([[(i, 1.0)] for i in range(m)])
    bs = np.arange(0, m + 1, 6, dtype=np.uint32)
    lu = _lu(oracle, m, rp, ci, va, bs)
    b = np.arange(1, m + 1, dtype=F)
    st, it, x = oracle.ccg(rp, ci, va, lu, bs, b, b.copy(), 100, 1e-6, 0.0)
    assert st == 3 and it == 0
    assert oracle.last_error() == "PCG(BlockJacobiGpu): dot(p,Ap) is zero (breakdown)"
    # M = diag(1,-1,1), b = (1,1,0): (r,u) = 0 with (w,u) = -2a != 0, so alpha = 0
    # and the next iteration finds gamma_prev = 0
    a = 0.5
    rp, ci, va = _csr([[(0, 1.0), (1, a)], [(0, a), (1, -1.0)], [(2, 1.0)]])
    bs = np.array([0, 1, 2, 3], np.uint32)
    lu = _lu(oracle, 3, rp, ci, va, bs)
    st, it, x = oracle.ccg(rp, ci, va, lu, bs, np.array([1, 1, 0], F), np.zeros(3, F), 10, 1e-6, 0.0)
    assert st == 4 and it == 1 and np.array_equal(x, np.zeros(3, F))
    assert oracle.last_error() == "PCG(BlockJacobiGpu): rz_old is zero (breakdown)"
