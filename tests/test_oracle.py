"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check compares the oracle with something other than its own formula:
the reference's known-answer self-tests (tests/golden/kats.json, from cli
main.rs:171-443), the reference's golden x_ref vectors (datasets 1-3 with the
commands.txt tolerances), hand-derived values that only the reference's
operation order produces, and numpy/scipy in FP64 within error bounds.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KATS = json.load(open(os.path.join(GOLDEN, "kats.json")))
F = np.float32
T24 = float(2 ** 24)


def csr(rows, n=None):
    """rows: list of [(col, val), ...] -> (row_ptr, col_idx, values)."""
    rp, ci, va = [0], [], []
    for r in rows:
        for c, v in r:
            ci.append(c)
            va.append(v)
        rp.append(len(ci))
    return np.array(rp, np.uint32), np.array(ci, np.uint32), np.array(va, np.float32)


# ------------------------------------------------------------------- KATs

def test_kat_vec(oracle):
    k = KATS["vec_test"]
    y = oracle.axpy(k["alpha"], np.full(k["n"], k["x"], F), np.full(k["n"], k["y"], F))
    assert np.all(np.abs(y - k["expected"]) < k["tol"])


def test_kat_dot(oracle):
    k = KATS["dot_test"]
    assert abs(oracle.dot(k["a"], k["b"]) - k["expected"]) < k["tol"]


def test_kat_spmv(oracle):
    k = KATS["spmv_test"]
    y = oracle.spmv(k["row_ptr"], k["col_idx"], k["values"], k["x"])
    assert np.all(np.abs(y - np.array(k["expected"], F)) < k["tol"])


def test_kat_block_jacobi(oracle):
    k = KATS["block_jacobi_test"]
    lu = np.zeros(2 * 36, F)
    for b in range(2):
        for i in range(6):
            lu[b * 36 + i * 6 + i] = 1.0
    z = oracle.block_jacobi(k["n"], lu, k["block_starts"], np.array(k["r"], F))
    assert np.all(np.abs(z - np.array(k["expected"], F)) < k["tol"])


def test_kat_update_scalars(oracle):
    k = KATS["pcg_update_scalars_test"]
    s = np.zeros(k["slots"], F)
    s[k["PAP"]], s[k["RZ_OLD"]], s[k["RZ_NEW"]] = k["pAp"], k["rz_old"], k["rz_new"]
    s = oracle.update_scalars(s, k["PAP"], k["RZ_NEW"], k["RZ_OLD"], k["ALPHA"], k["MINUS_ALPHA"],
                              k["BETA"])
    assert abs(s[k["ALPHA"]] - k["alpha"]) < k["tol"]
    assert abs(s[k["MINUS_ALPHA"]] - k["minus_alpha"]) < k["tol"]
    assert abs(s[k["BETA"]] - k["beta"]) < k["tol"]


def test_update_scalars_zero_guards(oracle):
    # pcg_update_scalars.wgsl:56-67: zero denominators give 0, not inf/nan
    s = oracle.update_scalars(np.array([0, 5, 10, 9, 9, 9], F), 0, 1, 2, 3, 4, 5)
    assert s[3] == 0 and s[4] == 0 and s[5] == 0.5
    s = oracle.update_scalars(np.array([2, 5, 0, 9, 9, 9], F), 0, 1, 2, 3, 4, 5)
    assert s[3] == 0 and s[5] == 0


# --------------------------------------------------------- dot tree order

def _chunk(vals_at, n=256):
    a = np.zeros(n, F)
    for i, v in vals_at.items():
        a[i] = v
    return a


def test_dot_stride_halving_tree(oracle):
    # a[0]=2^24, a[1]=a[3]=1 (b = 1): the stride-halving tree pairs (1,3) at
    # offset 2 -> 2, then 2^24 + 2 (exact).  A sequential sum gives 2^24
    # (each +1 rounds away), an adjacent-pair tree (0,1),(2,3) gives 2^24 too.
    a = _chunk({0: T24, 1: 1.0, 3: 1.0})
    assert oracle.dot(a, np.ones(256, F)) == T24 + 2
    seq = F(0)
    for v in a:
        seq = F(seq + v)
    assert seq == T24  # the alternative the oracle must not match


def test_dot_partial_level_tree(oracle):
    # the same pattern one level up: partials P0=2^24, P1=1, P3=1 over 4 chunks
    a = np.zeros(1024, F)
    a[0], a[256], a[768] = T24, 1.0, 1.0
    assert oracle.dot(a, np.ones(1024, F)) == T24 + 2


def test_dot_first_level_pairs_across_128(oracle):
    # offset 128 pairs t with t+128 first: a[0]=2^24, a[128]=1, a[1]=1:
    # (2^24+1)->2^24, then 2^24+1 -> 2^24 ; a sequential sum also gives 2^24,
    # but a[0]=2^24, a[1]=1, a[129]=1 pairs (1,129) -> 2 -> 2^24+2.
    assert oracle.dot(_chunk({0: T24, 1: 1.0, 129: 1.0}), np.ones(256, F)) == T24 + 2


def test_dot_empty_and_bound(oracle):
    assert oracle.dot(np.zeros(0, F), np.zeros(0, F)) == 0.0
    rng = np.random.default_rng(1)
    for n in (1, 255, 256, 257, 65536, 65537 * 3):
        a = rng.standard_normal(n).astype(F)
        b = rng.standard_normal(n).astype(F)
        exact = math.fsum((a.astype(np.float64) * b.astype(np.float64)).tolist())
        bound = 2 * (math.log2(n) + 3) * 2 ** -24 * float(np.sum(np.abs(a.astype(np.float64) * b)))
        assert abs(oracle.dot(a, b) - exact) <= bound + 1e-30


# ----------------------------------------------------------------- spmv

def test_spmv_sequential_row_order(oracle):
    # row [2^24, 1, 1] . [1,1,1]: (0 + 2^24) + 1 -> 2^24, + 1 -> 2^24 ;
    # any other order (reverse, pairwise) gives 2^24 + 2.
    rp, ci, va = csr([[(0, T24), (1, 1.0), (2, 1.0)], [(1, 1.0)], [(2, 1.0)]])
    y = oracle.spmv(rp, ci, va, np.ones(3, F))
    assert y[0] == T24


def test_spmv_no_contraction(oracle):
    # sum + v*x with sum = -(1+2^-11), v = x = 1+2^-12: fl(v*x) = 1+2^-11
    # (the 2^-24 tail rounds to even), so the two-rounding result is 0 while a
    # fused multiply-add would give 2^-24.
    e = 1.0 + 2.0 ** -12
    rp, ci, va = csr([[(0, -(1.0 + 2.0 ** -11)), (1, e)], [(1, 1.0)]])
    y = oracle.spmv(rp, ci, va, np.array([1.0, e], F))
    assert y[0] == 0.0


def test_spmv_matches_scipy(oracle):
    rng = np.random.default_rng(2)
    n = 3000
    A = sp.random(n, n, density=0.01, format="csr", random_state=3, dtype=np.float64)
    A.data = rng.integers(-8, 9, A.nnz).astype(np.float64)
    A.sort_indices()
    x = rng.integers(-8, 9, n).astype(F)
    y = oracle.spmv(A.indptr, A.indices, A.data.astype(F), x)
    assert np.array_equal(y.astype(np.float64), A @ x.astype(np.float64))  # exact on integers
    A.data = rng.standard_normal(A.nnz)
    xf = rng.standard_normal(n).astype(F)
    y = oracle.spmv(A.indptr, A.indices, A.data.astype(F), xf).astype(np.float64)
    A32 = A.astype(F).astype(np.float64)
    ref = A32 @ xf.astype(np.float64)
    absA = abs(A32) @ abs(xf.astype(np.float64))
    k = np.diff(A.indptr).max()
    assert np.all(np.abs(y - ref) <= (k + 1) * 2 ** -24 * absA + 1e-30)


def test_axpy_scale_two_roundings(oracle):
    e = F(1.0 + 2.0 ** -12)
    y = oracle.axpy(float(e), np.array([e], F), np.array([-(1.0 + 2.0 ** -11)], F))
    assert y[0] == 0.0  # FMA would give 2^-24
    rng = np.random.default_rng(4)
    x = rng.standard_normal(1000).astype(F)
    yy = rng.standard_normal(1000).astype(F)
    a = F(0.37)
    assert np.array_equal(oracle.axpy(float(a), x, yy), yy + a * x)
    assert np.array_equal(oracle.scale(float(a), x), x * a)


# ------------------------------------------------------------------- LU

def _doolittle_exact(A):
    n = len(A)
    M = [[Fraction(v) for v in row] for row in A]
    for k in range(n):
        for i in range(k + 1, n):
            M[i][k] = M[i][k] / M[k][k]
            for j in range(k + 1, n):
                M[i][j] -= M[i][k] * M[k][j]
    return M


def test_lu_exact_dyadic(oracle):
    # A = L U with small integer L (unit lower) and U (power-of-two pivots):
    # every step of the elimination is exact, so the packed factors must equal
    # L and U exactly -- fixes the in-place layout (strict lower = L).
    rng = np.random.default_rng(5)
    for m in range(1, 7):
        L = np.tril(rng.integers(-2, 3, (m, m)), -1) + np.eye(m, dtype=int)
        U = np.triu(rng.integers(-3, 4, (m, m)), 1) + np.diag(2 ** rng.integers(0, 3, m))
        A = L @ U
        mat = np.eye(6, dtype=F)
        mat[:m, :m] = A
        st, f = oracle.lu_factor_inplace_6(mat.ravel(), m)
        assert st == 0
        f = f.reshape(6, 6)
        ex = _doolittle_exact(A.tolist())
        for i in range(m):
            for j in range(m):
                assert f[i, j] == float(ex[i][j]), (m, i, j)
        assert np.array_equal(np.tril(f[:m, :m], -1), np.tril(L, -1).astype(F))
        assert np.array_equal(np.triu(f[:m, :m]), U.astype(F))
        assert np.array_equal(f[m:, m:], np.eye(6 - m, dtype=F))


def test_lu_zero_pivot_text(oracle):
    mat = np.eye(6, dtype=F)
    mat[0, 0] = 0
    st, _ = oracle.lu_factor_inplace_6(mat.ravel(), 6)
    assert st == 2 and oracle.last_error() == "LU zero pivot at k = 0"
    A = np.array([[1, 2, 0], [2, 4, 0], [0, 0, 1]], F)  # pivot 1 vanishes after elimination
    mat = np.eye(6, dtype=F)
    mat[:3, :3] = A
    st, _ = oracle.lu_factor_inplace_6(mat.ravel(), 3)
    assert st == 2 and oracle.last_error() == "LU zero pivot at k = 1"
    st, _ = oracle.lu_factor_inplace_6(np.eye(6, dtype=F).ravel(), 7)
    assert st == 1 and oracle.last_error() == "lu_factor_inplace_6: n must be <= 6, got 7"


def test_lu_random_reconstructs(oracle):
    rng = np.random.default_rng(6)
    for _ in range(50):
        A = rng.standard_normal((6, 6)) + 8 * np.eye(6)
        st, f = oracle.lu_factor_inplace_6(A.astype(F).ravel(), 6)
        f = f.reshape(6, 6).astype(np.float64)
        L = np.tril(f, -1) + np.eye(6)
        U = np.triu(f)
        assert np.allclose(L @ U, A.astype(F), rtol=0, atol=1e-5 * np.abs(A).max())


def test_build_lu_blocks_layout_and_fill(oracle):
    # blocks of size 3, 8 (oversize: leading 6x6 only), 1; duplicate entry
    # (0,0) appears twice in row 0: the last one wins (compute.rs:166-168)
    n = 12
    dense = np.zeros((n, n))
    rng = np.random.default_rng(7)
    for s, e in ((0, 3), (3, 11), (11, 12)):
        B = rng.standard_normal((e - s, e - s))
        dense[s:e, s:e] = B @ B.T + 4 * np.eye(e - s)
    dense[0, 5] = dense[5, 0] = 0.5  # outside every block: ignored
    rows = []
    for i in range(n):
        r = [(j, dense[i, j]) for j in range(n) if dense[i, j] != 0]
        if i == 0:
            r = [(0, 99.0)] + r  # earlier duplicate of (0,0)
        rows.append(r)
    rp, ci, va = csr(rows)
    bs = np.array([0, 3, 11, 12], np.uint32)
    st, lu = oracle.build_lu_blocks_6(n, rp, ci, va, bs)
    assert st == 0
    lu = lu.reshape(3, 6, 6).astype(np.float64)
    for b, (s, m) in enumerate(((0, 3), (3, 6), (11, 1))):
        L = np.tril(lu[b], -1) + np.eye(6)
        U = np.triu(lu[b])
        blk = np.eye(6)
        blk[:m, :m] = dense[s:s + m, s:s + m].astype(F)
        assert np.allclose(L @ U, blk, atol=1e-4 * np.abs(blk).max()), b
        assert np.array_equal(lu[b][m:, m:], np.eye(6 - m))  # identity fill untouched


def test_build_lu_blocks_errors(oracle):
    rp, ci, va = csr([[(0, 1.0)], [(1, 1.0)]])
    st, _ = oracle.build_lu_blocks_6(2, rp, ci, va, np.array([0, 1, 1, 2], np.uint32))
    assert st == 1 and oracle.last_error().endswith("block_starts must be strictly increasing")
    st, _ = oracle.build_lu_blocks_6(2, rp, ci, va, np.array([1, 2], np.uint32))
    assert st == 1 and oracle.last_error().endswith("block_starts must start at 0")
    st, _ = oracle.build_lu_blocks_6(2, rp, ci, va, np.array([0, 1], np.uint32))
    assert st == 1 and "last must equal n (2), got 1" in oracle.last_error()
    st, lu = oracle.build_lu_blocks_6(2, rp, ci, va, np.array([0], np.uint32))
    assert st == 0 and lu.size == 0  # fewer than 2 starts: Ok(vec![])


# ----------------------------------------------------------- block Jacobi

def test_block_jacobi_solves_blocks(oracle):
    rng = np.random.default_rng(8)
    sizes = [6, 3, 6, 1, 6, 2]
    bs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    n = int(bs[-1])
    dense = np.zeros((n, n))
    for s, e in zip(bs[:-1], bs[1:]):
        B = rng.standard_normal((e - s, e - s))
        dense[s:e, s:e] = B @ B.T + 3 * np.eye(e - s)
    A = sp.csr_matrix(dense.astype(F))
    st, lu = oracle.build_lu_blocks_6(n, A.indptr, A.indices, A.data, bs)
    r = rng.standard_normal(n).astype(F)
    z = oracle.block_jacobi(n, lu, bs, r).astype(np.float64)
    for s, e in zip(bs[:-1], bs[1:]):
        blk = dense[s:e, s:e].astype(F).astype(np.float64)
        ref = np.linalg.solve(blk, r[s:e].astype(np.float64))
        cond = np.linalg.cond(blk)
        assert np.allclose(z[s:e], ref, rtol=0, atol=60 * cond * 2 ** -24 * np.abs(ref).max())


def test_block_jacobi_oversize_rows_untouched(oracle):
    # a block of 8 rows: only its leading 6 rows are written (block_jacobi.wgsl:124-126)
    n = 8
    lu = np.eye(6, dtype=F).ravel() * 2  # U diag = 2
    bs = np.array([0, 8], np.uint32)
    z0 = np.full(n, -7.0, F)
    z = oracle.block_jacobi(n, lu, bs, np.arange(1, n + 1, dtype=F), z0)
    assert np.array_equal(z[:6], (np.arange(1, 7) / 2).astype(F))
    assert np.array_equal(z[6:], z0[6:])


# ------------------------------------------------------------------- PCG

def _lu(oracle, d):
    st, lu = oracle.build_lu_blocks_6(int(d["n"]), d["row_ptr"], d["col_idx"], d["values"],
                                      d["block_starts"])
    assert st == 0
    return lu


# commands.txt:26-78 (solve rel 1e-4 / abs 1e-7 / 2000 its; compare-x rtol 1e-3,
# atol 1e-2 / 1e0 / 1e1) and the iteration counts of the survey's independent
# numpy restatement (42 / 370 / 1273).
DATASETS = [(1, 1e-2, 42), (2, 1e0, 370), (3, 1e1, 1273)]


@pytest.mark.parametrize("k,atol,its", DATASETS)
def test_pcg_datasets_match_x_ref(oracle, k, atol, its):
    d = oracle.load_dataset(k)
    lu = _lu(oracle, d)
    st, it, x, _ = oracle.pcg(d["row_ptr"], d["col_idx"], d["values"], lu, d["block_starts"],
                              d["b"], d["x0"], 2000, 1e-4, 1e-7)
    assert st == 0 and it == its
    xr = d["x_ref"].astype(np.float64)
    err = x.astype(np.float64) - xr
    rel_l2 = np.linalg.norm(err) / max(np.linalg.norm(xr), 1e-30)
    rmse = np.sqrt(np.mean(err ** 2))
    assert rel_l2 <= 1e-3 or rmse <= atol  # compare-x PASS rule (cli main.rs:687-689)
    assert rel_l2 <= 2e-4
    # FP64 direct solve of the same FP32 matrix
    n = int(d["n"])
    A = sp.csr_matrix((d["values"].astype(np.float64), d["col_idx"], d["row_ptr"]), shape=(n, n))
    xd = spla.spsolve(A.tocsc(), d["b"].astype(np.float64))
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-3


def test_pcg_matches_textbook_fp64_pcg_on_well_conditioned(oracle):
    # independent textbook block-Jacobi PCG in FP64 (numpy): same stop rule on
    # the recurrence residual; a well-conditioned system makes the FP32
    # rounding irrelevant to the iteration count (+-1).
    rng = np.random.default_rng(9)
    n = 600
    B = sp.random(n, n, density=0.01, random_state=10)
    A = (B @ B.T + sp.eye(n) * 2.0).tocsr().astype(F)
    A.sort_indices()
    bs = np.arange(0, n + 1, 6, dtype=np.uint32)
    b = rng.standard_normal(n).astype(F)
    lu = _lu(oracle, {"n": n, "row_ptr": A.indptr.astype(np.uint32), "col_idx": A.indices.astype(np.uint32),
                      "values": A.data, "block_starts": bs})
    st, it, x, _ = oracle.pcg(A.indptr, A.indices, A.data, lu, bs, b, np.zeros(n, F), 500, 1e-6, 0.0)
    A64 = A.astype(np.float64)
    Minv = sp.block_diag([np.linalg.inv(A64[s:s + 6, s:s + 6].toarray()) for s in range(0, n, 6)])
    xx = np.zeros(n)
    r = b.astype(np.float64).copy()
    z = Minv @ r
    p = z.copy()
    rz = r @ z
    bb = b.astype(np.float64) @ b
    ref_it = None
    for k in range(500):
        q = A64 @ p
        a = rz / (p @ q)
        xx += a * p
        r -= a * q
        if r @ r <= 1e-12 * bb:
            ref_it = k + 1
            break
        z = Minv @ r
        rz, rz_old = r @ z, rz
        p = z + (rz / rz_old) * p
    assert st == 0 and ref_it is not None and abs(it - ref_it) <= 1
    assert np.linalg.norm(x - xx) / np.linalg.norm(xx) < 1e-5


def test_pcg_zero_rhs_is_ok0_and_x_untouched(oracle):
    d = oracle.load_dataset(1)
    lu = _lu(oracle, d)
    x0 = np.arange(int(d["n"]), dtype=F)
    st, it, x, _ = oracle.pcg(d["row_ptr"], d["col_idx"], d["values"], lu, d["block_starts"],
                              np.zeros(int(d["n"]), F), x0, 100, 1e-4, 0.0)
    assert st == 0 and it == 0 and np.array_equal(x, x0)


def test_pcg_breakdown_pap(oracle):
    # identity A, x0 = b: r0 = 0, p = 0, p.Ap = 0 -> breakdown after iteration 1
    n = 12
    rp, ci, va = csr([[(i, 1.0)] for i in range(n)])
    bs = np.arange(0, n + 1, 6, dtype=np.uint32)
    st, lu = oracle.build_lu_blocks_6(n, rp, ci, va, bs)
    b = np.arange(1, n + 1, dtype=F)
    st, it, x, _ = oracle.pcg(rp, ci, va, lu, bs, b, b.copy(), 100, 1e-6, 0.0)
    assert st == 3 and it == 1
    assert oracle.last_error() == "PCG(BlockJacobiGpu): dot(p,Ap) is zero (breakdown)"


def test_pcg_breakdown_rz(oracle):
    # indefinite diagonal M = diag(1,-1,1) and b = (1,1,0): r0.z0 = 0 while
    # p.Ap = -2a != 0 -> the rz_old check fires (compute.rs:476-478)
    a = 0.5
    rp, ci, va = csr([[(0, 1.0), (1, a)], [(0, a), (1, -1.0)], [(2, 1.0)]])
    bs = np.array([0, 1, 2, 3], np.uint32)
    st, lu = oracle.build_lu_blocks_6(3, rp, ci, va, bs)
    st, it, x, _ = oracle.pcg(rp, ci, va, lu, bs, np.array([1, 1, 0], F), np.zeros(3, F), 10, 1e-6, 0.0)
    assert st == 4 and it == 1 and np.array_equal(x, np.zeros(3, F))
    assert oracle.last_error() == "PCG(BlockJacobiGpu): rz_old is zero (breakdown)"


def test_pcg_not_converged_keeps_x0(oracle):
    d = oracle.load_dataset(2)
    lu = _lu(oracle, d)
    x0 = np.full(int(d["n"]), 0.25, F)
    for mi in (0, 1, 5):
        st, it, x, _ = oracle.pcg(d["row_ptr"], d["col_idx"], d["values"], lu, d["block_starts"],
                                  d["b"], x0, mi, 1e-4, 1e-7)
        assert st == 5 and np.array_equal(x, x0)
        assert oracle.last_error() == f"PCG(BlockJacobiGpu): did not converge in {mi} iterations"


def test_fp64_oracle_vs_scipy(oracle, hbj):
    # datasets 1-2: FP64 attainable relative residual is ~1e-10 even for a
    # direct solve (entries up to 1e12, heavy cancellation), so the FP64 PCG
    # is checked by its error against scipy's direct solve.
    for k in (1, 2):
        d = oracle.load_dataset(k)
        n = int(d["n"])
        A = sp.csr_matrix((d["values"].astype(np.float64), d["col_idx"], d["row_ptr"]), shape=(n, n))
        b = d["b"].astype(np.float64)
        xd = spla.spsolve(A.tocsc(), b)
        st, it, x = oracle.pcg64(d["row_ptr"], d["col_idx"], d["values"], d["block_starts"], b,
                                 np.zeros(n), 4000, 1e-14)
        tr = oracle.true_rel_residual64(d["row_ptr"], d["col_idx"], d["values"], b, x)
        assert abs(tr - np.linalg.norm(b - A @ x) / np.linalg.norm(b)) <= 1e-12 * max(tr, 1e-3)
        assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    # a beam lattice (cond ~1e7): residual-based convergence to 1e-12
    L = hbj.lattice(4, seed=3)
    n = L["n"]
    A = sp.csr_matrix((L["values"].astype(np.float64), L["col_idx"], L["row_ptr"]), shape=(n, n))
    b = L["rhs"].astype(np.float64)
    st, it, x = oracle.pcg64(L["row_ptr"], L["col_idx"], L["values"], L["block_starts"], b,
                             np.zeros(n), 5000, 1e-12)
    assert st == 0
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-12
    xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8


def test_oracle_residual_histories_match_the_fixture(oracle, hbj):
    """tests/golden/histories.npz (tests/golden/make_histories.py): the FP32
    r.r of every iteration on the three datasets and the 10^3 lattice, bit for
    bit, and the lattice generator's output by digest."""
    import hashlib
    fx = np.load(os.path.join(GOLDEN, "histories.npz"))
    L = hbj.lattice(10, seed=0)
    h = hashlib.sha256()
    for k in ("row_ptr", "col_idx", "values", "rhs", "block_starts"):
        h.update(np.ascontiguousarray(L[k]).tobytes())
    assert h.hexdigest() == str(fx["lattice10_sha256"])
    cases = []
    for k in (1, 2, 3):
        d = oracle.load_dataset(k)
        cases.append((f"ds{k}", d["row_ptr"], d["col_idx"], d["values"], d["block_starts"], d["b"],
                      d["x0"], 2000, 1e-4, 1e-7))
    cases.append(("lattice10", L["row_ptr"], L["col_idx"], L["values"], L["block_starts"], L["rhs"],
                  np.zeros(L["n"], F), 5000, 1e-6, 0.0))
    for name, rp, ci, va, bs, b, x0, mi, rel, ab in cases:
        lu = _lu(oracle, {"n": len(rp) - 1, "row_ptr": rp, "col_idx": ci, "values": va, "block_starts": bs})
        st, it, x, hist = oracle.pcg(rp, ci, va, lu, bs, b, x0, mi, rel, ab, history=True)
        assert [st, it] == fx[f"{name}_status_iterations"].tolist(), name
        assert np.array_equal(hist.view(np.uint32), fx[f"{name}_rr"].view(np.uint32)), name
