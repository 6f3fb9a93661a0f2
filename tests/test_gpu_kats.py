"""Single-op executors on the GPU: the reference's five self-tests (cli
main.rs:171-443) and bit-for-bit parity with the oracle on larger inputs."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KATS = json.load(open(os.path.join(GOLDEN, "kats.json")))
F = np.float32


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_vec_test(ctx, hbj):
    k = KATS["vec_test"]
    ex = hbj.VecOpsExecutor.create(ctx)
    y = ex.axpy_inplace(np.full(k["n"], k["x"], F), np.full(k["n"], k["y"], F), k["alpha"])
    assert np.all(np.abs(y - k["expected"]) < k["tol"])


def test_dot_test(ctx, hbj):
    k = KATS["dot_test"]
    ex = hbj.DotScalarExecutor.create(ctx, 3, 4)
    assert abs(ex.dot_scalar_into(k["a"], k["b"], 0) - k["expected"]) < k["tol"]


def test_spmv_test(ctx, hbj):
    k = KATS["spmv_test"]
    ex = hbj.SpmvExecutor.create(ctx, 3, k["row_ptr"], k["col_idx"], k["values"])
    y = ex.spmv(np.array(k["x"], F))
    assert np.all(np.abs(y - np.array(k["expected"], F)) < k["tol"])


def test_block_jacobi_test(ctx, hbj):
    k = KATS["block_jacobi_test"]
    lu = np.zeros(72, F)
    for b in range(2):
        for i in range(6):
            lu[b * 36 + i * 6 + i] = 1
    ex = hbj.BlockJacobiExecutor.create(ctx, k["n"], lu, k["block_starts"])
    z = ex.apply(np.array(k["r"], F))
    assert np.all(np.abs(z - np.array(k["expected"], F)) < k["tol"])


def test_pcg_update_scalars_test(ctx, hbj):
    k = KATS["pcg_update_scalars_test"]
    s = np.zeros(k["slots"], F)
    s[k["PAP"]], s[k["RZ_OLD"]], s[k["RZ_NEW"]] = k["pAp"], k["rz_old"], k["rz_new"]
    ex = hbj.PcgUpdateScalarsExecutor.create(ctx)
    o = ex.update_scalars(s, k["PAP"], k["RZ_NEW"], k["RZ_OLD"], k["ALPHA"], k["MINUS_ALPHA"], k["BETA"])
    assert o[k["ALPHA"]] == k["alpha"] and o[k["MINUS_ALPHA"]] == k["minus_alpha"]
    assert o[k["BETA"]] == k["beta"]
    z = ex.update_scalars(np.array([0, 5, 0, 9, 9, 9], F), 0, 1, 2, 3, 4, 5)
    assert z[3] == 0 and z[4] == 0 and z[5] == 0


@pytest.mark.parametrize("n", [1, 255, 256, 257, 4096, 65536, 65537, 196611, 20_000_000])
def test_dot_bitwise(ctx, hbj, oracle, n):
    print(f"dot n={n}", flush=True)
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n).astype(F)
    b = rng.standard_normal(n).astype(F)
    got = hbj.DotScalarExecutor.create(ctx, n, 1).dot(a, b)
    assert bits(got) == bits(oracle.dot(a, b))


def test_dot_tree_cases(ctx, hbj):
    ex = hbj.DotScalarExecutor.create(ctx, 1024, 1)
    a = np.zeros(256, F)
    a[0], a[1], a[3] = 2.0 ** 24, 1, 1
    assert ex.dot(a, np.ones(256, F)) == 2.0 ** 24 + 2
    a = np.zeros(1024, F)
    a[0], a[256], a[768] = 2.0 ** 24, 1, 1
    assert ex.dot(a, np.ones(1024, F)) == 2.0 ** 24 + 2
    assert ex.dot(np.zeros(0, F), np.zeros(0, F)) == 0.0


def test_axpy_scale_bitwise(ctx, hbj, oracle):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(100_003).astype(F)
    y = rng.standard_normal(100_003).astype(F)
    ex = hbj.VecOpsExecutor.create(ctx)
    assert np.array_equal(bits(ex.axpy_inplace(x, y, -0.731)), bits(oracle.axpy(-0.731, x, y)))
    assert np.array_equal(bits(ex.scale_inplace(x, 1.37)), bits(oracle.scale(1.37, x)))
    e = F(1 + 2.0 ** -12)
    assert ex.axpy_inplace(np.array([e], F), np.array([-(1 + 2.0 ** -11)], F), float(e))[0] == 0.0


@pytest.mark.parametrize("k", [1, 2, 3])
def test_spmv_bitwise_datasets(ctx, hbj, oracle, k):
    d = oracle.load_dataset(k)
    n = int(d["n"])
    rng = np.random.default_rng(k)
    x = rng.standard_normal(n).astype(F)
    ex = hbj.SpmvExecutor.create(ctx, n, d["row_ptr"], d["col_idx"], d["values"])
    assert np.array_equal(bits(ex.spmv(x)), bits(oracle.spmv(d["row_ptr"], d["col_idx"], d["values"], x)))


def test_spmv_bitwise_irregular_rows(ctx, hbj, oracle):
    # empty rows, a 50k-entry row spanning many staging chunks, order-sensitive rows
    rng = np.random.default_rng(5)
    n = 5000
    lens = rng.integers(0, 60, n)
    lens[::97] = 0
    lens[1234] = 50_000
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    ci = rng.integers(0, n, int(rp[-1])).astype(np.uint32)
    va = (rng.standard_normal(int(rp[-1])) * 10.0 ** rng.integers(-6, 7, int(rp[-1]))).astype(F)
    x = rng.standard_normal(n).astype(F)
    ex = hbj.SpmvExecutor.create(ctx, n, rp, ci, va)
    assert np.array_equal(bits(ex.spmv(x)), bits(oracle.spmv(rp, ci, va, x)))
    rp3 = np.array([0, 3, 4, 5], np.uint32)  # [2^24, 1, 1] row: sequential order only
    y = hbj.SpmvExecutor.create(ctx, 3, rp3, np.array([0, 1, 2, 1, 2], np.uint32),
                                np.array([2.0 ** 24, 1, 1, 1, 1], F)).spmv(np.ones(3, F))
    assert y[0] == 2.0 ** 24


@pytest.mark.parametrize("k", [2, 3])
def test_block_jacobi_bitwise_datasets(ctx, hbj, oracle, k):
    d = oracle.load_dataset(k)
    n = int(d["n"])
    st, lu = oracle.build_lu_blocks_6(n, d["row_ptr"], d["col_idx"], d["values"], d["block_starts"])
    r = np.random.default_rng(k).standard_normal(n).astype(F)
    ex = hbj.BlockJacobiExecutor.create(ctx, n, lu, d["block_starts"])
    assert np.array_equal(bits(ex.apply(r)), bits(oracle.block_jacobi(n, lu, d["block_starts"], r)))


def test_block_jacobi_mixed_and_oversize_blocks(ctx, hbj, oracle):
    # sizes 1..9 (oversize blocks keep stale rows), blocks straddling tile
    # edges, and many size-1 blocks in one tile (more than the LDS stage)
    rng = np.random.default_rng(9)
    sizes = list(rng.integers(1, 10, 300)) + [1] * 400 + [6] * 100
    bs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    n = int(bs[-1])
    nb = len(sizes)
    lu = rng.standard_normal(36 * nb).astype(F)
    for b in range(nb):
        for i in range(6):
            lu[b * 36 + i * 7] = F(3 + rng.random())
    r = rng.standard_normal(n).astype(F)
    z0 = np.full(n, 5.5, F)
    ex = hbj.BlockJacobiExecutor.create(ctx, n, lu, bs)
    got = ex.apply(r, z0)
    ref = oracle.block_jacobi(n, lu, bs, r, z0)
    assert np.array_equal(bits(got), bits(ref))
