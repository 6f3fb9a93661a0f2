"""The C ABI library loads, exports every declared entry point, and its
host-side functions (LU build, validation) agree with the oracle bit for bit."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(hbj_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol(hbj):
    L = hbj._lib.lib()
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert L.hbj_version().decode().startswith("hbj")


def test_no_device_error_without_gpu(hbj):
    import torch  # noqa: F401  (device counting only)
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(hbj.SolverError) as e:
        hbj.GpuContext.create(0)
    assert e.value.status_name == "NO_DEVICE"


def test_lu_factor_matches_oracle(hbj, oracle):
    rng = np.random.default_rng(11)
    for m in range(0, 7):
        for _ in range(20):
            A = rng.standard_normal((6, 6)).astype(np.float32) + 3 * np.eye(6, dtype=np.float32)
            st, ref = oracle.lu_factor_inplace_6(A.ravel(), m)
            got = hbj.lu_factor_inplace_6(A.ravel(), m)
            assert st == 0 and np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    A = np.eye(6, dtype=np.float32)
    A[2, 2] = 0
    with pytest.raises(hbj.SolverError) as e:
        hbj.lu_factor_inplace_6(A.ravel(), 6)
    assert str(e.value) == "LU zero pivot at k = 2" and e.value.status_name == "ZERO_PIVOT"


@pytest.mark.parametrize("k", [1, 2, 3])
def test_build_lu_blocks_matches_oracle(hbj, oracle, k):
    d = oracle.load_dataset(k)
    st, ref = oracle.build_lu_blocks_6(int(d["n"]), d["row_ptr"], d["col_idx"], d["values"],
                                       d["block_starts"])
    got = hbj.build_lu_blocks_from_csr_block_starts_6(int(d["n"]), d["row_ptr"], d["col_idx"],
                                                      d["values"], d["block_starts"])
    assert st == 0 and np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_build_lu_blocks_lattice_threaded_matches_oracle(hbj, oracle):
    L = hbj.lattice(24, seed=5, e_log10_half_range=1.5)  # 13824 blocks: multi-threaded path
    st, ref = oracle.build_lu_blocks_6(L["n"], L["row_ptr"], L["col_idx"], L["values"],
                                       L["block_starts"])
    got = hbj.build_lu_blocks_from_csr_block_starts_6(L["n"], L["row_ptr"], L["col_idx"],
                                                      L["values"], L["block_starts"])
    assert st == 0 and np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    big = hbj.lattice(42, seed=1)  # 74088 blocks >= the threading threshold
    st, ref = oracle.build_lu_blocks_6(big["n"], big["row_ptr"], big["col_idx"], big["values"],
                                       big["block_starts"])
    got = hbj.build_lu_blocks_from_csr_block_starts_6(big["n"], big["row_ptr"], big["col_idx"],
                                                      big["values"], big["block_starts"])
    assert st == 0 and np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_build_lu_blocks_errors_match_reference_texts(hbj, oracle):
    rp = np.array([0, 1, 2], np.uint32)
    ci = np.array([0, 1], np.uint32)
    va = np.array([1, 1], np.float32)
    cases = [np.array([0, 1, 1, 2], np.uint32), np.array([1, 2], np.uint32), np.array([0, 1], np.uint32)]
    for bs in cases:
        st, _ = oracle.build_lu_blocks_6(2, rp, ci, va, bs)
        with pytest.raises(hbj.SolverError) as e:
            hbj.build_lu_blocks_from_csr_block_starts_6(2, rp, ci, va, bs)
        assert e.value.status == st == 1
        assert str(e.value) == oracle.last_error()
    # zero pivot in the second block: first failing block reports
    va0 = np.array([1, 0], np.float32)
    with pytest.raises(hbj.SolverError) as e:
        hbj.build_lu_blocks_from_csr_block_starts_6(2, rp, ci, va0, np.array([0, 1, 2], np.uint32))
    assert str(e.value) == "LU zero pivot at k = 0"
    with pytest.raises(hbj.SolverError) as e:
        hbj.build_lu_blocks_from_csr_block_starts_6(3, rp, ci, va, np.array([0, 3], np.uint32))
    assert str(e.value) == "build_lu_blocks_from_csr_block_starts_6: row_ptr len must be n+1 (4), got 3"
