"""A created system's own SpMV (hbj_system_spmv: the iteration's kernel on the
resident layout, the chosen variant) against the oracle's row sums (or_spmv,
spmv.wgsl:43-64), bit for bit: every variant, the sliced layout with and
without its nonzero-only waves, the 50^3 lattice; and vectors holding Inf and
NaN, where a row's left-out +-0 entries must still make its sum NaN as the
full row's 0 * Inf does."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F = np.float32


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def same(y, yo):
    """NaN where the oracle has NaN, every other value bit for bit."""
    nan = np.isnan(yo)
    return np.array_equal(np.isnan(y), nan) and np.array_equal(bits(y)[~nan], bits(yo)[~nan])


def system(hbj, ctx, L):
    return hbj.System(ctx, L["row_ptr"], L["col_idx"], L["values"], L["block_starts"])


@pytest.mark.parametrize("nz", ["1", "0"])
def test_system_spmv_every_variant(ctx, hbj, oracle, monkeypatch, nz):
    monkeypatch.setenv("HBJ_SLICED_NZ", nz)
    L = hbj.lattice(12, seed=3)
    s = system(hbj, ctx, L)
    info = s.info()
    print(f"12^3 lattice, HBJ_SLICED_NZ={nz}: {info['sliced_nz_waves']} nonzero-only waves", flush=True)
    # nz=1: the clamped layer's waves (rows of one entry) keep every entry,
    # the others store their nonzeros only -- both kinds in one system
    assert (0 < info["sliced_nz_waves"] < 4 * info["n_tiles"]) == (nz == "1")
    assert (info["sliced_nz_waves"] == 0) == (nz == "0")
    x = np.random.default_rng(4).standard_normal(L["rhs"].size).astype(F)
    yo = oracle.spmv(L["row_ptr"], L["col_idx"], L["values"], x)
    for v in range(6):
        s.set_spmv_variant(v)
        assert np.array_equal(bits(s.spmv(x)), bits(yo)), f"variant {v}"
    s.close()


def test_system_spmv_lattice_50(ctx, hbj, oracle):
    """The benchmark's 50^3 lattice (30,352,200 entries, 69% of them explicit
    zeros) through both sliced variants."""
    L = hbj.lattice(50)
    s = system(hbj, ctx, L)
    assert s.info()["sliced_nz_waves"] > 0
    x = np.random.default_rng(50).uniform(-1.0, 1.0, L["rhs"].size).astype(F)
    yo = oracle.spmv(L["row_ptr"], L["col_idx"], L["values"], x)
    for v in (4, 5):
        print(f"50^3 lattice, variant {v}", flush=True)
        s.set_spmv_variant(v)
        assert np.array_equal(bits(s.spmv(x)), bits(yo))
    s.close()


@pytest.mark.parametrize("nz", ["1", "0"])
def test_system_spmv_nonfinite_vector(ctx, hbj, oracle, monkeypatch, nz):
    """+Inf, -Inf and NaN in x: rows with a nonzero entry in those columns
    give +-Inf or NaN, rows whose entry there is an explicit zero give NaN
    (0 * Inf), every other row its finite sum -- with the zeros left out of
    the stream as with every entry kept."""
    monkeypatch.setenv("HBJ_SLICED_NZ", nz)
    L = hbj.lattice(8, seed=2)
    rp, ci, va = L["row_ptr"], L["col_idx"], L["values"]
    n = rp.size - 1
    s = system(hbj, ctx, L)
    x = np.random.default_rng(8).standard_normal(n).astype(F)
    cols = [6 * 300 + 1, 6 * 301 + 4, 6 * 450 + 2]  # free nodes (not in the clamped layer)
    x[cols[0]], x[cols[1]], x[cols[2]] = np.inf, -np.inf, np.nan
    yo = oracle.spmv(rp, ci, va, x)
    # rows that are NaN only through an explicit zero in a non-finite column
    rows = np.repeat(np.arange(n), np.diff(rp))
    zero_hit = np.zeros(n, bool)
    zero_hit[rows[np.isin(ci, cols[:2]) & (va == 0.0)]] = True
    nonzero_hit = np.zeros(n, bool)
    nonzero_hit[rows[np.isin(ci, cols) & (va != 0.0)]] = True
    only_zero = zero_hit & ~nonzero_hit
    assert only_zero.sum() > 0 and np.isnan(yo[only_zero]).all()
    assert np.isinf(yo).any() and np.isfinite(yo).sum() > n // 2
    for v in (4, 5):
        s.set_spmv_variant(v)
        y = s.spmv(x)
        assert same(y, yo), f"variant {v}"
    # a finite vector afterwards: no flag is left over from the previous call
    x2 = np.random.default_rng(9).standard_normal(n).astype(F)
    assert np.array_equal(bits(s.spmv(x2)), bits(oracle.spmv(rp, ci, va, x2)))
    s.close()
