"""Host setup of the SpMV's tile-local layout (hbj_spmv_tile_tables), checked
against numpy: per 256-row tile, np.unique of its entries' columns is the list
(none when it has more than 2048 columns) and np.searchsorted gives each
entry's index.  The GPU parity tests check the same tables end to end (an
entry pointing at a wrong list position changes the SpMV's result)."""
import numpy as np
import pytest

OFF = 0xFFFFFFFF


def expected(rp, ci):
    n = rp.size - 1
    nt = (n + 255) // 256
    off, cnt = np.zeros(nt, np.uint64), np.zeros(nt, np.uint64)
    lidx = np.zeros(ci.size, np.uint16)
    lists, total = [], 0
    for t in range(nt):
        e0, e1 = rp[t * 256], rp[min(t * 256 + 256, n)]
        u = np.unique(ci[e0:e1])
        off[t] = total
        if u.size > 2048:
            cnt[t] = OFF
            continue
        cnt[t] = u.size
        lidx[e0:e1] = np.searchsorted(u, ci[e0:e1])
        lists.append(u)
        total += u.size
    tcols = np.concatenate(lists) if lists else np.zeros(0, np.uint32)
    return off, cnt, lidx, tcols


def random_csr(rng, n, per_row, cols, counts=None):
    """rows of 0..2*per_row entries (or the given counts); cols(rows) draws one
    column for each entry, given the entries' row indices"""
    if counts is None:
        counts = rng.integers(0, 2 * per_row + 1, n)
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum(counts)
    rows = np.repeat(np.arange(n), counts)
    return rp, np.asarray(cols(rows), np.int64).astype(np.uint32)


def cases():
    rng = np.random.default_rng(11)
    yield "random", random_csr(rng, 5000, 8, lambda r: rng.integers(0, 5000, r.size))
    # most tiles reference more than 2048 columns: no list
    yield "dense rows", random_csr(rng, 3000, 40, lambda r: rng.integers(0, 3000, r.size))
    # bands of 1601 and 2601 columns in alternate tiles: lists below and above
    # the 2048 limit
    yield "band", random_csr(rng, 20001, 30, lambda r: np.clip(
        r + (rng.random(r.size) * 2 - 1) * np.where(r // 256 % 2 == 0, 800, 1300), 0, 20000).astype(np.int64))
    # the first and last column in many rows: a span far wider than the entries
    yield "arrow", random_csr(rng, 20000, 3, lambda r: np.where(rng.random(r.size) < 0.3,
                                                                np.where(r % 2 == 0, 0, 19999),
                                                                rng.integers(0, 20000, r.size)))
    # columns at both ends of a range far wider than a tile's entries (the
    # builder sorts such a tile's columns instead of marking a bitmap)
    yield "two clusters", random_csr(rng, 1000000, 2,
                                     lambda r: np.where(rng.random(r.size) < 0.5, rng.integers(0, 8, r.size),
                                                        999999 - rng.integers(0, 8, r.size)))
    yield "repeated columns", random_csr(rng, 3000, 20, lambda r: rng.integers(0, 50, r.size))
    yield "empty rows", random_csr(rng, 4097, 1, lambda r: rng.integers(0, 4097, r.size),
                                   counts=np.where(np.arange(4097) % 3 == 0, 2, 0))
    yield "n=0", (np.zeros(1, np.uint32), np.zeros(0, np.uint32))
    yield "one empty row", (np.zeros(2, np.uint32), np.zeros(0, np.uint32))


@pytest.mark.parametrize("name", [c[0] for c in cases()])
def test_tile_tables_match_numpy(hbj, name):
    rp, ci = dict(cases())[name]
    off, cnt, lidx, tcols = hbj.spmv_tile_tables(rp, ci)
    e_off, e_cnt, e_lidx, e_tcols = expected(rp, ci)
    assert np.array_equal(cnt, e_cnt)
    assert np.array_equal(off, e_off)
    assert np.array_equal(lidx, e_lidx)
    assert np.array_equal(tcols, e_tcols)
    n_off = int(np.sum(cnt == OFF))
    if name == "dense rows":
        assert n_off == cnt.size
    if name == "band":
        assert 0 < n_off < cnt.size
    # every entry of a listed tile finds its own column
    n = rp.size - 1
    for t in np.flatnonzero(cnt != OFF):
        e0, e1 = rp[t * 256], rp[min(t * 256 + 256, n)]
        assert np.array_equal(tcols[off[t] + lidx[e0:e1]], ci[e0:e1])


@pytest.mark.parametrize("N", [3, 7, 12])
def test_tile_tables_lattice(hbj, N):
    L = hbj.lattice(N, seed=4)
    off, cnt, lidx, tcols = hbj.spmv_tile_tables(L["row_ptr"], L["col_idx"])
    e = expected(L["row_ptr"], L["col_idx"])
    assert all(np.array_equal(a, b) for a, b in zip((off, cnt, lidx, tcols), e))
    assert np.all(cnt != OFF)  # a lattice tile references a few hundred columns


def test_tile_tables_reject_bad_csr(hbj):
    with pytest.raises(hbj.SolverError, match="out-of-range column 7 for n_cols=4"):
        hbj.spmv_tile_tables(np.array([0, 1, 1, 2, 2], np.uint32), np.array([0, 7], np.uint32))
