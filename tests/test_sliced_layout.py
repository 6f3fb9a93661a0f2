"""Host setup of the sliced SpMV layout (hbj_sliced_layout_info), against the
rule restated in numpy: a 64-row wave whose rows all have at most 64 entries
stores only its nonzero values (two mask words per row ahead of them) when
that takes at most 80% of the bytes of all its entries, steps of 4 either
way.  The GPU tests check the layout's products bit for bit; this pins which
waves take which form and how many value words the layout holds."""
import numpy as np
import pytest

F = np.float32


def expected(rp, va, allow_nz=True):
    n = rp.size - 1
    lens = np.diff(rp).astype(np.int64)
    nzc = np.add.reduceat((va != 0).astype(np.int64), rp[:-1].astype(np.int64)) if va.size else np.zeros(n, np.int64)
    nzc = np.where(lens > 0, nzc, 0)  # reduceat repeats the next entry for empty rows
    nt = (n + 255) // 256
    waves = nz_waves = words = 0
    for w in range(4 * nt):
        r0, r1 = 64 * w, min(64 * w + 64, n)
        waves += 1
        if r0 >= n:
            continue
        maxlen = int(lens[r0:r1].max())
        maxnz = int(nzc[r0:r1].max())
        nq, nq_nz = (maxlen + 3) // 4, (maxnz + 3) // 4
        nz = allow_nz and 0 < maxlen <= 64 and 5 * (1024 * nq_nz + 512) <= 4 * 1024 * nq
        nz_waves += nz
        words += 256 * nq_nz + 128 if nz else 256 * nq
    return waves, nz_waves, words


@pytest.mark.parametrize("N", [4, 6, 9])
def test_lattice_waves_store_nonzeros(hbj, N):
    L = hbj.lattice(N, seed=N)
    rp, ci, va = L["row_ptr"], L["col_idx"], L["values"]
    info = hbj.sliced_layout_info(rp, ci, va)
    waves, nz_waves, words = expected(rp, va)
    assert info["sliced"] == 1
    assert (info["waves"], info["nz_waves"], info["value_words"]) == (waves, nz_waves, words)
    assert 0 < nz_waves < waves  # the clamped layer's waves (one entry per row) keep every entry
    full = hbj.sliced_layout_info(rp, ci, va, allow_nz=False)
    assert full["nz_waves"] == 0 and full["value_words"] == expected(rp, va, False)[2]
    assert info["value_words"] < 0.6 * full["value_words"]  # 69% of the entries are zeros
    assert info["run_words"] == full["run_words"]


def test_no_stored_zeros_keep_every_entry(hbj, oracle):
    """The reference's datasets store no zeros: nothing to leave out."""
    d = oracle.load_dataset(2)
    info = hbj.sliced_layout_info(d["row_ptr"], d["col_idx"], d["values"])
    assert info["nz_waves"] == 0


def _band(n, width, zero_every, rng):
    """Rows of `width` consecutive columns around the diagonal; every
    zero_every-th entry stored as +0.0 or -0.0."""
    rows, cols = [], []
    for r in range(n):
        c0 = max(0, min(r - width // 2, n - width))
        cols.extend(range(c0, c0 + width))
        rows.extend([r] * width)
    va = rng.uniform(0.5, 1.5, len(cols)).astype(F)
    if zero_every:
        va[::zero_every] = 0.0
        va[1::2 * zero_every] = -0.0
    rp = np.arange(0, n * width + 1, width, dtype=np.uint32)
    return rp, np.asarray(cols, np.uint32), va


@pytest.mark.parametrize("width,zero_every,want_nz", [
    (40, 2, True),    # half the entries are zeros: nonzero-only waves
    (40, 10, False),  # a tenth: not worth the masks
    (80, 2, False),   # rows longer than a mask covers
])
def test_band_matrices(hbj, width, zero_every, want_nz):
    rp, ci, va = _band(1000, width, zero_every, np.random.default_rng(width + zero_every))
    info = hbj.sliced_layout_info(rp, ci, va)
    waves, nz_waves, words = expected(rp, va)
    assert info["sliced"] == 1
    assert (info["nz_waves"], info["value_words"]) == (nz_waves, words)
    assert (nz_waves > 0) == want_nz
