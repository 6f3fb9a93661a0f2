// setup.cpp -- host-side system setup: CSR / block_starts validation, the
// 6x6 no-pivot LU block builder (with the reference's error texts) and the
// tile-local column tables of the SpMV.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "hbj_internal.hpp"

namespace hbj {

static thread_local std::string g_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
}
void clear_error() { g_error.clear(); }

// loaders.rs:84-105 (CSR) and :164-184 (block_starts)
hbj_status validate_case(uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                         const uint32_t* col_idx, const uint32_t* bs, uint32_t n_starts) {
  if (!row_ptr || (nnz && !col_idx) || !bs) {
    set_error("null pointer argument");
    return HBJ_E_INVALID;
  }
  if (row_ptr[0] != 0) {
    set_error("CSR row_ptr must start with 0");
    return HBJ_E_INVALID;
  }
  if (row_ptr[n] != nnz) {
    set_error("CSR row_ptr must end with nnz (%u), got Some(%u)", nnz, row_ptr[n]);
    return HBJ_E_INVALID;
  }
  for (uint32_t i = 0; i < n; ++i) {
    if (row_ptr[i + 1] < row_ptr[i]) {
      set_error("CSR row_ptr must be non-decreasing");
      return HBJ_E_INVALID;
    }
  }
  uint32_t max_col = 0;
  for (uint32_t k = 0; k < nnz; ++k) max_col = col_idx[k] > max_col ? col_idx[k] : max_col;
  if (n > 0 && nnz > 0 && max_col >= n) {
    set_error("CSR col_idx contains out-of-range column %u for n_cols=%u", max_col, n);
    return HBJ_E_INVALID;
  }
  if (n_starts == 0) {
    set_error("block_starts must not be empty");
    return HBJ_E_INVALID;
  }
  if (bs[0] != 0) {
    set_error("block_starts must start with 0, got %u", bs[0]);
    return HBJ_E_INVALID;
  }
  if (bs[n_starts - 1] != n) {
    set_error("block_starts must end with n (%u), got %u", n, bs[n_starts - 1]);
    return HBJ_E_INVALID;
  }
  for (uint32_t w = 0; w + 1 < n_starts; ++w) {
    if (bs[w + 1] <= bs[w]) {
      set_error("block_starts must be strictly increasing");
      return HBJ_E_INVALID;
    }
  }
  return HBJ_OK;
}

// Runs fn(i0, i1) over [0, count) on up to 16 threads (one when count is
// below `serial`).
template <typename Fn>
static void parallel_ranges(uint32_t count, uint32_t serial, Fn fn) {
  unsigned nth = count >= serial ? std::thread::hardware_concurrency() : 1u;
  if (nth == 0) nth = 1;
  if (nth > 16) nth = 16;
  if (nth == 1) {
    fn(0u, count);
    return;
  }
  std::vector<std::thread> th;
  const uint32_t per = (count + nth - 1) / nth;
  for (unsigned t = 0; t < nth; ++t) {
    const uint32_t a = t * per, b = (a + per) < count ? a + per : count;
    if (a < b) th.emplace_back(fn, a, b);
  }
  for (auto& x : th) x.join();
}

// One tile's table: its sorted distinct columns into u (false and u cleared
// when there are more than kLocalMax), each entry's index in u into lidx.
// Distinct columns are bits of a bitmap over [cmin, cmax], so the list comes
// out sorted and an entry's index is the number of set bits below its column
// (a per-word prefix count plus a popcount): linear in the entries and the
// span / 64.  A tile whose columns spread over more than 64 times twice its
// entry count (plus 4096 words) sorts its columns instead.
static bool tile_table(const uint32_t* col, uint32_t e0, uint32_t e1, std::vector<uint64_t>& bits,
                       std::vector<uint32_t>& rank, std::vector<uint32_t>& u, uint16_t* lidx) {
  u.clear();
  if (e0 == e1) return true;
  uint32_t cmin = col[e0], cmax = col[e0];
  for (uint32_t e = e0 + 1; e < e1; ++e) {
    cmin = col[e] < cmin ? col[e] : cmin;
    cmax = col[e] > cmax ? col[e] : cmax;
  }
  const uint32_t words = ((cmax - cmin) >> 6) + 1;
  if ((uint64_t)words > 2ull * (e1 - e0) + 4096ull) {
    u.assign(col + e0, col + e1);
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    if (u.size() > kLocalMax) {
      u.clear();
      return false;
    }
    for (uint32_t e = e0; e < e1; ++e)
      lidx[e] = (uint16_t)(std::lower_bound(u.begin(), u.end(), col[e]) - u.begin());
    return true;
  }
  bits.assign(words, 0ull);
  for (uint32_t e = e0; e < e1; ++e) {
    const uint32_t c = col[e] - cmin;
    bits[c >> 6] |= 1ull << (c & 63u);
  }
  rank.resize(words);
  uint32_t total = 0;
  for (uint32_t w = 0; w < words; ++w) {
    rank[w] = total;
    total += (uint32_t)__builtin_popcountll(bits[w]);
  }
  if (total > kLocalMax) return false;
  u.reserve(total);
  for (uint32_t w = 0; w < words; ++w)
    for (uint64_t m = bits[w]; m; m &= m - 1)
      u.push_back(cmin + (w << 6) + (uint32_t)__builtin_ctzll(m));
  for (uint32_t e = e0; e < e1; ++e) {
    const uint32_t c = col[e] - cmin;
    lidx[e] = (uint16_t)(rank[c >> 6] + (uint32_t)__builtin_popcountll(bits[c >> 6] & ((1ull << (c & 63u)) - 1ull)));
  }
  return true;
}

void build_tile_local(uint32_t n, const uint32_t* row_ptr, const uint32_t* col,
                      std::vector<uint16_t>& lidx, std::vector<uint2>& tloc,
                      std::vector<uint32_t>& tcols) {
  const uint32_t nt = (n + kTile - 1) / kTile;
  const uint32_t nnz = n ? row_ptr[n] : 0u;
  std::vector<std::vector<uint32_t>> lists(nt);
  tloc.assign(nt, make_uint2(0u, 0u));
  lidx.assign((size_t)nnz + 8u, 0u);
  parallel_ranges(nt, 64, [&](uint32_t t0, uint32_t t1) {
    std::vector<uint64_t> bits;
    std::vector<uint32_t> rank;
    for (uint32_t t = t0; t < t1; ++t) {
      const uint32_t r0 = t * kTile, r1 = std::min<uint32_t>(r0 + kTile, n);
      // entries of a tile without a list keep index 0
      if (!tile_table(col, row_ptr[r0], row_ptr[r1], bits, rank, lists[t], lidx.data()))
        tloc[t].y = kLocalOff;
    }
  });
  size_t total = 0;
  for (uint32_t t = 0; t < nt; ++t) {
    tloc[t].x = (uint32_t)total;
    if (tloc[t].y != kLocalOff) tloc[t].y = (uint32_t)lists[t].size();
    total += lists[t].size();
  }
  tcols.resize(total ? total : 1u, 0u);
  parallel_ranges(nt, 64, [&](uint32_t t0, uint32_t t1) {
    for (uint32_t t = t0; t < t1; ++t)
      std::copy(lists[t].begin(), lists[t].end(), tcols.begin() + tloc[t].x);
  });
}

// A value a kSliceNz wave stores: anything but +0.0 and -0.0 (bits, so that
// denormals count as nonzero whatever the host's floating-point mode).
static inline bool nonzero(float v) {
  uint32_t b;
  std::memcpy(&b, &v, 4);
  return (b & 0x7FFFFFFFu) != 0u;
}

// Sliced SpMV layout (variant 4, layout in hbj_device.hpp) from the tile-local
// tables.  False, with the outputs cleared, when the system does not
// qualify: a tile without a column list, a row with more than kMaxRuns runs
// or more than kSlicedMaxLen entries, or padding that would take the values past
// 1.25 nnz + 1,024 per tile.  With allow_nz, a wave whose rows have at most
// kNzMaxLen entries stores only their nonzero values (kSliceNz) when that
// takes at most 80% of the bytes of all of them.
bool build_sliced(uint32_t n, const uint32_t* row_ptr, const float* val,
                  const std::vector<uint16_t>& lidx, const std::vector<uint2>& tloc,
                  bool allow_nz, std::vector<float>& sval, std::vector<uint32_t>& runs,
                  std::vector<SliceHdr>& shdr) {
  const uint32_t nt = (n + kTile - 1) / kTile, ns = nt * (uint32_t)kWaves;
  sval.clear();
  runs.clear();
  shdr.assign(ns, SliceHdr{0u, 0u, 0u, 0u});
  for (uint32_t t = 0; t < nt; ++t)
    if (tloc[t].y == kLocalOff) return false;
  // a row's runs: maximal stretches of entries whose list positions step by
  // one, packed (end << 16) | (4 delta as 16 bits); |4 delta| < 2^15 since
  // rows have at most kSlicedMaxLen entries and lists at most kLocalMax
  auto row_runs = [&](uint32_t r, uint32_t* out) -> uint32_t {
    const uint32_t e0 = row_ptr[r], e1 = row_ptr[r + 1];
    uint32_t k = 0;
    for (uint32_t e = e0; e < e1; ++e) {
      const uint32_t pos = e - e0;
      if (e == e0 || lidx[e] != (uint32_t)lidx[e - 1] + 1u) {
        if (k == kMaxRuns) return kMaxRuns + 1u;
        const int32_t dlt = (int32_t)lidx[e] - (int32_t)pos;
        out[k++] = (uint32_t)(uint16_t)(int16_t)(4 * dlt);
      }
      out[k - 1] = (out[k - 1] & 0xFFFFu) | ((pos + 1u) << 16);  // end of the current run
    }
    return k;
  };
  // entries of row r whose value is not +-0 (the ones a kSliceNz wave stores)
  auto stored_count = [&](uint32_t r) -> uint32_t {
    uint32_t c = 0;
    for (uint32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) c += nonzero(val[e]) ? 1u : 0u;
    return c;
  };
  std::vector<uint8_t> bad(nt, 0);
  parallel_ranges(nt, 64, [&](uint32_t t0, uint32_t t1) {
    uint32_t tmp[kMaxRuns], prev[kMaxRuns];
    for (uint32_t t = t0; t < t1; ++t) {
      for (uint32_t w = 0; w < (uint32_t)kWaves; ++w) {
        const uint32_t r0 = t * kTile + 64u * w;
        uint32_t maxlen = 0, maxruns = 0, kprev = 0, maxnz = 0;
        bool group = r0 + 64u <= n;  // a full wave whose groups of 6 share runs
        for (uint32_t l = 0; l < 64u && r0 + l < n; ++l) {
          const uint32_t r = r0 + l, len = row_ptr[r + 1] - row_ptr[r];
          const uint32_t k = row_runs(r, tmp);
          if (len > kSlicedMaxLen || k > kMaxRuns) {
            bad[t] = 1;
            group = false;
            continue;
          }
          if (l > 0 && r % 6u != 0u &&
              (k != kprev || !std::equal(tmp, tmp + k, prev)))
            group = false;
          std::copy(tmp, tmp + k, prev);
          kprev = k;
          maxlen = len > maxlen ? len : maxlen;
          maxruns = k > maxruns ? k : maxruns;
          const uint32_t nzc = stored_count(r);
          maxnz = nzc > maxnz ? nzc : maxnz;
        }
        SliceHdr& h = shdr[t * kWaves + w];
        const uint32_t nq = (maxlen + 3u) / 4u, nq_nz = (maxnz + 3u) / 4u;
        // values only: 1,024 bytes per step of 4, plus 512 bytes of masks
        const bool nz = allow_nz && maxlen > 0 && maxlen <= kNzMaxLen &&
                        5u * (1024u * nq_nz + 512u) <= 4u * 1024u * nq;
        h.nq = nz ? nq_nz : nq;
        h.nr = maxruns | (group ? kRunsPerGroup : 0u) | (nz ? kSliceNz : 0u);
      }
    }
  });
  // groups of 6 a full wave's rows fall in
  auto groups = [](uint32_t r0) { return (r0 + 63u) / 6u - r0 / 6u + 1u; };
  uint64_t vtot = 0, rtot = 0;
  for (uint32_t i = 0; i < ns; ++i) {
    if (bad[i / kWaves]) {
      shdr.clear();
      return false;
    }
    shdr[i].voff = (uint32_t)vtot;
    shdr[i].roff = (uint32_t)rtot;
    vtot += 256ull * shdr[i].nq + ((shdr[i].nr & kSliceNz) ? 128u : 0u);
    const uint32_t nr = shdr[i].nr & ~(kRunsPerGroup | kSliceNz);
    rtot += (shdr[i].nr & kRunsPerGroup) ? (uint64_t)groups(i * 64u) * nr : 64ull * nr;
  }
  const uint64_t nnz = n ? row_ptr[n] : 0u;
  if (vtot > nnz + nnz / 4 + 1024ull * nt || vtot >= (1ull << 30) || rtot >= (1ull << 30)) {
    shdr.clear();
    return false;
  }
  sval.assign(vtot ? vtot : 1u, 0.0f);
  runs.assign(rtot ? rtot : 1u, kRunUnused);
  parallel_ranges(nt, 64, [&](uint32_t t0, uint32_t t1) {
    uint32_t tmp[kMaxRuns];
    for (uint32_t t = t0; t < t1; ++t) {
      for (uint32_t w = 0; w < (uint32_t)kWaves; ++w) {
        const SliceHdr& h = shdr[t * kWaves + w];
        const uint32_t r0 = t * kTile + 64u * w;
        for (uint32_t l = 0; l < 64u && r0 + l < n; ++l) {
          const uint32_t r = r0 + l, e0 = row_ptr[r], len = row_ptr[r + 1] - e0;
          if (h.nr & kSliceNz) {
            uint64_t mask = 0;
            uint32_t j = 0;
            for (uint32_t s = 0; s < len; ++s) {
              if (!nonzero(val[e0 + s])) continue;
              mask |= 1ull << s;
              sval[h.voff + 128u + 4u * (64u * (j / 4u) + l) + j % 4u] = val[e0 + s];
              ++j;
            }
            const uint32_t mw[2] = {(uint32_t)mask, (uint32_t)(mask >> 32)};
            std::memcpy(&sval[h.voff + l], &mw[0], 4);
            std::memcpy(&sval[h.voff + 64u + l], &mw[1], 4);
          } else {
            for (uint32_t s = 0; s < len; ++s)
              sval[h.voff + 4u * (64u * (s / 4u) + l) + s % 4u] = val[e0 + s];
          }
          const uint32_t k = row_runs(r, tmp);
          if (h.nr & kRunsPerGroup) {
            const uint32_t nbs = groups(r0), g = r / 6u - r0 / 6u;
            for (uint32_t j = 0; j < k; ++j) runs[h.roff + nbs * j + g] = tmp[j];
          } else {
            for (uint32_t j = 0; j < k; ++j) runs[h.roff + 64u * j + l] = tmp[j];
          }
        }
      }
    }
  });
  return true;
}

}  // namespace hbj

using namespace hbj;

extern "C" {

const char* hbj_last_error(void) { return g_error.c_str(); }

// compute.rs:37-71
hbj_status hbj_lu_factor_inplace_6(float* mat, uint32_t n) {
  if (!mat) {
    set_error("lu_factor_inplace_6: mat must have len >= 36");
    return HBJ_E_INVALID;
  }
  if (n > 6) {
    set_error("lu_factor_inplace_6: n must be <= 6, got %u", n);
    return HBJ_E_INVALID;
  }
  for (uint32_t k = 0; k < n; ++k) {
    const float pivot = mat[k * 6 + k];
    if (pivot == 0.0f) {
      set_error("LU zero pivot at k = %u", k);
      return HBJ_E_ZERO_PIVOT;
    }
    for (uint32_t i = k + 1; i < n; ++i) mat[i * 6 + k] /= pivot;
    for (uint32_t i = k + 1; i < n; ++i) {
      const float l = mat[i * 6 + k];
      if (l == 0.0f) continue;  // compute.rs:62
      for (uint32_t j = k + 1; j < n; ++j) mat[i * 6 + j] -= l * mat[k * 6 + j];
    }
  }
  return HBJ_OK;
}

// compute.rs:84-181.  Blocks are independent, so large systems are split
// over host threads; the first failing block (lowest index) reports.
hbj_status hbj_build_lu_blocks_6(uint32_t n, const uint32_t* row_ptr, uint32_t row_ptr_len,
                                 const uint32_t* col_idx, uint32_t col_idx_len,
                                 const float* values, uint32_t values_len,
                                 const uint32_t* bs, uint32_t n_starts, float* lu_out) {
  clear_error();
  if (row_ptr_len != n + 1) {
    set_error("build_lu_blocks_from_csr_block_starts_6: row_ptr len must be n+1 (%u), got %u",
              n + 1, row_ptr_len);
    return HBJ_E_INVALID;
  }
  const uint32_t nnz = row_ptr[n];
  if (col_idx_len != nnz || values_len != nnz) {
    set_error("build_lu_blocks_from_csr_block_starts_6: nnz mismatch: row_ptr says %u, col_idx %u, values %u",
              nnz, col_idx_len, values_len);
    return HBJ_E_INVALID;
  }
  if (n_starts < 2) return HBJ_OK;
  if (bs[0] != 0) {
    set_error("build_lu_blocks_from_csr_block_starts_6: block_starts must start at 0");
    return HBJ_E_INVALID;
  }
  if (bs[n_starts - 1] != n) {
    set_error("build_lu_blocks_from_csr_block_starts_6: block_starts last must equal n (%u), got %u",
              n, bs[n_starts - 1]);
    return HBJ_E_INVALID;
  }
  for (uint32_t w = 0; w + 1 < n_starts; ++w) {
    if (bs[w + 1] <= bs[w]) {
      set_error("build_lu_blocks_from_csr_block_starts_6: block_starts must be strictly increasing");
      return HBJ_E_INVALID;
    }
  }
  const uint32_t nb = n_starts - 1;
  std::memset(lu_out, 0, sizeof(float) * 36ull * nb);
  // first block index whose factorisation failed, and where
  auto work = [&](uint32_t b0, uint32_t b1, int64_t* fail_blk, uint32_t* fail_k) {
    for (uint32_t blk = b0; blk < b1; ++blk) {
      const uint32_t off = bs[blk], end = bs[blk + 1];
      if (end > n || off >= end) continue;  // compute.rs:141-143
      const uint32_t m = end - off < 6 ? end - off : 6;
      float mat[36] = {0};
      for (int i = 0; i < 6; ++i) mat[i * 6 + i] = 1.0f;  // identity fill (:151-154)
      for (uint32_t il = 0; il < m; ++il) {
        const uint32_t i = off + il;
        for (uint32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
          const uint32_t j = col_idx[k];
          if (j >= off && j < off + m) mat[il * 6 + (j - off)] = values[k];  // last write wins
        }
      }
      // lu_factor_inplace_6, inlined to keep the failing pivot index
      for (uint32_t k = 0; k < m; ++k) {
        const float pivot = mat[k * 6 + k];
        if (pivot == 0.0f) {
          *fail_blk = blk;
          *fail_k = k;
          return;
        }
        for (uint32_t i = k + 1; i < m; ++i) mat[i * 6 + k] /= pivot;
        for (uint32_t i = k + 1; i < m; ++i) {
          const float l = mat[i * 6 + k];
          if (l == 0.0f) continue;
          for (uint32_t j = k + 1; j < m; ++j) mat[i * 6 + j] -= l * mat[k * 6 + j];
        }
      }
      std::memcpy(lu_out + 36ull * blk, mat, sizeof mat);
    }
  };
  unsigned nt = nb >= 8192 ? std::thread::hardware_concurrency() : 1u;
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  std::vector<int64_t> fb(nt, -1);
  std::vector<uint32_t> fk(nt, 0);
  if (nt == 1) {
    work(0, nb, &fb[0], &fk[0]);
  } else {
    std::vector<std::thread> th;
    const uint32_t per = (nb + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
      const uint32_t b0 = t * per, b1 = (b0 + per) < nb ? b0 + per : nb;
      if (b0 >= b1) continue;
      th.emplace_back(work, b0, b1, &fb[t], &fk[t]);
    }
    for (auto& x : th) x.join();
  }
  for (unsigned t = 0; t < nt; ++t) {
    if (fb[t] >= 0) {
      set_error("LU zero pivot at k = %u", fk[t]);
      return HBJ_E_ZERO_PIVOT;
    }
  }
  return HBJ_OK;
}

hbj_status hbj_spmv_tile_tables(uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                                const uint32_t* col_idx, uint32_t* tile_off, uint32_t* tile_cnt,
                                uint16_t* lidx, uint32_t* tcols, uint64_t tcols_cap,
                                uint64_t* n_tcols) {
  clear_error();
  const uint32_t bs[2] = {0u, n};
  const hbj_status v = validate_case(n, nnz, row_ptr, col_idx, bs, n ? 2u : 1u);
  if (v != HBJ_OK) return v;
  const uint32_t nt = (n + kTile - 1) / kTile;
  if ((nt && (!tile_off || !tile_cnt)) || (nnz && !lidx) || !n_tcols) {
    set_error("hbj_spmv_tile_tables: null output");
    return HBJ_E_INVALID;
  }
  std::vector<uint16_t> lh;
  std::vector<uint2> tl;
  std::vector<uint32_t> tc;
  build_tile_local(n, row_ptr, col_idx, lh, tl, tc);
  uint64_t total = 0;
  for (uint32_t t = 0; t < nt; ++t) total += tl[t].y == kLocalOff ? 0u : tl[t].y;
  *n_tcols = total;
  if (total > tcols_cap || (total && !tcols)) {
    set_error("hbj_spmv_tile_tables: tcols holds %llu columns, %llu needed",
              (unsigned long long)tcols_cap, (unsigned long long)total);
    return HBJ_E_INVALID;
  }
  for (uint32_t t = 0; t < nt; ++t) {
    tile_off[t] = tl[t].x;
    tile_cnt[t] = tl[t].y;
  }
  if (nnz) std::memcpy(lidx, lh.data(), 2ull * nnz);
  if (total) std::memcpy(tcols, tc.data(), 4ull * total);
  return HBJ_OK;
}

hbj_status hbj_sliced_layout_info(uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                                  const uint32_t* col_idx, const float* values, int allow_nz,
                                  hbj_sliced_info* out) {
  clear_error();
  const uint32_t bs[2] = {0u, n};
  const hbj_status v = validate_case(n, nnz, row_ptr, col_idx, bs, n ? 2u : 1u);
  if (v != HBJ_OK) return v;
  if (!out || (nnz && !values)) {
    set_error("hbj_sliced_layout_info: null argument");
    return HBJ_E_INVALID;
  }
  std::vector<uint16_t> lh;
  std::vector<uint2> tl;
  std::vector<uint32_t> tc, runs;
  std::vector<float> sval;
  std::vector<SliceHdr> shdr;
  build_tile_local(n, row_ptr, col_idx, lh, tl, tc);
  std::memset(out, 0, sizeof *out);
  out->sliced = build_sliced(n, row_ptr, values, lh, tl, allow_nz != 0, sval, runs, shdr) ? 1u : 0u;
  if (out->sliced) {
    out->waves = (uint32_t)shdr.size();
    for (const SliceHdr& h : shdr) out->nz_waves += (h.nr & kSliceNz) ? 1u : 0u;
    out->value_words = sval.size();
    out->run_words = runs.size();
  }
  return HBJ_OK;
}

}  // extern "C"
