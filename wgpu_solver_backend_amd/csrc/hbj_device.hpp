// hbj_device.hpp -- layouts shared by the HIP kernels and the host runtime.
//
// A "system" is one SPD CSR matrix with its block-Jacobi factors and PCG
// vectors, all resident in HBM.  Kernels never see a system directly: they
// see an array of SysDesc (read-only pointers and sizes) plus a tile table
// that maps every workgroup to (system, 256-row tile).  One system solved
// alone is a table of one system; a batch or the serving loop is a table of
// many.  Per-system mutable scalars (alpha, beta, r.z, status, ...) live in
// SysState on the device, so no host round trip sits inside an iteration.
#pragma once
#include <stdint.h>

namespace hbj {

// Rows per tile.  Equals the reference's dot-product workgroup width
// (dot_partials.wgsl:37 @workgroup_size(256)): each tile's dot partial is the
// reference's partial for the same 256 rows, bit for bit.
constexpr int kTile = 256;
// block_jacobi.wgsl:60-61
constexpr int kBlock = 6;
constexpr int kLuStride = 36;
// CSR entries each lane stages per pass of the SpMV wave loop (FP32 path;
// the FP64 residual path stages half as many 8-byte products).
constexpr int kPer = 8;
constexpr int kWaves = kTile / 64;
// Blocks whose factors a tile stages in LDS at once (a tile meeting more
// stages them in groups).  A 256-row tile meets at most 44 blocks of 6 rows,
// so 48 stages a 6-row system's tile in one group with 6.9 KB per buffer.
constexpr int kLuStage = 48;
// Tiles of one system per workgroup of k_pcg_finish (its reductions are
// redundant per workgroup, so fewer workgroups do less of them).
constexpr uint32_t kFinishTiles = 4;

// SysState.status while a system is still iterating; any other value is a
// final hbj_status.
constexpr int32_t kRunning = -1;
// Most tiles one system may have (the first reduction level must fit LDS).
constexpr uint32_t kMaxTiles = 65536;
// Tile-local SpMV tables: a tile whose entries reference at most kLocalMax
// distinct columns gathers from an LDS copy of those x entries; others
// (TileHdr.cnt == kLocalOff) gather from HBM.
constexpr uint32_t kLocalMax = 2048;
constexpr uint32_t kLocalOff = 0xFFFFFFFFu;

// Per 256-row tile, read by one scalar load at the start of the SpMV: where
// its column list is, and the entry ranges of its four 64-row waves, so the
// entry stream and the list's loads go out together one round trip later
// (no row_ptr round trip in between).
struct alignas(32) TileHdr {
  uint32_t off;    // offset of the tile's column list in tcols
  uint32_t cnt;    // its length, or kLocalOff
  uint32_t eb[5];  // row_ptr at the tile's rows 0, 64, 128, 192 and its end (clamped)
  uint32_t goff;   // 1 + offset of the list's groups in tgrp, or 0: a list made of
                   // whole groups of 6 columns (6g .. 6g+5) is also kept as its
                   // groups g, which the sliced SpMV stages from
};

// Sliced SpMV layout (variant 4) of a system whose every tile has a column
// list and whose rows' entries form at most kMaxRuns runs of consecutive
// list positions (a FEA matrix's dense node blocks): per 64-row wave of a
// tile, the values step-major -- value of step s of the wave's row l at
// sval[voff + 4 (64 (s / 4) + l) + s % 4], nq steps of 4, zero padded -- and
// per row its runs at runs[roff + 64 k + l], k < nr: (end position of run k
// in the row) << 16 | (4 delta as a 16-bit two's complement number), delta
// = list position of the run's first entry minus its position in the row;
// unused runs kRunUnused.  Entry s of a row is list position s + delta(first
// run whose end exceeds s) -- the first run k with (s << 16 | 0xFFFF) <
// runs[k] -- and the row's length is the end of its last run.  A full wave
// whose rows of each group of 6 (rows 6g .. 6g+5, a node's dofs) have
// identical runs stores them once per group (kRunsPerGroup set in nr):
// runs[roff + nbs k + (row0 + l) / 6 - row0 / 6] with nbs = (row0 + 63) / 6 -
// row0 / 6 + 1 groups.
//
// A wave whose rows have at most 64 entries each and many explicit zeros
// (kSliceNz set in nr) stores only its rows' nonzero values: sval[voff + l]
// and sval[voff + 64 + l] hold (as bits) the low and high words of row l's
// mask of stored entries (bit s: entry s of the row is not +-0), and the
// stored values follow step-major from voff + 128, nq steps of 4 counting
// stored entries.  The t-th stored value of a row is its entry at the
// position of the mask's t-th set bit.  Skipping an entry whose value is +-0
// leaves the row sum's bits unchanged while the multiplied vector is finite
// there (the sum starts at +0.0 and is never -0, so adding a +-0 product is
// exact); the SpMV checks the staged vector and restores the NaN such an
// entry would produce when it is not.
constexpr uint32_t kMaxRuns = 8;
constexpr uint32_t kRunUnused = 0xFFFF0000u;
constexpr uint32_t kSlicedMaxLen = 8192;  // entries per row (4 delta fits 16 bits)
constexpr uint32_t kRunsPerGroup = 0x100u;
constexpr uint32_t kSliceNz = 0x200u;
constexpr uint32_t kNzMaxLen = 64;  // entries per row a mask covers
struct SliceHdr {
  uint32_t voff, roff, nq, nr;  // nr: run count | kRunsPerGroup | kSliceNz
};

enum Phase : int32_t {
  kPhInit = 0,      // r0 = b - A x0, z0 = M^-1 r0, p0 = z0, rz_old = r0.z0
  kPhIter = 1,      // PCG iteration
  kPhIrResid = 2,   // IR: x64 += x32 / scale (after an inner solve), then
                    // r64 = b64 - A x64 (FP64), decide, restart the inner PCG
                    // on b32 = r64 * scale at x32 = 0
  // One-reduction PCG (Chronopoulos-Gear), an alternative to the reference's
  // PCG: two launches per iteration, one reduction point
  kPhCcgInit = 3,   // q = A x0, ||b||^2; then r0 = b - q, u0 = M^-1 r0, p = s = 0
  kPhCcg = 4,       // w = A u with (r,u), (w,u), (r,r); then alpha, beta, the
                    // stop test, p, s, x, r and u = M^-1 r
  // The two-launch iteration (the finish of one launch set folded into the
  // next set's SpMV): the states before a launch set's first SpMV
  kPhStart = 5,     // q = A x0 and b.b pending (then kPhInit)
  kPhIrStart = 6,   // IR: the FP64 residual pending (then kPhIrResid)
};

// SysDesc.flags
constexpr uint32_t kFlagOversize = 1u;  // some block has more than 6 rows
constexpr uint32_t kFlagIr = 2u;        // iterative refinement enabled
constexpr uint32_t kFlagDot64 = 4u;     // inner dots accumulated in FP64
constexpr uint32_t kFlagBlocks6 = 8u;   // block b is rows [6b, min(6b + 6, n)): block
                                        // ranges and starts are arithmetic
constexpr uint32_t kFlagSliced = 16u;   // the sliced SpMV layout (variant 4) exists
constexpr uint32_t kFlagTwoLaunch = 32u; // the two-launch iteration (k_pcg_fspmv + k_pcg_update):
                                        // the refinement residual folds the correction on the
                                        // fly into x64b, which k_pcg_update copies to x64
constexpr uint32_t kFlagL1 = 64u;       // a single system of more than 4,096 tiles in the
                                        // three-launch iteration: k_level1 reduces each chunk
                                        // of 256 partials of slots 0-2 into l1 / l164 before
                                        // the kernel that needs their totals

struct alignas(128) SysState {
  // FP32 PCG scalars, the reference's mailbox slots (compute.rs:271-277)
  float rz_old, p_ap, alpha, neg_alpha;
  float rr, rz_new, beta, b_norm2;
  float rel_tol2, abs_tol2;
  int32_t status;      // kRunning or final hbj_status
  int32_t phase;       // Phase
  uint32_t iter;       // iterations of the current (inner) PCG
  uint32_t max_iter;   // per inner PCG
  // iterative refinement
  uint32_t outer, max_outer;
  uint32_t total_iter;
  uint32_t sets;       // launch sets (spmv, update, finish) that found the system running
  double target2;      // (target relative residual)^2
  double rr64, bb64;   // ||r64||^2, ||b64||^2 of the last residual
  double scale;        // b32 = r64 * scale
  double inv_scale;    // x64 += x32 * inv_scale
  double rz64_old, p_ap64;  // FP64-accumulated dots (kFlagDot64)
};

struct SysDesc {
  const uint32_t* row_ptr;   // n+1
  const uint32_t* col;       // nnz
  const float* val;          // nnz
  const uint32_t* bstart;    // nb+1
  const float* lu;           // 36*nb, row-major packed LU per block
  float* x;                  // solution (FP32 PCG) / correction (IR inner)
  float* p;                  // search direction (the two-launch iteration: of even
  float* pn;                 // launch sets; pn: of odd ones -- otherwise pn == p)
  float* q;                  // A p
  float* z;
  float* b;                  // right-hand side of the FP32 PCG
  float* part;               // 3 * n_tiles per-tile dot partials
  double* part64;            // 5 * n_tiles FP64 partials (FP64-accumulated dots; the
                             // FP64 residual's r.r and b.b at 3 and 4 * n_tiles)
  double* x64;               // IR solution
  double* x64b;              // IR, two-launch iteration: x64 with the inner solve's
                             // correction folded in, written by the residual SpMV
  const double* b64;         // IR right-hand side
  double* r64;               // IR residual
  float* r[2];               // r, double-buffered: a launch set of parity par reads
                             // r[par] and writes r[par ^ 1]
  float* s[2];               // one-reduction PCG: s = A p, double-buffered like r
  const uint2* tile_blk;     // per tile: first block touching it, number of such blocks
  const uint16_t* lidx;      // nnz (+8 zero padding): entry's index in its tile's column list
  const TileHdr* thdr;       // per tile: column list and wave entry ranges
  const uint32_t* tcols;     // concatenated sorted distinct column lists of the tiles
  const uint32_t* tgrp;      // the lists made of whole groups of 6, as groups (TileHdr.goff)
  const float* sval;         // sliced layout (kFlagSliced): values,
  const uint32_t* runs;      // per-row runs,
  const SliceHdr* shdr;      // and per wave of each tile its offsets and sizes
  float* l1;                 // kFlagL1: first-level sums of the FP32 partials, slot s
  double* l164;              // chunk g at [s * ceil(n_tiles / 256) + g]; FP64 ones
  SysState* st;
  uint32_t n, nb, n_tiles, flags;
};

}  // namespace hbj
