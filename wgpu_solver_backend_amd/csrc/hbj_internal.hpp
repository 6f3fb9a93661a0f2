// hbj_internal.hpp -- declarations shared by the host-side translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "hbj.h"
#include "hbj_device.hpp"

namespace hbj {

// What one launch of a PCG kernel covers: a descriptor array, a tile table
// and whether to apply the XCD-contiguous tile remap.
struct LaunchSet {
  const SysDesc* descs;
  const uint2* tiles;
  uint32_t n_tiles;
  const uint2* ftiles;    // k_pcg_finish's table (many systems): the entries of `tiles`
  uint32_t n_ftiles;      // whose tile is a multiple of kFinishTiles
  uint32_t remap;
  uint32_t n_sys;
  uint32_t spmv_variant;  // streaming variant of the SpMV kernel (bit-identical results)
  const SysDesc* one;     // host copy of the only system's descriptor, or null (table)
  uint32_t dot64;         // the systems accumulate dots in FP64 (kFlagDot64, all alike)
  uint32_t par;           // iteration parity: which of a system's two SysState copies is read
  uint32_t update_groups; // k_pcg_update_grp: two tiles per workgroup for a single
                          // system, kFinishTiles for a table
  uint32_t st_stride;     // distance between a system's two SysState copies
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;  // kernel timing (optional)
};
// SpMV variants: 0 dword loads with HBM gathers, 1 16-byte loads with HBM
// gathers, 2 / 3 tile-local columns in LDS with 512 / 1024 entries per pass,
// 4 the sliced layout (systems that have one; others as 2).
constexpr uint32_t kSpmvVariantCount = 6;
// Tile partials a kernel can prefetch into registers before it reads a
// system's state (the kernels' level_fetch holds up to this many).
constexpr uint32_t kFetchMaxTiles = 4096;
constexpr uint32_t kSpmvDefault = 4;  // sliced where the system has the layout, else as 2

void launch_pcg_spmv(const LaunchSet& ls, hipStream_t s);
void launch_pcg_update(const LaunchSet& ls, hipStream_t s);
void launch_pcg_finish(const LaunchSet& ls, hipStream_t s);
// the two-launch iteration: the previous set's finish and this set's SpMV
void launch_pcg_fspmv(const LaunchSet& ls, hipStream_t s);
// one-reduction PCG (Chronopoulos-Gear): SpMV with its three partials, update
void launch_ccg_spmv(const LaunchSet& ls, hipStream_t s);
void launch_ccg_update(const LaunchSet& ls, hipStream_t s);
void launch_true_resid(const SysDesc* d, const float* x, double* part2, uint32_t n_tiles,
                       hipStream_t s);

void launch_stream_copy(const void* src, void* dst, uint64_t n4, hipStream_t s);
// LU blocks of a CSR matrix on the device (block_starts already validated);
// *fail (set to ~0 first) receives (first failing block << 32 | pivot index)
void launch_build_lu6(uint32_t n, const uint32_t* row_ptr, const uint32_t* col, const float* val,
                      const uint32_t* bs, uint32_t nb, float* lu, unsigned long long* fail,
                      hipStream_t s);
void launch_op_axpy(uint32_t n, float alpha, const float* x, float* y, hipStream_t s);
void launch_op_scale(uint32_t n, float beta, float* x, hipStream_t s);
void launch_op_dot(uint32_t n, const float* a, const float* b, float* part, float* scratch,
                   float* out, hipStream_t s);
void launch_op_spmv(uint32_t n, const uint32_t* rp, const uint32_t* col, const float* val,
                    const float* x, float* y, hipStream_t s);
void launch_op_block_jacobi(uint32_t n, uint32_t nb, const float* lu, const uint32_t* bs,
                            const float* r, float* z, hipStream_t s);
void launch_op_update_scalars(float* slots, uint32_t a, uint32_t b, uint32_t c, uint32_t d,
                              uint32_t e, uint32_t f, hipStream_t s);

// Tile-local column tables of the SpMV for a validated CSR matrix: per
// 256-row tile the sorted distinct columns of its entries (tcols from
// tloc[t].x, tloc[t].y of them; kLocalOff and no list when there are more
// than kLocalMax) and per entry its index in that list (lidx, nnz + 8 with
// zero padding; 0 for entries of tiles without a list).
void build_tile_local(uint32_t n, const uint32_t* row_ptr, const uint32_t* col,
                      std::vector<uint16_t>& lidx, std::vector<uint2>& tloc,
                      std::vector<uint32_t>& tcols);

// Sliced SpMV layout (variant 4) from the tile-local tables; false when the
// system does not qualify (outputs cleared).  allow_nz: waves may store only
// their nonzero values (kSliceNz).
bool build_sliced(uint32_t n, const uint32_t* row_ptr, const float* val,
                  const std::vector<uint16_t>& lidx, const std::vector<uint2>& tloc,
                  bool allow_nz, std::vector<float>& sval, std::vector<uint32_t>& runs,
                  std::vector<SliceHdr>& shdr);

// The no-alloc dot's device scratch (noalloc.hip): noalloc_init resolves
// its address on `device` (current) once; noalloc_scratch returns it.
hbj_status noalloc_init(int device);
float* noalloc_scratch(int device);

// thread-local error text (hbj_last_error)
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

// CSR + block_starts validation with the reference's messages
// (loaders.rs:84-105, 164-184).
hbj_status validate_case(uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                         const uint32_t* col_idx, const uint32_t* block_starts,
                         uint32_t n_starts);

}  // namespace hbj
