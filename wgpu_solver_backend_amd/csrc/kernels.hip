// kernels.hip -- gfx950 kernels of the block-Jacobi PCG.
//
// One PCG iteration of every active system is three launches over a tile
// table (workgroup -> (system, 256-row tile)):
//
//   k_pcg_spmv    q = A p (each wave streams its 64 rows' CSR entries
//                 coalesced, products parked in the wave's LDS slice, each
//                 row summed by its owner lane), the tile's p.q partial.
//   k_pcg_update  every workgroup reduces the p.q partials to alpha =
//                 rz_old / p.q, then x += alpha p, r -= alpha q, z = M^-1 r
//                 (6x6 LU solves, factors staged in LDS), r.r and r.z
//                 partials.
//   k_pcg_finish  every workgroup reduces r.r and r.z, computes beta, the
//                 stop test and the state transition (one workgroup writes
//                 it), then p = z + beta p over its rows.
//
// The same three kernels run the init step (r0 = b - A x0 ...) and, for
// iterative refinement, the FP64 residual / restart / fold steps, chosen
// per system by SysState.phase.
// An optional one-reduction PCG (k_ccg_spmv, k_ccg_update) replaces the three
// with two; it is not the reference's operation order.
//
// Arithmetic follows the reference operation for operation (the file is
// compiled with -ffp-contract=off, IEEE division):
//   spmv.wgsl:55-61          q_i = ((0 + v_k0 x_c0) + v_k1 x_c1) + ...  in row order
//   dot_partials.wgsl:46-69  256-slot stride-halving tree per 256 rows
//   dot_reduce.wgsl:47-68    repeated 256-way trees over the partials
//   axpy*.wgsl:41            y + a*x, one rounding each
//   scale + axpy (compute.rs:442-457)  p = (p*beta) + z, two roundings
//   block_jacobi.wgsl:93-122 forward then backward substitution
//   pcg_update_scalars.wgsl:56-67  alpha, beta with zero guards
// so an FP32 solve here matches the CPU restatement bit for bit.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "hbj_device.hpp"
#include "hbj_internal.hpp"

namespace hbj {

// ------------------------------------------------------------------ helpers

// Workgroups are dispatched round-robin over the 8 XCDs.  Remap so that XCD k
// runs a contiguous 1/8 of the tiles: neighbouring rows gather neighbouring x
// entries, which then stay in that XCD's L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t g) {
  const uint32_t per = g / 8u, rem = g % 8u;
  const uint32_t xcd = b % 8u, k = b / 8u;
  return xcd < rem ? xcd * (per + 1u) + k : rem * (per + 1u) + (xcd - rem) * per + k;
}

// Orders LDS writes before LDS reads of other lanes of the same wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane l receives lane l + OFF's 32-bit value, for the lanes a tree step
// keeps (l < OFF): offsets 32 and 16 cross rows of 16 lanes with gfx950's
// v_permlane32_swap / v_permlane16_swap (swapping a value with itself
// leaves lane l + OFF's value in lane l of the second result), offsets 8..1
// shift inside a row with DPP row_shl.  All VALU: no LDS round trip, unlike
// a ds_bpermute shuffle.  Other lanes receive values no tree step uses.
// kLean keeps ds_bpermute for offsets 32 and 16: a swap overwrites both of
// its registers, so the compiler copies the value first, and a kernel whose
// occupancy is set by its register count can lose a wave per SIMD to that.
template <int OFF, bool kLean>
__device__ __forceinline__ uint32_t lane_down_u32(uint32_t v) {
  static_assert(OFF == 32 || OFF == 16 || OFF == 8 || OFF == 4 || OFF == 2 || OFF == 1, "offset");
  if constexpr (kLean && (OFF == 32 || OFF == 16)) return (uint32_t)__shfl_down((int)v, OFF, 64);
  else if constexpr (OFF == 32) return __builtin_amdgcn_permlane32_swap(v, v, false, false)[1];
  else if constexpr (OFF == 16) return __builtin_amdgcn_permlane16_swap(v, v, false, false)[1];
  else return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x100 + OFF, 0xF, 0xF, false);
}
template <int OFF, bool kLean, typename T>
__device__ __forceinline__ T lane_down(T v) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, lane_down_u32<OFF, kLean>(__builtin_bit_cast(uint32_t, v)));
  } else {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = lane_down_u32<OFF, kLean>((uint32_t)u);
    const uint64_t hi = lane_down_u32<OFF, kLean>((uint32_t)(u >> 32));
    return __builtin_bit_cast(T, (hi << 32) | lo);
  }
}
// Offsets 32..1 of the reference's stride-halving tree inside one wave:
// w[l] = w[l] + w[l + off] for l < off.  Result in lane 0.
template <typename T, bool kLean = false>
__device__ __forceinline__ T wave_tail(T w) {
  w = w + lane_down<32, kLean>(w);
  w = w + lane_down<16, kLean>(w);
  w = w + lane_down<8, kLean>(w);
  w = w + lane_down<4, kLean>(w);
  w = w + lane_down<2, kLean>(w);
  w = w + lane_down<1, kLean>(w);
  return w;
}

// The reference's 256-slot tree (dot_partials.wgsl:57-69) computed by one
// wave: lane l holds slots l, l+64, l+128, l+192 as v0..v3.  Offset 128 pairs
// (l, l+128) and (l+64, l+192); offset 64 pairs the two results; offsets
// 32..1 are lane exchanges.  Same pairs, same order.  Result in lane 0.
template <typename T, bool kLean = false>
__device__ __forceinline__ T wave_tree256(T v0, T v1, T v2, T v3) {
  return wave_tail<T, kLean>((v0 + v2) + (v1 + v3));
}

// The same tree over a workgroup of 256 threads (slot t = thread t): offsets
// 128 and 64 through LDS, 32..1 inside wave 0.  Result valid in thread 0.
template <typename T>
__device__ __forceinline__ T tree256(T v, T* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  if (t < 128) sh[t] = sh[t] + sh[t + 128];
  __syncthreads();
  T w = T(0);
  if (t < 64) w = wave_tail(sh[t] + sh[t + 64]);
  __syncthreads();
  return w;
}

// Two independent trees in one pass (same pairing for each component).
template <typename T, bool kLean = false>
__device__ __forceinline__ void tree256x2(T& a, T& b, T* sha, T* shb) {
  const int t = threadIdx.x;
  sha[t] = a;
  shb[t] = b;
  __syncthreads();
  if (t < 128) {
    sha[t] = sha[t] + sha[t + 128];
    shb[t] = shb[t] + shb[t + 128];
  }
  __syncthreads();
  T wa = T(0), wb = T(0);
  if (t < 64) {
    wa = wave_tail<T, kLean>(sha[t] + sha[t + 64]);
    wb = wave_tail<T, kLean>(shb[t] + shb[t + 64]);
  }
  __syncthreads();
  a = wa;
  b = wb;
}

// One level of dot_scalar_exec.rs:204-237: dst[g] = tree of chunk g (256
// values, zero padded) for g < ceil(len/256).  Chunk g is reduced by wave
// g % 4; each wave issues the loads of four of its chunks before reducing
// them.
template <typename T, bool kLean = false>
__device__ __forceinline__ void level_pass(const T* __restrict__ src, uint32_t len, T* dst) {
  constexpr int B = 4;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t out_len = (len + 255u) / 256u;
  for (uint32_t g0 = wave; g0 < out_len; g0 += 4u * B) {
    T v[B][4];
#pragma unroll
    for (int m = 0; m < B; ++m) {
      const uint32_t g = g0 + 4u * m;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t i = g * 256u + lane + 64u * j;
        v[m][j] = (g < out_len && i < len) ? src[i] : T(0);
      }
    }
#pragma unroll
    for (int m = 0; m < B; ++m) {
      const uint32_t g = g0 + 4u * m;
      if (g >= out_len) continue;
      const T r = wave_tree256<T, kLean>(v[m][0], v[m][1], v[m][2], v[m][3]);
      if (lane == 0) dst[g] = r;
    }
  }
}

// The first level of reduce_partials for 1 < len <= kFetchMax partials in
// two steps, so that a kernel can issue the loads (level_fetch) before it
// reads anything else and reduce them afterwards (reduce_fetched).  Same
// chunks and trees as level_pass; the result is reduce_partials' result.
constexpr uint32_t kFetchMax = 4096;
static_assert(kFetchMax == kFetchMaxTiles, "the host's prefetch limit is the kernels'");
template <typename T>
__device__ __forceinline__ void level_fetch(const T* __restrict__ src, uint32_t len, T (&v)[4][4]) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t i = (wave + 4u * m) * 256u + lane + 64u * j;
      v[m][j] = i < len ? src[i] : T(0);
    }
}

template <typename T, bool kLean = false>
__device__ __forceinline__ T reduce_fetched(const T (&v)[4][4], uint32_t len, T* lds) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t out_len = (len + 255u) / 256u;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const uint32_t g = wave + 4u * m;
    if (g < out_len) {
      const T r = wave_tree256<T, kLean>(v[m][0], v[m][1], v[m][2], v[m][3]);
      if (lane == 0) lds[g] = r;
    }
  }
  __syncthreads();
  if (out_len > 1u) {
    if (wave == 0) {
      T w[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t i = lane + 64u * j;
        w[j] = i < out_len ? lds[i] : T(0);
      }
      const T r = wave_tree256<T, kLean>(w[0], w[1], w[2], w[3]);
      if (lane == 0) lds[0] = r;
    }
    __syncthreads();
  }
  return lds[0];
}

// Three reduce_fetched in one pass, with the same trees and so the same
// results: every wave reduces its chunks of all three, one barrier, the
// second levels in waves 0, 1 and 2 side by side, one barrier.
template <typename T, bool kLean = false>
__device__ __forceinline__ void reduce_fetched3(const T (&va)[4][4], const T (&vb)[4][4],
                                                const T (&vc)[4][4], uint32_t len, T* la, T* lb,
                                                T* lc, T& a, T& b, T& c) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t out_len = (len + 255u) / 256u;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const uint32_t g = wave + 4u * m;
    if (g < out_len) {
      const T ra = wave_tree256<T, kLean>(va[m][0], va[m][1], va[m][2], va[m][3]);
      const T rb = wave_tree256<T, kLean>(vb[m][0], vb[m][1], vb[m][2], vb[m][3]);
      const T rc = wave_tree256<T, kLean>(vc[m][0], vc[m][1], vc[m][2], vc[m][3]);
      if (lane == 0) {
        la[g] = ra;
        lb[g] = rb;
        lc[g] = rc;
      }
    }
  }
  __syncthreads();
  if (out_len > 1u) {
    if (wave < 3u) {
      T* l = wave == 0u ? la : wave == 1u ? lb : lc;
      T w[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t i = lane + 64u * j;
        w[j] = i < out_len ? l[i] : T(0);
      }
      const T r = wave_tree256<T, kLean>(w[0], w[1], w[2], w[3]);
      if (lane == 0) l[0] = r;
    }
    __syncthreads();
  }
  a = la[0];
  b = lb[0];
  c = lc[0];
}

// dot_scalar_exec.rs:204-237: while more than one partial is left, reduce
// chunks of 256 with the tree; with exactly one partial there is no pass
// (:210).  Run by all 256 threads of a workgroup.  The first level's results
// go to lds when they fit (at most 256, i.e. up to 65,536 partials -- every
// PCG system) and otherwise to `scratch` (global, ceil(len/256) values),
// which a second level reduces into lds.  Returns the value in every thread.
template <typename T, bool kLean = false>
__device__ __forceinline__ T reduce_partials(const T* __restrict__ part, uint32_t len, T* lds,
                                             T* scratch = nullptr) {
  if (len <= 1u) return part[0];
  uint32_t out_len = (len + 255u) / 256u;
  if (out_len > 256u) {
    level_pass<T, kLean>(part, len, scratch);
    __threadfence_block();
    __syncthreads();
    len = out_len;
    out_len = (len + 255u) / 256u;
    level_pass<T, kLean>(scratch, len, lds);
  } else {
    level_pass<T, kLean>(part, len, lds);
  }
  __syncthreads();
  if (out_len > 1u) {  // out_len <= 256: one more level, one wave
    if ((threadIdx.x >> 6) == 0) {
      const uint32_t lane = threadIdx.x & 63u;
      T v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t i = lane + 64u * j;
        v[j] = i < out_len ? lds[i] : T(0);
      }
      const T r = wave_tree256<T, kLean>(v[0], v[1], v[2], v[3]);
      if (lane == 0) lds[0] = r;
    }
    __syncthreads();
  }
  return lds[0];
}

// The last level of reduce_partials over g first-level sums that k_level1
// already computed (kFlagL1): the same tree, so the same result.  Run by all
// 256 threads of a workgroup.
template <typename T, bool kLean = false>
__device__ __forceinline__ T reduce_last(const T* __restrict__ l1, uint32_t g, T* lds) {
  if (g <= 1u) return l1[0];
  if ((threadIdx.x >> 6) == 0) {
    const uint32_t lane = threadIdx.x & 63u;
    T v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t i = lane + 64u * j;
      v[j] = i < g ? l1[i] : T(0);
    }
    const T r = wave_tree256<T, kLean>(v[0], v[1], v[2], v[3]);
    if (lane == 0) lds[0] = r;
  }
  __syncthreads();
  return lds[0];
}

// Values the code knows to be equal in every lane (a system's state, a
// tile's block range), moved to scalar registers.  The compiler cannot always
// prove it (a load from memory the kernel also writes, an LDS read), and a
// buffer descriptor or branch built from a value it takes as divergent turns
// into a per-lane loop.
__device__ __forceinline__ uint32_t uni(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ int32_t uni(int32_t v) {
  return (int32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
}
// (the builtin returns int: each half goes back through uint32_t so that
// widening to 64 bits cannot sign-extend it)
template <typename T>
__device__ __forceinline__ T* uni(T* p) {
  const uint64_t a = reinterpret_cast<uint64_t>(p);
  const uint64_t lo = uni((uint32_t)a);
  const uint64_t hi = uni((uint32_t)(a >> 32));
  return reinterpret_cast<T*>((hi << 32) | lo);
}
// A struct read from memory, every dword moved to scalar registers.
template <typename T>
__device__ __forceinline__ T uni_copy(const T& src) {
  static_assert(sizeof(T) % 4 == 0, "dword copy");
  T out;
  const uint32_t* in = reinterpret_cast<const uint32_t*>(&src);
  uint32_t* o = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) o[i] = uni(in[i]);
  return out;
}

// The search direction of a launch set of parity par in the two-launch
// iteration: d.p for even sets, d.pn for odd ones (the same buffer in the
// three-launch form).  Both pointers are moved to scalar registers before
// the select: a select between the addresses of the two fields makes the
// compiler place the whole descriptor in scratch memory.
__device__ __forceinline__ float* p_of(const SysDesc& d, uint32_t par) {
  float* const even = uni(d.p);
  float* const odd = uni(d.pn);
  return par ? odd : even;
}

struct TileRef {
  SysDesc d;
  uint32_t tile, row0, nrows;
};

// Arguments of the tile kernels.  A single system travels by value in `one`
// (kOne kernels: tile = workgroup, no table or descriptor load before the
// first data load); otherwise `tiles` maps workgroups to (system, tile) and
// `descs` holds the systems.
//
// Each system has two SysState copies, st and st + st_stride (st_stride =
// the engine's system capacity): the kernels of an iteration with parity
// `par` read copy par, and only k_pcg_finish's state writer writes the next
// iteration's state into copy par ^ 1, so no kernel writes state that
// another workgroup of it reads.
struct TileArgs {
  SysDesc one;
  const SysDesc* descs;
  const uint2* tiles;
  uint32_t n_tiles, remap;
  uint32_t par, st_stride;
};

// This workgroup's table entry (blockIdx.x, or with `remap` its
// XCD-contiguous renumbering); false past the table.
template <bool kOne>
__device__ __forceinline__ bool get_tile(const TileArgs& a, TileRef& tr) {
  const uint32_t b = a.remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
  if (b >= a.n_tiles) return false;
  if constexpr (kOne) {
    tr.d = a.one;
    tr.tile = b;
  } else {
    const uint2 td = a.tiles[b];
    tr.d = uni_copy(a.descs[uni(td.x)]);
    tr.tile = uni(td.y);
  }
  tr.row0 = tr.tile * (uint32_t)kTile;
  const uint32_t left = tr.d.n - tr.row0;
  tr.nrows = left < (uint32_t)kTile ? left : (uint32_t)kTile;
  return true;
}

// ------------------------------------------------------------ SpMV wave body

// CSR bounds of this thread's row row0 + threadIdx.x of a tile (0, 0 past
// the tile's last row).  The wave loops below take them as arguments so a
// kernel can issue these loads before anything else it waits for.
__device__ __forceinline__ void lane_bounds(const uint32_t* __restrict__ row_ptr, uint32_t row0,
                                            uint32_t nrows, uint32_t& mb, uint32_t& me) {
  mb = 0u;
  me = 0u;
  if (threadIdx.x < nrows) {
    mb = row_ptr[row0 + threadIdx.x];
    me = row_ptr[row0 + threadIdx.x + 1u];
  }
}

// spmv.wgsl:52-63 for the 64 rows of one wave of a tile: lane l owns row
// row0 + 64*wave + l (bounds [mb, me), from lane_bounds) and sums its
// products sequentially from 0.0.  The wave streams its rows' CSR entries in
// coalesced chunks (the next chunk's loads are issued before the current
// chunk is summed); each product v_k * x[c_k] is rounded once (the WGSL
// sum + values[k]*x[j] without contraction) and parked in the wave's LDS
// slice.  No workgroup barrier: waves run independently.
//
// All loads here are raw buffer loads: 32-bit offsets (no 64-bit address
// pairs to assemble per lane) and a hardware range check, so a position past
// the wave's last entry reads 0 instead of faulting and every load can be
// unconditional.  Such positions produce products that are never summed.
// (A load under a branch makes the compiler wait for it where the branch
// joins, which would serialise the loads the loops keep in flight.)  Byte
// offsets are 32-bit: vectors hold fewer than 2^30 FP32 or 2^29 FP64
// entries, which the host checks.

// Raw buffer resource over [p, p + bytes) (gfx9 word 3: 32-bit data format,
// stride 0); loads at offsets >= bytes return 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
constexpr int kNt = 2;  // cache policy of the streaming loads: non-temporal

template <typename T>
__device__ __forceinline__ T buf_gather(__amdgpu_buffer_rsrc_t r, uint32_t i);
template <>
__device__ __forceinline__ float buf_gather<float>(__amdgpu_buffer_rsrc_t r, uint32_t i) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(i << 2), 0, 0));
}
template <>
__device__ __forceinline__ double buf_gather<double>(__amdgpu_buffer_rsrc_t r, uint32_t i) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(i << 3), 0, 0));
}

// A prefetch of up to kLightMax partials for the table kernels that
// accumulate in FP64 (the [4][4] prefetch would hold 32 registers per slot):
// thread t loads slot t of chunks 0 and 1 (partials t and 256 + t) with
// range-checked buffer loads (0 past len; no branch, so nothing waits for
// them before they are used), before the kernel reads the state.
// reduce_light then runs reduce_partials' trees over them: each chunk's tree
// as tree256 (slot t = thread t, the pairs of wave_tree256), and for two
// chunks the second level over the chunk sums in one wave -- the same bits.
// For 1 < len <= kLightMax.
constexpr uint32_t kLightMax = 512;
__device__ __forceinline__ void light_fetch(const double* src, uint32_t len, bool on, double (&v)[2]) {
  const __amdgpu_buffer_rsrc_t r = buf_rsrc(src, on ? len * 8u : 0u);
  v[0] = buf_gather<double>(r, threadIdx.x);
  v[1] = buf_gather<double>(r, 256u + threadIdx.x);
}

// reduce_partials over the light_fetch values of one slot (a) or two (a, b:
// both trees in one pass).  lds_a, lds_b: 256 values each; s_c: 4 values.
template <bool kTwo>
__device__ __forceinline__ void reduce_light(const double (&va)[2], const double (&vb)[2],
                                             uint32_t len, double* lds_a, double* lds_b,
                                             double* s_c, double& a, double& b) {
  const uint32_t out_len = (len + 255u) / 256u;  // 1 or 2
#pragma unroll
  for (uint32_t k = 0; k < 2u; ++k) {
    if (k >= out_len) break;
    double x = va[k], y = kTwo ? vb[k] : 0.0;
    if constexpr (kTwo) tree256x2<double>(x, y, lds_a, lds_b);
    else x = tree256<double>(x, lds_a);
    if (threadIdx.x == 0) {
      s_c[k] = x;
      if constexpr (kTwo) s_c[2u + k] = y;
    }
  }
  __syncthreads();
  if (out_len > 1u) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (wave < (kTwo ? 2u : 1u)) {  // wave 0: slot a, wave 1: slot b
      const double* c = s_c + 2u * wave;
      const double w0 = lane < out_len ? c[lane] : 0.0;
      const double r = wave_tree256<double>(w0, 0.0, 0.0, 0.0);
      if (lane == 0) s_c[2u * wave] = r;
    }
    __syncthreads();
  }
  a = s_c[0];
  if constexpr (kTwo) b = s_c[2];
}


typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));


// Element access through a raw buffer resource over a whole array: the base
// stays in scalar registers and a lane supplies one 32-bit byte offset
// instead of a 64-bit address (fewer vector registers; arrays hold fewer
// than 2^30 4-byte or 2^29 8-byte elements, which the host checks).
template <typename T>
__device__ __forceinline__ T bld(const T* base, uint32_t i) {
  const __amdgpu_buffer_rsrc_t r = buf_rsrc(base, 0xFFFFFFFFu);
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(i * 8u), 0, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(i * 4u), 0, 0));
}
template <typename T>
__device__ __forceinline__ void bst(T* base, uint32_t i, T v) {
  const __amdgpu_buffer_rsrc_t r = buf_rsrc(base, 0xFFFFFFFFu);
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)(i * 8u), 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)(i * 4u), 0, 0);
}

// The wave's entry range [wb, we); false when it has no rows or no entries.
// First tile row of this wave's 64 rows in a 256-thread tile (thread t owns
// row t).
__device__ __forceinline__ uint32_t own_wave_row0() { return uni((uint32_t)(threadIdx.x >> 6)) * 64u; }

// Entry range [wb, we) of the 64 rows [wr0, wr0 + 64) of a tile with nrows
// rows; mb/me are lane l's bounds of row wr0 + l.  False if the wave has no
// rows or no entries.
__device__ __forceinline__ bool wave_range(uint32_t mb, uint32_t me, uint32_t nrows, uint32_t wr0,
                                           uint32_t& wb, uint32_t& we) {
  if (wr0 >= nrows) return false;
  const uint32_t wrows = (nrows - wr0) < 64u ? nrows - wr0 : 64u;
  // lane reads into scalar registers: the range is wave-uniform by
  // construction, and buffer descriptors built from it must be too
  wb = uni(mb);
  we = (uint32_t)__builtin_amdgcn_readlane((int)me, (int)wrows - 1);
  return wb < we;
}

// Owner-lane sums of the entries [cs, ce) parked in the slice.
template <typename TA>
__device__ __forceinline__ TA sum_slice(TA sum, const TA* s_prod_wave, uint32_t mb, uint32_t me,
                                        uint32_t cs, uint32_t ce) {
  const uint32_t lo = mb > cs ? mb : cs;
  const uint32_t hi = me < ce ? me : ce;
#pragma unroll 4
  for (uint32_t e = lo; e < hi; ++e) sum = sum + s_prod_wave[e - cs];
  return sum;
}

// dword loads, 64*PER entries per pass (lane l takes entries cs + 64k + l).
// g(c) returns the multiplied vector's entry at column c (one or more buffer
// gathers; every column of the matrix is a valid argument).
template <typename TA, int PER, typename G>
__device__ __forceinline__ TA spmv_wave(uint32_t mb, uint32_t me,
                                        const uint32_t* __restrict__ col,
                                        const float* __restrict__ val, G g, uint32_t nrows,
                                        TA* s_prod_wave) {
  constexpr uint32_t chunk = 64u * PER;
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t wb, we;
  if (!wave_range(mb, me, nrows, own_wave_row0(), wb, we)) return TA(0);
  const __amdgpu_buffer_rsrc_t rc = buf_rsrc(col + wb, (we - wb) * 4u);
  const __amdgpu_buffer_rsrc_t ra = buf_rsrc(val + wb, (we - wb) * 4u);
  uint32_t c[PER];
  float a[PER];
  auto load = [&](uint32_t cs) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int off = (int)((cs - wb + (uint32_t)k * 64u + lane) * 4u);
      c[k] = __builtin_amdgcn_raw_buffer_load_b32(rc, off, 0, kNt);
      a[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(ra, off, 0, kNt));
    }
  };
  load(wb);
  TA sum = TA(0);
  for (uint32_t cs = wb; cs < we; cs += chunk) {
    const uint32_t ce = (we - cs) < chunk ? we : cs + chunk;
    TA pr[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) pr[k] = (TA)a[k] * (TA)g(c[k]);
    load(cs + chunk);  // prefetch the next pass
#pragma unroll
    for (int k = 0; k < PER; ++k) s_prod_wave[k * 64 + lane] = pr[k];
    wave_sync();
    sum = sum_slice(sum, s_prod_wave, mb, me, cs, ce);
    wave_sync();
  }
  return sum;
}

// The same wave loop with 16-byte loads: each lane streams K groups of four
// consecutive (col, value) entries per pass (chunk = 256*K entries, lane l
// group k = entries cs + 4(64k + l) + 0..3).  Passes start at the wave's
// first entry rounded down to a multiple of four, so loads are 16-byte
// aligned; the range of the wave's resources reaches four entries past its
// last one (the CSR arrays carry four entries of zero padding), so a group
// that holds the wave's last entry is read whole.  Products land in LDS in
// entry order, so the owner lanes sum exactly as in spmv_wave.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4v __attribute__((ext_vector_type(4)));

struct GroupSrc {
  __amdgpu_buffer_rsrc_t rc, ra;  // columns and values from entry cs0 on
  uint32_t cs0;
};

__device__ __forceinline__ GroupSrc group_src(const uint32_t* col, const float* val, uint32_t cs0,
                                              uint32_t we) {
  const uint32_t bytes = (we - cs0 + 4u) * 4u;
  return GroupSrc{buf_rsrc(col + cs0, bytes), buf_rsrc(val + cs0, bytes), cs0};
}

template <int K>
__device__ __forceinline__ void load_groups(const GroupSrc& g, uint32_t cs, u32x4 (&c)[K],
                                            f32x4v (&a)[K]) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int off = (int)((cs - g.cs0 + 4u * ((uint32_t)k * 64u + lane)) * 4u);
    c[k] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(g.rc, off, 0, kNt));
    a[k] = __builtin_bit_cast(f32x4v, __builtin_amdgcn_raw_buffer_load_b128(g.ra, off, 0, kNt));
  }
}

template <typename TA, int K>
__device__ __forceinline__ void park_groups(TA* s_prod_wave, const TA (&pr)[K][4]) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int j = 0; j < 4; ++j) s_prod_wave[4u * ((uint32_t)k * 64u + lane) + (uint32_t)j] = pr[k][j];
}

template <typename TA, int K, typename G>
__device__ __forceinline__ TA spmv_wave_vec(uint32_t mb, uint32_t me,
                                            const uint32_t* __restrict__ col,
                                            const float* __restrict__ val, G g, uint32_t nrows,
                                            TA* s_prod_wave) {
  constexpr uint32_t chunk = 256u * K;
  uint32_t wb, we;
  if (!wave_range(mb, me, nrows, own_wave_row0(), wb, we)) return TA(0);
  const GroupSrc gs = group_src(col, val, wb & ~3u, we);
  u32x4 c[K];
  f32x4v a[K];
  load_groups<K>(gs, gs.cs0, c, a);
  TA sum = TA(0);
  for (uint32_t cs = gs.cs0; cs < we; cs += chunk) {
    const uint32_t ce = (we - cs) < chunk ? we : cs + chunk;
    TA pr[K][4];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int j = 0; j < 4; ++j) pr[k][j] = (TA)a[k][j] * (TA)g(c[k][j]);
    load_groups<K>(gs, cs + chunk, c, a);  // prefetch the next pass
    park_groups<TA, K>(s_prod_wave, pr);
    wave_sync();
    sum = sum_slice(sum, s_prod_wave, mb, me, cs, ce);
    wave_sync();
  }
  return sum;
}

// Tile-local form (FP32): the tile's distinct columns are listed at setup
// (d.tcols, sorted) and every entry carries its 16-bit index into that list
// (d.lidx).  stage() copies the multiplied vector at the listed columns into
// LDS (s_v, coalesced: the list is sorted); the entries then gather from
// LDS, and stream 6 bytes each (value + local index) instead of 8.  The
// first pass's loads are issued before stage() so they overlap it.
// stage() holds workgroup barriers and runs in every wave, including waves
// without rows.  Products are a_k * v[c_k] with the same v values, summed
// as in spmv_wave.

// [wb, we): the wave's entries (wave-uniform); mb/me: the lane's row.
template <int K, typename Stage>
__device__ __forceinline__ float spmv_wave_local(uint32_t wb, uint32_t we, uint32_t mb, uint32_t me,
                                                 const uint16_t* __restrict__ lidx,
                                                 const float* __restrict__ val, const float* s_v,
                                                 float* s_prod_wave, Stage stage) {
  constexpr uint32_t chunk = 256u * K;
  const uint32_t lane = threadIdx.x & 63u;
  const bool has = wb < we;
  const uint32_t cs0 = has ? (wb & ~3u) : 0u, wend = has ? we : 0u;
  // ranges reach four entries past the wave's last one (zero padding)
  const __amdgpu_buffer_rsrc_t ri = buf_rsrc(lidx + cs0, (wend - cs0 + 4u) * 2u);
  const __amdgpu_buffer_rsrc_t ra = buf_rsrc(val + cs0, (wend - cs0 + 4u) * 4u);
  struct Pass {
    u32x2 ci[K];  // four 16-bit local indices per group
    f32x4v a[K];
  };
  auto load = [&](Pass& b, uint32_t cs) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t e = cs - cs0 + 4u * ((uint32_t)k * 64u + lane);
      b.ci[k] = __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(ri, (int)(e * 2u), 0, kNt));
      b.a[k] = __builtin_bit_cast(f32x4v, __builtin_amdgcn_raw_buffer_load_b128(ra, (int)(e * 4u), 0, kNt));
    }
  };
  float sum = 0.0f;
  // products of pass b (entries [cs, cs + chunk)), parked and summed by the
  // owner lanes; `next` is loaded in between
  auto run = [&](const Pass& b, uint32_t cs, Pass& next, uint32_t next_cs) {
    const uint32_t ce = (we - cs) < chunk ? we : cs + chunk;
    float pr[K][4];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      pr[k][0] = b.a[k][0] * s_v[b.ci[k][0] & 0xFFFFu];
      pr[k][1] = b.a[k][1] * s_v[b.ci[k][0] >> 16];
      pr[k][2] = b.a[k][2] * s_v[b.ci[k][1] & 0xFFFFu];
      pr[k][3] = b.a[k][3] * s_v[b.ci[k][1] >> 16];
    }
    load(next, next_cs);
    park_groups<float, K>(s_prod_wave, pr);
    wave_sync();
    sum = sum_slice(sum, s_prod_wave, mb, me, cs, ce);
    wave_sync();
  };
  Pass p0;
  load(p0, cs0);
  stage();  // the tile's LDS copy of the vector, while the first pass loads
  if (!has) return 0.0f;
  for (uint32_t cs = cs0; cs < we; cs += chunk) run(p0, cs, p0, cs + chunk);  // prefetch next
  return sum;
}

// What sliced_issue loads for a wave of a sliced tile before the state is
// read: resources, sizes and flags from the slice header, the row's runs,
// and its first quad of values (a) -- its first kPreQuads with kQ = kPreQuads
// (ax holds quads 1 ..).
constexpr int kPreQuads = 4;
struct SlicedPre {
  __amdgpu_buffer_rsrc_t rv, rm;  // values; masks (kSliceNz waves, else empty)
  uint32_t nq, nr, nz;
  uint32_t run[kMaxRuns];
  f32x4v a;
  f32x4v ax[kPreQuads - 1];
};

// Byte offset into s_v of entry st of a row with runs pk (packed as in
// hbj_device.hpp), the search unrolled over the slice's R = nr runs (a
// wave-uniform count): the first run k with (st << 16 | 0xFFFF) < pk[k]
// holds 4 delta in its low 16 bits, and the entry's list position is st +
// delta.  Rows with fewer runs hold kRunUnused in their unused slots.
template <uint32_t R>
__device__ __forceinline__ uint32_t run_off(uint32_t st, const uint32_t (&pk)[kMaxRuns]) {
  const uint32_t key = (st << 16) | 0xFFFFu;
  uint32_t sel = pk[R - 1];
#pragma unroll
  for (int k = (int)R - 2; k >= 0; --k) sel = key < pk[k] ? pk[k] : sel;
  return 4u * st + (uint32_t)(int32_t)(int16_t)(sel & 0xFFFFu);
}

// Quads [q0, q1) of a sliced row (below); entry st is at run_off<R>(st).
// kGuard: some lane may reach its
// row's end inside these quads, so steps from len on (padding) are not
// added; without it every step is below every lane's length.  `a` holds quad
// q0's values on entry and quad q1's on return.  (Loading 2-4 quads ahead
// instead measured slower in the fused SpMV: the registers cost a wave per
// SIMD, 46.4 -> 50.0 us per 50^3 iteration at 3.)
template <uint32_t R, bool kGuard>
__device__ __forceinline__ void sliced_quad(const f32x4v v, uint32_t q,
                                            const uint32_t (&pk)[kMaxRuns], uint32_t len,
                                            const float* s_v, float& sum) {
  const char* s_b = reinterpret_cast<const char*>(s_v);
#pragma unroll
  for (uint32_t i = 0; i < 4u; ++i) {
    const uint32_t st = 4u * q + i;
    const uint32_t off = run_off<R>(st, pk);
    if constexpr (kGuard) {
      const bool in = st < len;
      const float x = *reinterpret_cast<const float*>(s_b + (in ? off : 0u));
      const float pr = v[i] * x;
      sum = in ? sum + pr : sum;
    } else {
      const float x = *reinterpret_cast<const float*>(s_b + off);
      sum = sum + v[i] * x;
    }
  }
}

template <uint32_t R, bool kGuard, int kAux>
__device__ __forceinline__ void sliced_steps(__amdgpu_buffer_rsrc_t rv, uint32_t q0, uint32_t q1,
                                             f32x4v& a, const uint32_t (&pk)[kMaxRuns],
                                             uint32_t len, const float* s_v, uint32_t lane,
                                             float& sum) {
  for (uint32_t q = q0; q < q1; ++q) {
    const f32x4v v = a;
    a = __builtin_bit_cast(f32x4v, __builtin_amdgcn_raw_buffer_load_b128(
                                       rv, (int)(((q + 1u) * 64u + lane) * 16u), 0, kAux));
    sliced_quad<R, kGuard>(v, q, pk, len, s_v, sum);
  }
}

// A row's stream when its first kQ quads were loaded by sliced_issue (kQ =
// kPreQuads: issued before the state, so a short row is in flight in full
// during the workgroup's earlier phases); with kQ = 1 quad q + 1 is loaded
// while quad q is summed.
template <uint32_t R, int kAux, int kQ>
__device__ __forceinline__ float sliced_row(__amdgpu_buffer_rsrc_t rv, uint32_t qa, uint32_t nq,
                                            const SlicedPre& p, const uint32_t (&pk)[kMaxRuns],
                                            uint32_t len, const float* s_v, uint32_t lane) {
  float sum = 0.0f;
  f32x4v a = p.a;
  if constexpr (kQ == 1) {
    sliced_steps<R, false, kAux>(rv, 0u, qa, a, pk, len, s_v, lane, sum);
    sliced_steps<R, true, kAux>(rv, qa, nq, a, pk, len, s_v, lane, sum);
  } else {
    if (nq > 0u) sliced_quad<R, true>(p.a, 0u, pk, len, s_v, sum);
#pragma unroll
    for (uint32_t q = 1; q < (uint32_t)kQ; ++q)
      if (nq > q) sliced_quad<R, true>(p.ax[q - 1], q, pk, len, s_v, sum);
    if (nq > (uint32_t)kQ) {
      a = __builtin_bit_cast(f32x4v, __builtin_amdgcn_raw_buffer_load_b128(
                                         rv, (int)(((uint32_t)kQ * 64u + lane) * 16u), 0, kAux));
      sliced_steps<R, true, kAux>(rv, (uint32_t)kQ, nq, a, pk, len, s_v, lane, sum);
    }
  }
  return sum;
}

// Quads [q0, q1) of the stored values of a kSliceNz row: stored value t is
// the row's entry st, the lowest set bit left in its mask (lo: entries
// 0-31, hi: 32-63), whose list position comes from the runs as above.  The
// products of the stored entries are added in entry order from 0.0, as the
// full row's; its +-0 entries are left out (see nz_poisoned).  kGuard: steps
// from the row's stored count cnt on are padding.
template <uint32_t R, bool kGuard>
__device__ __forceinline__ void nz_quad(const f32x4v v, uint32_t q, const uint32_t (&pk)[kMaxRuns],
                                        uint32_t cnt, uint32_t& lo, uint32_t& hi, const float* s_v,
                                        float& sum) {
  const char* s_b = reinterpret_cast<const char*>(s_v);
#pragma unroll
  for (uint32_t i = 0; i < 4u; ++i) {
    const bool in_lo = lo != 0u;
    const uint32_t w = in_lo ? lo : hi;
    const uint32_t st = (uint32_t)__builtin_ctzg(w, 32) + (in_lo ? 0u : 32u);
    const uint32_t w1 = w & (w - 1u);
    lo = in_lo ? w1 : lo;
    hi = in_lo ? hi : w1;
    const uint32_t off = run_off<R>(st, pk);
    if constexpr (kGuard) {
      const bool in = 4u * q + i < cnt;
      const float x = *reinterpret_cast<const float*>(s_b + (in ? off : 0u));
      const float pr = v[i] * x;
      sum = in ? sum + pr : sum;
    } else {
      const float x = *reinterpret_cast<const float*>(s_b + off);
      sum = sum + v[i] * x;
    }
  }
}

template <uint32_t R, bool kGuard, int kAux>
__device__ __forceinline__ void nz_steps(__amdgpu_buffer_rsrc_t rv, uint32_t q0, uint32_t q1,
                                         f32x4v& a, const uint32_t (&pk)[kMaxRuns], uint32_t cnt,
                                         uint32_t& lo, uint32_t& hi, const float* s_v,
                                         uint32_t lane, float& sum) {
  for (uint32_t q = q0; q < q1; ++q) {
    const f32x4v v = a;
    a = __builtin_bit_cast(f32x4v, __builtin_amdgcn_raw_buffer_load_b128(
                                       rv, (int)(((q + 1u) * 64u + lane) * 16u), 0, kAux));
    nz_quad<R, kGuard>(v, q, pk, cnt, lo, hi, s_v, sum);
  }
}

template <uint32_t R, int kAux, int kQ>
__device__ __forceinline__ float nz_row(__amdgpu_buffer_rsrc_t rv, uint32_t qa, uint32_t nq,
                                        const SlicedPre& p, const uint32_t (&pk)[kMaxRuns],
                                        uint32_t cnt, uint32_t lo, uint32_t hi, const float* s_v,
                                        uint32_t lane) {
  float sum = 0.0f;
  f32x4v a = p.a;
  if constexpr (kQ == 1) {
    nz_steps<R, false, kAux>(rv, 0u, qa, a, pk, cnt, lo, hi, s_v, lane, sum);
    nz_steps<R, true, kAux>(rv, qa, nq, a, pk, cnt, lo, hi, s_v, lane, sum);
  } else {
    if (nq > 0u) nz_quad<R, true>(p.a, 0u, pk, cnt, lo, hi, s_v, sum);
#pragma unroll
    for (uint32_t q = 1; q < (uint32_t)kQ; ++q)
      if (nq > q) nz_quad<R, true>(p.ax[q - 1], q, pk, cnt, lo, hi, s_v, sum);
    if (nq > (uint32_t)kQ) {
      a = __builtin_bit_cast(f32x4v, __builtin_amdgcn_raw_buffer_load_b128(
                                         rv, (int)(((uint32_t)kQ * 64u + lane) * 16u), 0, kAux));
      nz_steps<R, true, kAux>(rv, (uint32_t)kQ, nq, a, pk, cnt, lo, hi, s_v, lane, sum);
    }
  }
  return sum;
}

// Whether a +-0 entry that a kSliceNz row leaves out meets a vector value
// that is not finite: its product, and then the full row's sum, would be
// NaN.  Runs only when the tile's staging found such a value (never on a
// converging solve); the masks are read again rather than kept in registers.
__device__ __forceinline__ bool nz_poisoned(__amdgpu_buffer_rsrc_t rm, const uint32_t (&pk)[kMaxRuns],
                                            uint32_t len, const float* s_v, uint32_t lane) {
  const char* s_b = reinterpret_cast<const char*>(s_v);
  const uint32_t lo = __builtin_amdgcn_raw_buffer_load_b32(rm, (int)(lane * 4u), 0, 0);
  const uint32_t hi = __builtin_amdgcn_raw_buffer_load_b32(rm, (int)((64u + lane) * 4u), 0, 0);
  bool bad = false;
  for (uint32_t st = 0; st < len && st < kNzMaxLen; ++st) {
    if ((((st < 32u ? lo : hi) >> (st & 31u)) & 1u) != 0u) continue;
    const float x = *reinterpret_cast<const float*>(s_b + run_off<kMaxRuns>(st, pk));
    bad = bad || !__builtin_isfinite(x);
  }
  return bad;
}

// Sliced form (variant 4, layout in hbj_device.hpp): lane l sums its own
// row, wave row0 + l, step by step in CSR order, sum = sum + v_s x_c(s) from
// 0.0, each product rounded once -- the reference's row sum.  The values
// stream step-major (16-byte loads, consecutive lanes read consecutive 16
// bytes); an entry's position in the tile's column list comes from the
// row's runs, so no per-entry index is read; x comes from the tile's LDS
// copy s_v.
//
// The matrix's part does not depend on the system's state, so it can be
// issued in two steps around the state's load: sliced_hdr (the wave's slice
// header, a scalar load) and sliced_issue (the row's runs and first
// values); sliced_rows then stages the vector and streams the rest.


// Row l's mask of stored entries in a kSliceNz wave, low and high words (0
// elsewhere: the resource is empty there).  Loaded after the state, with the
// staging: before it, two more registers held across the state's wait cost
// the fused SpMV a wave per SIMD.
template <int kAux>
__device__ __forceinline__ uint2 nz_mask(const SlicedPre& p) {
  const uint32_t lane = threadIdx.x & 63u;
  return make_uint2(__builtin_amdgcn_raw_buffer_load_b32(p.rm, (int)(lane * 4u), 0, kAux),
                    __builtin_amdgcn_raw_buffer_load_b32(p.rm, (int)((64u + lane) * 4u), 0, kAux));
}

__device__ __forceinline__ SliceHdr sliced_hdr(const SysDesc& d, uint32_t tile) {
  const uint32_t w = uni((uint32_t)(threadIdx.x >> 6));
  return uni_copy(d.shdr[tile * (uint32_t)kWaves + w]);
}

// row0: the wave's first row (its runs may be stored once per group of 6).
// kAux: cache policy of the value loads -- non-temporal (variant 4), or the
// default, which keeps a layout that fits the GPU's last-level cache there
// from one SpMV to the next (variant 5).
template <int kAux, int kQ = 1>
__device__ __forceinline__ void sliced_issue(const SysDesc& d, const SliceHdr& h, uint32_t row0,
                                             SlicedPre& p) {
  const uint32_t lane = threadIdx.x & 63u;
  const bool grp = (h.nr & kRunsPerGroup) != 0u;
  const bool nz = (h.nr & kSliceNz) != 0u;
  p.nq = h.nq;
  p.nr = h.nr & ~(kRunsPerGroup | kSliceNz);
  p.nz = nz ? 1u : 0u;
  p.rm = buf_rsrc(d.sval + h.voff, nz ? 512u : 0u);
  p.rv = buf_rsrc(d.sval + h.voff + (nz ? 128u : 0u), h.nq * 1024u);
  const uint32_t g0 = row0 / 6u;
  const uint32_t stride = grp ? (row0 + 63u) / 6u - g0 + 1u : 64u;
  const uint32_t slot = grp ? (row0 + lane) / 6u - g0 : lane;
  const __amdgpu_buffer_rsrc_t rr = buf_rsrc(d.runs + h.roff, p.nr * stride * 4u);
#pragma unroll
  for (uint32_t k = 0; k < kMaxRuns; ++k)
    p.run[k] = __builtin_amdgcn_raw_buffer_load_b32(rr, (int)((k * stride + slot) * 4u), 0, 0);
  p.a = __builtin_bit_cast(f32x4v, __builtin_amdgcn_raw_buffer_load_b128(p.rv, (int)(lane * 16u), 0, kAux));
  // quads past the slice's nq lie outside the resource and read 0
#pragma unroll
  for (int q = 1; q < kQ; ++q)
    p.ax[q - 1] = __builtin_bit_cast(f32x4v, __builtin_amdgcn_raw_buffer_load_b128(
                                                 p.rv, (int)(((uint32_t)q * 64u + lane) * 16u), 0, kAux));
}

// stage() fills s_v (the tile's LDS copy of the vector) while the loads of
// sliced_issue are in flight.
template <int kAux, int kQ = 1, typename Stage>
__device__ __forceinline__ float sliced_rows(const SlicedPre& p, const float* s_v, Stage stage) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint2 m = nz_mask<kAux>(p);
  stage();
  uint32_t pk[kMaxRuns];
  uint32_t len = 0;
#pragma unroll
  for (uint32_t k = 0; k < kMaxRuns; ++k) {
    pk[k] = k < p.nr ? p.run[k] : kRunUnused;
    const uint32_t end = pk[k] >> 16;
    len = (end != 0xFFFFu && end > len) ? end : len;
  }
  // the slice's shortest row: quads below it need no padding guard
  uint32_t ml = len;
  ml = min(ml, lane_down_u32<32, true>(ml));
  ml = min(ml, lane_down_u32<16, true>(ml));
  ml = min(ml, lane_down_u32<8, true>(ml));
  ml = min(ml, lane_down_u32<4, true>(ml));
  ml = min(ml, lane_down_u32<2, true>(ml));
  ml = min(ml, lane_down_u32<1, true>(ml));
  if (p.nz) {
    const uint32_t lo = m.x, hi = m.y;
    const uint32_t cnt = (uint32_t)__builtin_popcount(lo) + (uint32_t)__builtin_popcount(hi);
    uint32_t mc = cnt;
    mc = min(mc, lane_down_u32<32, true>(mc));
    mc = min(mc, lane_down_u32<16, true>(mc));
    mc = min(mc, lane_down_u32<8, true>(mc));
    mc = min(mc, lane_down_u32<4, true>(mc));
    mc = min(mc, lane_down_u32<2, true>(mc));
    mc = min(mc, lane_down_u32<1, true>(mc));
    const uint32_t qz = min(uni(mc) / 4u, p.nq);
    float sum;
    switch (p.nr) {
      case 0:
      case 1: sum = nz_row<1, kAux, kQ>(p.rv, qz, p.nq, p, pk, cnt, lo, hi, s_v, lane); break;
      case 2: sum = nz_row<2, kAux, kQ>(p.rv, qz, p.nq, p, pk, cnt, lo, hi, s_v, lane); break;
      case 3: sum = nz_row<3, kAux, kQ>(p.rv, qz, p.nq, p, pk, cnt, lo, hi, s_v, lane); break;
      case 4: sum = nz_row<4, kAux, kQ>(p.rv, qz, p.nq, p, pk, cnt, lo, hi, s_v, lane); break;
      case 5: sum = nz_row<5, kAux, kQ>(p.rv, qz, p.nq, p, pk, cnt, lo, hi, s_v, lane); break;
      case 6: sum = nz_row<6, kAux, kQ>(p.rv, qz, p.nq, p, pk, cnt, lo, hi, s_v, lane); break;
      case 7: sum = nz_row<7, kAux, kQ>(p.rv, qz, p.nq, p, pk, cnt, lo, hi, s_v, lane); break;
      default: sum = nz_row<kMaxRuns, kAux, kQ>(p.rv, qz, p.nq, p, pk, cnt, lo, hi, s_v, lane);
    }
    // the staging's flags: some staged value is not finite (workgroup-uniform)
    const uint32_t* fl = reinterpret_cast<const uint32_t*>(s_v + kLocalMax);
    if (uni(fl[0] | fl[1] | fl[2] | fl[3]) != 0u && nz_poisoned(p.rm, pk, len, s_v, lane))
      sum = __builtin_nanf("");
    return sum;
  }
  const uint32_t qa = min(uni(ml) / 4u, p.nq);
  switch (p.nr) {
    case 0:
    case 1: return sliced_row<1, kAux, kQ>(p.rv, qa, p.nq, p, pk, len, s_v, lane);
    case 2: return sliced_row<2, kAux, kQ>(p.rv, qa, p.nq, p, pk, len, s_v, lane);
    case 3: return sliced_row<3, kAux, kQ>(p.rv, qa, p.nq, p, pk, len, s_v, lane);
    case 4: return sliced_row<4, kAux, kQ>(p.rv, qa, p.nq, p, pk, len, s_v, lane);
    case 5: return sliced_row<5, kAux, kQ>(p.rv, qa, p.nq, p, pk, len, s_v, lane);
    case 6: return sliced_row<6, kAux, kQ>(p.rv, qa, p.nq, p, pk, len, s_v, lane);
    case 7: return sliced_row<7, kAux, kQ>(p.rv, qa, p.nq, p, pk, len, s_v, lane);
    default: return sliced_row<kMaxRuns, kAux, kQ>(p.rv, qa, p.nq, p, pk, len, s_v, lane);
  }
}

// Each wave's flag after the staging barrier, at s_v[kLocalMax + wave] (the
// LDS copy of the vector is declared with kWaves words beyond kLocalMax): 1
// when a value it staged is not finite (a kSliceNz row's left-out zeros
// then need nz_poisoned), else 0.
__device__ __forceinline__ void flag_staged(bool bad, float* s_v) {
  const bool any = __ballot(bad) != 0ull;
  if ((threadIdx.x & 63u) == 0u)
    reinterpret_cast<uint32_t*>(s_v + kLocalMax)[threadIdx.x >> 6] = any ? 1u : 0u;
}

// s_v[i] = g(tcols[off + i]) for i < m (m <= kLocalMax), then a barrier.
// Run by the whole workgroup.  Loads are unconditional (positions past m
// re-read the list's first entry).
template <typename G>
__device__ __forceinline__ void stage_local(const uint32_t* __restrict__ tcols, uint32_t off,
                                            uint32_t m, G g, float* s_v) {
  constexpr int kSteps = (int)(kLocalMax / kTile);
  const uint32_t t = threadIdx.x;
  bool bad = false;
  if (m > 0u) {
    uint32_t c[kSteps];
#pragma unroll
    for (int k = 0; k < kSteps; ++k) {
      const uint32_t i = t + (uint32_t)k * kTile;
      c[k] = tcols[off + (i < m ? i : 0u)];
    }
    float x[kSteps];
#pragma unroll
    for (int k = 0; k < kSteps; ++k) x[k] = g(c[k]);
#pragma unroll
    for (int k = 0; k < kSteps; ++k) {
      const uint32_t i = t + (uint32_t)k * kTile;
      if (i < m) s_v[i] = x[k];
      bad = bad || (i < m && !__builtin_isfinite(x[k]));
    }
  }
  flag_staged(bad, s_v);
  __syncthreads();
}

// stage_local for a list made of whole groups of 6 columns, read from its
// groups: entry i is column 6 grp[goff + i / 6] + i % 6 (m a multiple of 6).
template <typename G>
__device__ __forceinline__ void stage_groups(const uint32_t* __restrict__ grp, uint32_t goff,
                                             uint32_t m, G g, float* s_v) {
  constexpr int kSteps = (int)(kLocalMax / kTile);
  const uint32_t t = threadIdx.x;
  uint32_t c[kSteps];
#pragma unroll
  for (int k = 0; k < kSteps; ++k) {
    const uint32_t i = t + (uint32_t)k * kTile, ic = i < m ? i : 0u, j = ic / 6u;
    c[k] = 6u * grp[goff + j] + (ic - 6u * j);
  }
  float x[kSteps];
#pragma unroll
  for (int k = 0; k < kSteps; ++k) x[k] = g(c[k]);
  bool bad = false;
#pragma unroll
  for (int k = 0; k < kSteps; ++k) {
    const uint32_t i = t + (uint32_t)k * kTile;
    if (i < m) s_v[i] = x[k];
    bad = bad || (i < m && !__builtin_isfinite(x[k]));
  }
  flag_staged(bad, s_v);
  __syncthreads();
}

// Streaming variants of the FP32 SpMV (all bit-identical; the autotuner
// picks one per system): 0 = dword loads, 512 entries per pass, gathers from
// HBM; 1 = 16-byte loads, 1024, gathers from HBM; 2 = tile-local columns,
// 512; 3 = tile-local, 1024 (2 and 3 use variant 0 on tiles without a local
// column list); 4 = sliced, values read non-temporally; 5 = sliced, values
// read through the caches (4 and 5 run 2 on systems without the sliced
// layout).  kSlice = LDS floats per wave.
template <int V> struct SpmvCfg;
template <> struct SpmvCfg<0> { static constexpr int kSlice = 512; };
template <> struct SpmvCfg<1> { static constexpr int kSlice = 1024; };
template <> struct SpmvCfg<2> { static constexpr int kSlice = 512; };
template <> struct SpmvCfg<3> { static constexpr int kSlice = 1024; };
template <> struct SpmvCfg<4> { static constexpr int kSlice = 512; };  // sliced; else as 2
template <> struct SpmvCfg<5> { static constexpr int kSlice = 512; };  // sliced, cached; else as 2

// --------------------------------------------------- block-Jacobi tile body

// Block range of a tile: from the system's per-tile table, or by binary
// search over block starts when there is none (thread 0, shared via s_blk).
__device__ __forceinline__ uint2 tile_blocks(const uint32_t* __restrict__ bs,
                                             const uint2* __restrict__ tile_blk, uint32_t tile,
                                             uint32_t nb, uint32_t row0, uint32_t nrows,
                                             uint32_t* s_blk) {
  if (tile_blk) return tile_blk[tile];
  if (threadIdx.x == 0) {
    // largest b with bs[b] <= row: first and last block touching the tile
    uint32_t lo = 0, hi = nb;
    while (hi - lo > 1u) {
      const uint32_t mid = (lo + hi) >> 1;
      if (bs[mid] <= row0) lo = mid; else hi = mid;
    }
    const uint32_t bf = lo;
    const uint32_t last_row = row0 + nrows - 1u;
    hi = nb;
    while (hi - lo > 1u) {
      const uint32_t mid = (lo + hi) >> 1;
      if (bs[mid] <= last_row) lo = mid; else hi = mid;
    }
    s_blk[0] = bf;
    s_blk[1] = lo - bf + 1u;
  }
  __syncthreads();
  return make_uint2(s_blk[0], s_blk[1]);
}

// Copies the factors of blocks [bfirst+g, bfirst+g+cnt) (cnt <= kLuStage)
// into s_lu, kLuStride words per block in order, with buffer loads that
// write LDS directly (gfx950 "buffer_load_dword ... lds": lane l of a wave
// writes word M0 + 4 l), so the copy holds no registers and needs no store
// instructions.  Range-checked: positions past the group write 0.  The
// loads complete asynchronously; a reader first calls lds_dma_wait() and
// then the barrier that publishes s_lu.  s_lu holds kLuPer * kTile words.
constexpr int kLuPer = (kLuStage * kLuStride + kTile - 1) / kTile;
__device__ __forceinline__ void bj_stage(const float* __restrict__ lu, uint32_t bfirst,
                                         uint32_t g, uint32_t cnt, float* s_lu) {
  const uint32_t len = cnt * kLuStride;
  const __amdgpu_buffer_rsrc_t r = buf_rsrc(lu + (size_t)(bfirst + g) * kLuStride, len * 4u);
  const uint32_t wave = uni((uint32_t)(threadIdx.x >> 6));
#pragma unroll
  for (int k = 0; k < kLuPer; ++k) {
    const uint32_t i = threadIdx.x + (uint32_t)k * kTile;
    auto* dst = (__attribute__((address_space(3))) void*)(s_lu + (uint32_t)k * kTile + wave * 64u);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, dst, 4, i * 4u, 0, 0, 0);
  }
}
// Waits for this thread's outstanding vector-memory loads, the LDS-writing
// ones included (s_waitcnt vmcnt(0)).
__device__ __forceinline__ void lds_dma_wait() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Rows on each side of a tile whose r a straddling block's solve reads: the
// leading (at most 6) rows of a block that meets the tile lie within 5 rows
// of it.
constexpr uint32_t kHalo = 8;

// z for the rows [row0, row0+nrows) of one tile (block_jacobi.wgsl:63-127).
// Thread t solves block blk.x+t (blk.y blocks meet the tile; [bs_s, bs_e)
// are that block's start and end, loaded by the caller).  Factors are
// staged through LDS in groups of kLuStage blocks, the first group by the
// caller (bj_stage(..., 0, ...) and lds_dma_wait() before the barrier that
// publishes s_rh).
// s_rh[row + kHalo - row0] holds r for rows [row0-kHalo, row0+nrows+kHalo):
// the tile's own rows and the halo rows of straddling blocks (recomputed by
// the owner's formula, bit-identical to the owner's value).
__device__ __forceinline__ void bj_tile(const float* __restrict__ lu, uint2 blk, uint32_t n,
                                        uint32_t row0, uint32_t nrows, uint32_t bs_s,
                                        uint32_t bs_e, const float* s_rh, float* s_z, float* s_lu) {
  const uint32_t t = threadIdx.x;
  const uint32_t bfirst = blk.x, nblk = blk.y;
  for (uint32_t g = 0; g < nblk; g += kLuStage) {
    const uint32_t cnt = (nblk - g) < (uint32_t)kLuStage ? nblk - g : (uint32_t)kLuStage;
    if (g > 0) {
      bj_stage(lu, bfirst, g, cnt, s_lu);
      lds_dma_wait();
      __syncthreads();
    }
    const uint32_t j = t - g;
    if (t >= g && j < cnt) {
      const uint32_t s = bs_s, e = bs_e;
      if (s < n && e > s) {  // block_jacobi.wgsl:74-76
        const uint32_t m = (e - s) < (uint32_t)kBlock ? e - s : (uint32_t)kBlock;
        if (s + m > row0) {
          // All 36 factors and 6 right-hand sides are read unconditionally
          // (the LDS slots exist for any m; kHalo >= kBlock - 1 keeps R in
          // s_rh), so the reads issue together instead of one guarded read
          // at a time.  Rows i >= m compute values no row < m uses: the
          // forward pass only reads y[jj < i], and the backward pass keeps
          // the running sum for jj >= m.  Every row < m performs the same
          // operations in the same order as the plain guarded loops.
          const float* Ls = s_lu + j * (uint32_t)kLuStride;
          const float* R = s_rh + (s + kHalo - row0);
          float L[kBlock * kBlock], y[kBlock], xx[kBlock];
#pragma unroll
          for (int k = 0; k < kBlock * kBlock; ++k) L[k] = Ls[k];
#pragma unroll
          for (int i = 0; i < kBlock; ++i) {
            float sum = R[i];
#pragma unroll
            for (int jj = 0; jj < i; ++jj) sum = sum - L[i * 6 + jj] * y[jj];
            y[i] = sum;
          }
#pragma unroll
          for (int i = kBlock - 1; i >= 0; --i) {
            float sum = y[i];
#pragma unroll
            for (int jj = i + 1; jj < kBlock; ++jj) {
              const float u = sum - L[i * 6 + jj] * xx[jj];
              sum = (uint32_t)jj < m ? u : sum;
            }
            xx[i] = sum / L[i * 6 + i];
          }
#pragma unroll
          for (int i = 0; i < kBlock; ++i) {
            const uint32_t row = s + i;
            if ((uint32_t)i < m && row >= row0 && row < row0 + nrows) s_z[row - row0] = xx[i];
          }
        }
      }
    }
    __syncthreads();
  }
}

// Halo row handled by thread t < 2*kHalo (rows just before and just after the
// tile); false when there is none.
__device__ __forceinline__ bool halo_row(uint32_t row0, uint32_t nrows, uint32_t n, uint32_t& hrow,
                                         uint32_t& slot) {
  const uint32_t t = threadIdx.x;
  if (t < kHalo) {
    if (row0 + t < kHalo) return false;
    hrow = row0 + t - kHalo;
    slot = t;
    return true;
  }
  if (t < 2u * kHalo) {
    hrow = row0 + nrows + (t - kHalo);
    slot = nrows + t;
    return hrow < n;
  }
  return false;
}

// ------------------------------------------------------------ PCG kernels
//
// Per iteration three launches:
//   k_pcg_spmv     (tiles)   q = A p and the tile's p.Ap partial [init:
//                            q = A x0 and b.b; refinement: the FP64 residual]
//   k_pcg_update   (tiles)   every workgroup reduces the p.Ap partials itself
//                            (the same tree, so alpha is bit-identical in all
//                            of them), then x += alpha p, r -= alpha q,
//                            z = M^-1 r and the tile's r.r and r.z partials
//                            (k_pcg_update_grp: the same over two consecutive
//                            tiles per workgroup, four in a table)
//   k_pcg_finish   (tiles)   every workgroup reduces r.z (the state writer and
//                            refinement also r.r), then beta, breakdown
//                            checks, stop test and the state transition in
//                            the reference's order; then p = z + beta p over
//                            four tiles [refinement: the fold]
// Tiles write plain per-tile partials that later kernels read after the
// kernel boundary; no workgroup waits on or synchronises with another.
// Within a kernel no workgroup writes state another workgroup of it reads:
// the state is double-buffered by iteration parity, every kernel reads copy
// `par`, k_pcg_update's workgroup 0 records alpha, p.Ap and ||b||^2 into it
// (fields no workgroup of k_pcg_update reads), and k_pcg_finish's tile-0
// workgroup writes the next state into copy par ^ 1.
//
// Every kernel issues the loads that do not depend on the system's state
// (row bounds, own-row vector entries, factors, partials) before it reads
// SysState, so a workgroup waits for one memory round trip where the state
// and its data would otherwise be two or three dependent ones.

// per-wave product slices: SpmvCfg<V>::kSlice floats, or 64*kPer/2 doubles
template <int V>
struct SpmvLds {
  static constexpr int kSliceF = SpmvCfg<V>::kSlice > kPer * 64 ? SpmvCfg<V>::kSlice : kPer * 64;
  static constexpr bool kLocal = V >= 2;
  static constexpr int K = V == 3 ? 4 : 2;  // groups per lane and pass (local)
  static_assert(kWaves * kSliceF / 2 >= 2 * kTile, "reduction scratch aliases the slices");
};

// Refinement residual of one tile: r64 = b64 - A x64 in FP64 (each FP32
// entry times an FP64 x rounded once, summed from 0.0 in the row's entry
// order) and the tile's FP64 partials of |r64|^2 and |b64|^2.  ax: the
// thread's row of A x64, from the CSR arrays (spmv_ir_resid) or from the
// sliced layout (spmv_ir_resid_sliced) -- the same products added in the
// same order, so the same bits.
__device__ __forceinline__ void ir_resid_tail(const SysDesc& d, const TileRef& tr, double ax,
                                              double xown, double* s_prod) {
  double* s_ra = s_prod;  // tile reductions reuse the slices after a barrier
  double* s_rb = s_prod + kTile;
  const uint32_t t = threadIdx.x, row = tr.row0 + t;
  const bool own = t < tr.nrows;
  __syncthreads();
  // the two-launch iteration: the row's x64 with the correction folded in
  if (own && (d.flags & kFlagTwoLaunch)) bst(d.x64b, row, xown);
  double rr = 0.0, bb = 0.0;
  if (own) {
    const double bi = d.b64[row];
    const double ri = bi - ax;
    d.r64[row] = ri;
    rr = ri * ri;
    bb = bi * bi;
  }
  tree256x2<double>(rr, bb, s_ra, s_rb);
  if (t == 0) {
    d.part64[3u * d.n_tiles + tr.tile] = rr;
    d.part64[4u * d.n_tiles + tr.tile] = bb;
  }
}

// x64 at column c as the residual sees it: with fold (the two-launch
// iteration after an inner solve), x64 + x32 * inv_scale -- the fold
// k_pcg_finish applies in the three-launch form, the same operations.
struct X64Eff {
  __amdgpu_buffer_rsrc_t r64, r32;
  double inv_scale;
  bool fold;
  __device__ __forceinline__ double operator()(uint32_t c) const {
    const double v = buf_gather<double>(r64, c);
    const float xs = buf_gather<float>(r32, c);
    return fold ? v + (double)xs * inv_scale : v;
  }
};
__device__ __forceinline__ X64Eff x64_eff(const SysDesc& d, bool fold, double inv_scale) {
  return X64Eff{buf_rsrc(d.x64, d.n * 8u), buf_rsrc(d.x, d.n * 4u), inv_scale, fold};
}

template <int V>
__device__ __forceinline__ void spmv_ir_resid(const SysDesc& d, const TileRef& tr, uint32_t mb,
                                              uint32_t me, double* s_prod, bool fold = false,
                                              double inv_scale = 0.0) {
  constexpr int kSliceF = SpmvLds<V>::kSliceF;
  const uint32_t wave = threadIdx.x >> 6;
  const X64Eff xg = x64_eff(d, fold, inv_scale);
  double* slice = s_prod + wave * (kSliceF / 2);
  const double xown = xg(threadIdx.x < tr.nrows ? tr.row0 + threadIdx.x : tr.row0);
  const double ax = spmv_wave<double, kPer / 2>(mb, me, d.col, d.val, xg, tr.nrows, slice);
  ir_resid_tail(d, tr, ax, xown, s_prod);
}

// The refinement residual of a tile with the sliced layout, which needs no
// CSR arrays: lane l walks its row's steps in entry order; step st's value
// is at sval step-major, its column-list position is st + delta of the
// first run ending after st, and the column comes from the tile's list (or
// its groups of 6).  x64 is gathered from HBM: this runs once per
// refinement step, not per PCG iteration.  Lanes without a row (past the
// system, or rows without entries) have no runs and sum nothing.
__device__ __forceinline__ void spmv_ir_resid_sliced(const SysDesc& d, const TileRef& tr,
                                                     const TileHdr& hdr, const SlicedPre& p,
                                                     double* s_prod, bool fold = false,
                                                     double inv_scale = 0.0) {
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t pk[kMaxRuns];
  uint32_t len = 0;
#pragma unroll
  for (uint32_t k = 0; k < kMaxRuns; ++k) {
    pk[k] = k < p.nr ? p.run[k] : kRunUnused;
    const uint32_t end = pk[k] >> 16;
    len = (end != 0xFFFFu && end > len) ? end : len;
  }
  const X64Eff xg = x64_eff(d, fold, inv_scale);
  const double xown = xg(threadIdx.x < tr.nrows ? tr.row0 + threadIdx.x : tr.row0);
  double ax = 0.0;
  const uint2 m = nz_mask<0>(p);
  uint32_t j = 0;  // kSliceNz: stored values so far; the row's +-0 entries are
                   // added as 0.0 (the full row's operations, NaN included)
  for (uint32_t st = 0; st < len; ++st) {
    const bool stored = !p.nz || (((st < 32u ? m.x : m.y) >> (st & 31u)) & 1u) != 0u;
    const uint32_t e = p.nz ? j : st;
    j += stored ? 1u : 0u;
    const float vs = __builtin_bit_cast(
        float, __builtin_amdgcn_raw_buffer_load_b32(p.rv, (int)(((e / 4u) * 64u + lane) * 16u + (e % 4u) * 4u), 0, 0));
    const float v = stored ? vs : 0.0f;
    const uint32_t pos = run_off<kMaxRuns>(st, pk) / 4u;
    uint32_t c;
    if (hdr.goff) {
      const uint32_t j = pos / 6u;
      c = 6u * d.tgrp[hdr.goff - 1u + j] + (pos - 6u * j);
    } else {
      c = d.tcols[hdr.off + pos];
    }
    ax = ax + (double)v * xg(c);
  }
  ir_resid_tail(d, tr, ax, xown, s_prod);
}

// q of the tile's rows stored, and the tile's dot partial: dot(b,b) at init
// (compute.rs:235-242), dot(p,Ap) per iteration (:353-360); a = b[row] at
// init, p[row] per iteration (read by the caller).
__device__ __forceinline__ void spmv_tail(const SysDesc& d, const TileRef& tr, float q, bool init,
                                          float a, double* s_prod) {
  double* s_ra = s_prod;
  double* s_rb = s_prod + kTile;
  const uint32_t t = threadIdx.x, row = tr.row0 + t;
  const bool own = t < tr.nrows;
  __syncthreads();  // all waves are done with their slices
  float w = 0.0f;
  double w64 = 0.0;
  if (own) {
    bst(d.q, row, q);
    const float c = init ? a : q;
    w = a * c;
    w64 = (double)a * (double)c;
  }
  if ((d.flags & kFlagDot64) && !init) {
    const double pd = tree256<double>(w64, s_rb);
    if (t == 0) d.part64[tr.tile] = pd;
  } else {
    const float pd = tree256<float>(w, reinterpret_cast<float*>(s_ra));
    if (t == 0) d.part[tr.tile] = pd;
  }
}

// One tile of a system with the sliced layout (variant 4).  The tile and
// slice headers are read with the state, and the rows' runs and first
// values are issued before the state is tested: none of them depends on it.
// A path of its own, so that these registers are not live in the other
// variants' code.
template <int kAux>
__device__ __forceinline__ void spmv_tile_sliced(const TileArgs& ta, const TileRef& tr,
                                                 double* s_prod, float* s_v) {
  const SysDesc& d = tr.d;
  const SysState* st = d.st + ta.par * ta.st_stride;
  const TileHdr hdr = uni_copy(d.thdr[tr.tile]);
  const SliceHdr sh = sliced_hdr(d, tr.tile);
  const int32_t status = uni(st->status), phase = uni(st->phase);
  SlicedPre pre;
  // the three-launch SpMV (systems of more than kFetchMax tiles, whose
  // layouts stream from HBM) issues a row's first kPreQuads quads at once
  sliced_issue<kAux, kPreQuads>(d, sh, tr.row0 + 64u * uni((uint32_t)(threadIdx.x >> 6)), pre);
  if (status != kRunning) return;
  if (phase == kPhIrResid) {
    spmv_ir_resid_sliced(d, tr, hdr, pre, s_prod);
    return;
  }
  if (phase != kPhInit && phase != kPhIter) return;
  const bool init = phase == kPhInit;
  // the multiplied vector: x0 at init (compute.rs:294), p per iteration
  const __amdgpu_buffer_rsrc_t rv = buf_rsrc(init ? d.x : d.p, d.n * 4u);
  auto vec_at = [&](uint32_t c) -> float { return buf_gather<float>(rv, c); };
  // the row's own b or p for the dot, issued before the stream
  const uint32_t rowc = threadIdx.x < tr.nrows ? tr.row0 + threadIdx.x : tr.row0;
  const float a = init ? bld(d.b, rowc) : vec_at(rowc);
  const float q = sliced_rows<kAux, kPreQuads>(pre, s_v, [&]() {
    if (hdr.goff) stage_groups(d.tgrp, hdr.goff - 1u, hdr.cnt, vec_at, s_v);
    else stage_local(d.tcols, hdr.off, hdr.cnt, vec_at, s_v);
  });
  spmv_tail(d, tr, q, init, a, s_prod);
}

// One tile of k_pcg_spmv.
template <int V>
__device__ __forceinline__ void spmv_tile(const TileArgs& ta, const TileRef& tr, double* s_prod,
                                          float* s_v) {
  constexpr int kSliceF = SpmvLds<V>::kSliceF;
  constexpr bool kLocal = SpmvLds<V>::kLocal;
  const SysDesc& d = tr.d;
  if constexpr (V >= 4) {
    if (d.flags & kFlagSliced) {
      spmv_tile_sliced<V == 4 ? kNt : 0>(ta, tr, s_prod, s_v);
      return;
    }
  }
  const SysState* st = d.st + ta.par * ta.st_stride;
  const uint32_t t = threadIdx.x, wave = t >> 6;
  // the tile header (tile-local variants) goes out with the state's loads
  TileHdr hdr{};
  if constexpr (kLocal) hdr = uni_copy(d.thdr[tr.tile]);
  uint32_t mb, me;
  lane_bounds(d.row_ptr, tr.row0, tr.nrows, mb, me);
  const int32_t status = uni(st->status), phase = uni(st->phase);
  if (status != kRunning) return;

  if (phase == kPhIrResid) {
    spmv_ir_resid<V>(d, tr, mb, me, s_prod);
    return;
  }
  if (phase != kPhInit && phase != kPhIter) return;
  // (the refinement residual above returns on its own: joining it with this
  // path before the entry stream starts made the compiler wait for every
  // outstanding load there, the row bounds included)
  const bool init = phase == kPhInit;
  // the multiplied vector: x0 at init (compute.rs:294), p per iteration
  const __amdgpu_buffer_rsrc_t rv = buf_rsrc(init ? d.x : d.p, d.n * 4u);
  auto vec_at = [&](uint32_t c) -> float { return buf_gather<float>(rv, c); };
  float* slice = reinterpret_cast<float*>(s_prod) + wave * kSliceF;
  float q;
  if constexpr (kLocal) {
    if (hdr.cnt != kLocalOff) {
      // the wave's entries from the header (selects, not an index: a
      // runtime index into the struct would place it in scratch memory)
      const uint32_t w = uni(wave);
      const uint32_t wb = w == 0u ? hdr.eb[0] : w == 1u ? hdr.eb[1] : w == 2u ? hdr.eb[2] : hdr.eb[3];
      const uint32_t we = w == 0u ? hdr.eb[1] : w == 1u ? hdr.eb[2] : w == 2u ? hdr.eb[3] : hdr.eb[4];
      q = spmv_wave_local<SpmvLds<V>::K>(wb, we, mb, me, d.lidx, d.val, s_v, slice, [&]() {
        stage_local(d.tcols, hdr.off, hdr.cnt, vec_at, s_v);
      });
    } else {
      q = spmv_wave<float, kPer>(mb, me, d.col, d.val, vec_at, tr.nrows, slice);
    }
  } else if constexpr (V == 1) {
    q = spmv_wave_vec<float, 4>(mb, me, d.col, d.val, vec_at, tr.nrows, slice);
  } else {
    q = spmv_wave<float, kPer>(mb, me, d.col, d.val, vec_at, tr.nrows, slice);
  }
  const uint32_t row = tr.row0 + threadIdx.x;
  spmv_tail(d, tr, q, init, threadIdx.x < tr.nrows ? (init ? bld(d.b, row) : vec_at(row)) : 0.0f, s_prod);
}

// One workgroup per table entry.  (A persistent grid in which every
// workgroup loops over the same number of tiles removes the partly occupied
// last round, but the loop keeps lane-index arithmetic live across tiles --
// 48 -> 80 VGPRs for variant 2 -- and each workgroup waits for its tiles'
// staging one after the other; measured slower, 38.3 vs 36.8 us per launch
// on the 50^3 lattice.)
template <int V, bool kOne>
__global__ __launch_bounds__(kTile) void k_pcg_spmv(const TileArgs ta) {
  __shared__ __align__(16) double s_prod[kWaves * SpmvLds<V>::kSliceF / 2];
  __shared__ float s_v[SpmvLds<V>::kLocal ? kLocalMax + kWaves : 1];  // tile-local copy of the vector
  TileRef tr;
  if (!get_tile<kOne>(ta, tr)) return;
  spmv_tile<V>(ta, tr, s_prod, s_v);
}

// The total of partial slot `slot` (0: p.Ap or b.b, 1: r.r, 2: r.z) of
// system d: reduce_partials over its n_tiles partials, or, with kFlagL1, the
// last level over the first-level sums k_level1 stored (the same bits).
template <typename T, bool kLean = false>
__device__ __forceinline__ T slot_total(const SysDesc& d, uint32_t slot, const T* part, T* lds) {
  const uint32_t nt = d.n_tiles;
  if (d.flags & kFlagL1) {
    const uint32_t g = (nt + 255u) / 256u;
    const T* l1 = std::is_same<T, double>::value ? reinterpret_cast<const T*>(d.l164)
                                                 : reinterpret_cast<const T*>(d.l1);
    return reduce_last<T, kLean>(l1 + slot * g, g, lds);
  }
  return reduce_partials<T, kLean>(part, nt, lds);
}

// First level of reduce_partials -- each chunk of 256 partials reduced by
// the tree of level_pass -- for the slots the next kernel of a kFlagL1
// system needs, computed once instead of in every one of its workgroups:
// which 0 (after the SpMV): slot 0 in FP32, and in FP64 as well for a
// system that accumulates its dots in FP64; which 1 (after the update):
// slots 1 and 2 in the system's dot type.  Wave w of workgroup b reduces
// chunk (4 b + w) mod g of job (4 b + w) / g.
__global__ __launch_bounds__(kTile) void k_level1(const SysDesc d, uint32_t which) {
  const uint32_t nt = d.n_tiles, g = (nt + 255u) / 256u;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t w = blockIdx.x * 4u + uni((uint32_t)(threadIdx.x >> 6));
  const bool dot64 = (d.flags & kFlagDot64) != 0u;
  const uint32_t jobs = which == 0u ? (dot64 ? 2u : 1u) : 2u;
  if (w >= jobs * g) return;
  const uint32_t job = w / g, c = w - job * g;
  const uint32_t slot = which == 0u ? 0u : 1u + job;
  const bool f64 = which == 0u ? job == 1u : dot64;
  if (f64) {
    double v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t i = c * 256u + lane + 64u * j;
      v[j] = i < nt ? d.part64[slot * nt + i] : 0.0;
    }
    const double r = wave_tree256<double>(v[0], v[1], v[2], v[3]);
    if (lane == 0) d.l164[slot * g + c] = r;
  } else {
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t i = c * 256u + lane + 64u * j;
      v[j] = i < nt ? d.part[slot * nt + i] : 0.0f;
    }
    const float r = wave_tree256<float>(v[0], v[1], v[2], v[3]);
    if (lane == 0) d.l1[slot * g + c] = r;
  }
}

// The share of the reduction still needed to reach the refinement's
// target that a step's inner solve aims for when that is looser than the
// caller's inner tolerance: the last step of a 1e-8 refinement in steps of
// 1e-3, which starts near 1e-6, then solves to about 8e-3 instead of 1e-3.
// Falling short costs one more residual and a few iterations of a step that
// needs little.  Measured (256-system batch to 1e-8, inner 1e-3): 1,145
// inner iterations per system with the caller's tolerance throughout, 1,074
// at 0.5, 1,048 at 0.8.
#ifndef HBJ_IR_MARGIN
#define HBJ_IR_MARGIN 0.8
#endif
constexpr double kIrMargin = HBJ_IR_MARGIN;

// Refinement decision from the FP64 residual partials (k_pcg_spmv, phase
// kPhIrResid): 0 = converged, 5 = out of outer steps, kRunning = restart the
// inner solve on r64 * (1/||r64||).  Computed identically by every workgroup
// of k_pcg_update and of k_pcg_finish.
template <bool kLean = false>
__device__ __forceinline__ int32_t ir_decide(const SysDesc& d, double target2, uint32_t outer,
                                             uint32_t max_outer, double* lds_a, double* lds_b,
                                             double& rr, double& bb) {
  rr = reduce_partials<double, kLean>(d.part64 + 3u * d.n_tiles, d.n_tiles, lds_a);
  bb = reduce_partials<double, kLean>(d.part64 + 4u * d.n_tiles, d.n_tiles, lds_b);
  if (bb == 0.0 || rr <= target2 * bb) return 0;
  if (outer >= max_outer) return 5;  // HBJ_E_NOT_CONVERGED
  return kRunning;
}

// kDot64: the table's systems accumulate their dots in FP64 (refinement);
// a table never mixes the two.
//
// Loads come in two rounds, each issued in full before any of it is used
// (every load is unconditional: threads without a row read row0): first the
// state with what does not depend on it (x, q, z, tile blocks, and the p.Ap
// partials of an iteration, reduced to alpha right after), then what the
// state points to (r) with p, the block starts and the factors.
template <bool kOne, bool kDot64>
__global__ __launch_bounds__(kTile) void k_pcg_update(const TileArgs ta) {
  __shared__ float s_rh[kTile + 2 * kHalo];
  __shared__ float s_z[kTile];
  __shared__ float s_lu[kLuPer * kTile];
  __shared__ double s_ra[kTile];
  __shared__ double s_rb[kTile];
  __shared__ uint32_t s_blk[2];
  using TP = typename std::conditional<kDot64, double, float>::type;
  TileRef tr;
  if (!get_tile<kOne>(ta, tr)) return;
  const SysDesc& d = tr.d;
  SysState* st = d.st + ta.par * ta.st_stride;
  const uint32_t t = threadIdx.x, row = tr.row0 + t, nt = d.n_tiles;
  const bool own = t < tr.nrows;
  const uint32_t rowc = own ? row : tr.row0;
  uint32_t hrow = 0, hslot = 0;
  const bool halo = halo_row(tr.row0, tr.nrows, d.n, hrow, hslot);
  const uint32_t hrowc = halo ? hrow : tr.row0;
  // first load round
  const TP* part_pap = kDot64 ? reinterpret_cast<const TP*>(d.part64)
                              : reinterpret_cast<const TP*>(d.part);
  const bool pre = nt > 1u && nt <= kFetchMax;
  TP vp[4][4];
  if (pre) level_fetch<TP>(part_pap, nt, vp);
  const float xi = bld(d.x, rowc), qi = bld(d.q, rowc), hq = bld(d.q, hrowc);
  // rows no block writes keep z
  const float zi_old = (d.flags & kFlagOversize) ? bld(d.z, rowc) : 0.0f;
  // r of the previous iteration (the source of an iteration's update), p
  // (a select, not d.r[par]: a runtime index into the by-value descriptor
  // would copy it to scratch memory)
  const float* __restrict__ r_in = ta.par ? d.r[1] : d.r[0];
  float* __restrict__ r_out = ta.par ? d.r[0] : d.r[1];
  const float ri_prev = bld(r_in, rowc), rh_prev = bld(r_in, hrowc);
  const float pi = bld(p_of(d, ta.par), rowc);
  // blocks touching the tile, their row ranges and factors: arithmetic with
  // 6-row blocks, so the factor loads go out now; otherwise after the table
  // lookup below
  const bool b6 = (d.flags & kFlagBlocks6) != 0u;
  uint2 tb = make_uint2(0u, 0u);
  uint32_t bs_s = 0, bs_e = 0;
  if (b6) {
    const uint32_t first = tr.row0 / 6u, last = (tr.row0 + tr.nrows - 1u) / 6u;
    tb = make_uint2(first, last - first + 1u);
    const uint32_t bi = t < tb.y ? tb.x + t : tb.x;
    bs_s = 6u * bi;
    bs_e = 6u * bi + 6u < d.n ? 6u * bi + 6u : d.n;
    bj_stage(d.lu, tb.x, 0, tb.y < (uint32_t)kLuStage ? tb.y : (uint32_t)kLuStage, s_lu);
  }
  const uint2 blk = (!b6 && d.tile_blk) ? d.tile_blk[tr.tile] : make_uint2(0u, 0u);
  const int32_t status = uni(st->status), phase = uni(st->phase);
  const float rz_old = st->rz_old;
  const double rz64_old = st->rz64_old, target2 = st->target2;
  const uint32_t outer = uni(st->outer), max_outer = uni(st->max_outer);
  if (status != kRunning) return;
  if (phase != kPhInit && phase != kPhIter && phase != kPhIrResid) return;
  const bool iter = phase == kPhIter, restart = phase == kPhIrResid;
  // the two-launch iteration: the residual left the folded x64 in x64b
  if (restart && (d.flags & kFlagTwoLaunch) && own) bst(d.x64, row, bld(d.x64b, row));

  // alpha (pcg_update_scalars.wgsl:59-62, 70-71), ||b||^2 at init, or the
  // refinement restart's scale
  float alpha = 0.0f;
  double scale = 0.0;
  if (iter) {
    TP* lds = reinterpret_cast<TP*>(s_ra);
    // lean trees: this kernel's occupancy is set by its register count
    const TP tot =
        pre ? reduce_fetched<TP, true>(vp, nt, lds) : slot_total<TP, true>(d, 0u, part_pap, lds);
    if constexpr (kDot64) {
      alpha = tot != 0.0 ? (float)(rz64_old / tot) : 0.0f;
      if (tr.tile == 0 && t == 0) {
        st->p_ap = (float)tot;
        st->p_ap64 = tot;
        st->alpha = alpha;
      }
    } else {
      alpha = tot != 0.0f ? rz_old / tot : 0.0f;
      if (tr.tile == 0 && t == 0) {
        st->p_ap = tot;
        st->alpha = alpha;
      }
    }
  } else if (!restart) {
    const float bn = slot_total<float, true>(d, 0u, d.part, reinterpret_cast<float*>(s_ra));
    if (tr.tile == 0 && t == 0) st->b_norm2 = bn;
    if (bn == 0.0f) return;  // Ok(0), x untouched (compute.rs:252-254)
  } else {
    double rr, bb;
    if (ir_decide<true>(d, target2, outer, max_outer, s_ra, s_rb, rr, bb) != kRunning) return;
    scale = 1.0 / sqrt(rr);
  }
  __syncthreads();  // s_ra / s_rb are reused below
  const float na = -alpha;

  // second load round, only where the state chooses another r source (b at init;
  // r64 * scale = b32 at a refinement restart) or the blocks are general
  float rs = ri_prev, rhs = rh_prev;
  if (restart) {
    rs = (float)(bld(d.r64, rowc) * scale);
    rhs = (float)(bld(d.r64, hrowc) * scale);
  } else if (!iter) {
    rs = bld(d.b, rowc);
    rhs = bld(d.b, hrowc);
  }
  if (!b6) {
    const uint2 tb0 = d.tile_blk ? blk : tile_blocks(d.bstart, nullptr, tr.tile, d.nb, tr.row0,
                                                     tr.nrows, s_blk);
    tb = make_uint2(uni(tb0.x), uni(tb0.y));
    const uint32_t bi = t < tb.y ? tb.x + t : tb.x;
    bs_s = bld(d.bstart, bi);
    bs_e = bld(d.bstart, bi + 1u);
    bj_stage(d.lu, tb.x, 0, tb.y < (uint32_t)kLuStage ? tb.y : (uint32_t)kLuStage, s_lu);
  }

  // r of this iteration: compute.rs:395-403 (iteration), :294,301-308 (init)
  auto rnew = [&](float r, float q) -> float {
    if (iter) return r + na * q;
    if (!restart) return r + (-1.0f) * q;
    return r;
  };
  const float rn = own ? rnew(rs, qi) : 0.0f;
  if (own) {
    if (iter) {
      bst(d.x, row, xi + alpha * pi);  // compute.rs:386-394
    } else if (restart) {
      bst(d.b, row, rn);
      bst(d.x, row, 0.0f);
    }
    bst(r_out, row, rn);
    s_rh[t + kHalo] = rn;
    s_z[t] = zi_old;
  }
  if (halo) s_rh[hslot] = rnew(rhs, hq);
  lds_dma_wait();  // the factor copy into s_lu
  __syncthreads();
  bj_tile(d.lu, tb, d.n, tr.row0, tr.nrows, bs_s, bs_e, s_rh, s_z, s_lu);
  float zi = 0.0f;
  if (own) {
    zi = s_z[t];
    bst(d.z, row, zi);
  }
  // r.r (compute.rs:406-413) and r.z (:419-426)
  if constexpr (kDot64) {
    double a = own ? (double)rn * (double)rn : 0.0;
    double b = own ? (double)rn * (double)zi : 0.0;
    tree256x2<double, true>(a, b, s_ra, s_rb);
    if (t == 0) {
      d.part64[nt + tr.tile] = a;
      d.part64[2u * nt + tr.tile] = b;
    }
  } else {
    float a = own ? rn * rn : 0.0f;
    float b = own ? rn * zi : 0.0f;
    tree256x2<float, true>(a, b, reinterpret_cast<float*>(s_ra), reinterpret_cast<float*>(s_rb));
    if (t == 0) {
      d.part[nt + tr.tile] = a;
      d.part[2u * nt + tr.tile] = b;
    }
  }
}

// k_pcg_update with kUT consecutive tiles of one system per workgroup: the
// p.Ap partials are reduced once per workgroup instead of once per tile,
// and fewer workgroups are dispatched (a 50^3 system's 1,465 pairs fit in
// one round of the GPU).  Each tile's operations are k_pcg_update's.  A
// tile's loads are issued after the previous tile is done (holding two
// tiles' would cost a wave per SIMD).  kOne: tiles kUT*b + u; a table:
// entries of the engine's group table (tile0 a multiple of kUT).
// Waves per SIMD the compiler must fit: the table update with FP64 dots
// holds 69 registers under a cap of 7 (73 without; 20.6 KB of LDS also
// allows 7 workgroups per CU), the FP32 instances would spill under it.
constexpr int update_grp_waves(bool one, bool dot64) { return !one && dot64 ? 7 : 1; }
template <bool kOne, bool kDot64, uint32_t kUT>
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(update_grp_waves(kOne, kDot64), 8)))
void k_pcg_update_grp(const TileArgs ta) {
  __shared__ float s_rh[kTile + 2 * kHalo];
  __shared__ float s_z[kTile];
  __shared__ float s_lu[kLuPer * kTile];   // the factors of even tiles of the group
  __shared__ float s_lu2[kLuPer * kTile];  // and of odd ones
  __shared__ double s_ra[kTile];
  __shared__ double s_rb[kTile];
  __shared__ double s_c[4];
  __shared__ uint32_t s_blk[2];
  using TP = typename std::conditional<kDot64, double, float>::type;
  TileRef tr;
  if constexpr (kOne) {
    tr.d = ta.one;
    tr.tile = kUT * (ta.remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x);
    if (tr.tile >= tr.d.n_tiles) return;
  } else {
    if (!get_tile<false>(ta, tr)) return;  // tr.tile: the group's first tile
  }
  const SysDesc& d = tr.d;
  const uint32_t nt = d.n_tiles, t = threadIdx.x, tile0 = tr.tile;
  SysState* st = d.st + ta.par * ta.st_stride;
  const float* __restrict__ p_cur = ta.par ? d.pn : d.p;
  const float* __restrict__ r_in = ta.par ? d.r[1] : d.r[0];
  float* __restrict__ r_out = ta.par ? d.r[0] : d.r[1];
  const bool b6 = (d.flags & kFlagBlocks6) != 0u;
  struct Ld {
    uint32_t tile, row0, nrows, row, rowc, hrow, hslot, hrowc, bs_s, bs_e;
    bool own, halo;
    float xi, qi, hq, zi_old, ri_prev, rh_prev, pi;
    uint2 tb, blk;
  };
  // first load round of a tile: everything that does not depend on the state
  auto load = [&](uint32_t tile, Ld& L, float* lu_buf) {
    L.tile = tile;
    L.row0 = tile * (uint32_t)kTile;
    const uint32_t left = d.n - L.row0;
    L.nrows = left < (uint32_t)kTile ? left : (uint32_t)kTile;
    L.own = t < L.nrows;
    L.row = L.row0 + t;
    L.rowc = L.own ? L.row : L.row0;
    L.hrow = 0;
    L.hslot = 0;
    L.halo = halo_row(L.row0, L.nrows, d.n, L.hrow, L.hslot);
    L.hrowc = L.halo ? L.hrow : L.row0;
    L.xi = bld(d.x, L.rowc);
    L.qi = bld(d.q, L.rowc);
    L.hq = bld(d.q, L.hrowc);
    L.zi_old = (d.flags & kFlagOversize) ? bld(d.z, L.rowc) : 0.0f;
    L.ri_prev = bld(r_in, L.rowc);
    L.rh_prev = bld(r_in, L.hrowc);
    L.pi = bld(p_cur, L.rowc);
    L.tb = make_uint2(0u, 0u);
    L.bs_s = L.bs_e = 0;
    if (b6) {
      const uint32_t first = L.row0 / 6u, last = (L.row0 + L.nrows - 1u) / 6u;
      L.tb = make_uint2(first, last - first + 1u);
      const uint32_t bi = t < L.tb.y ? L.tb.x + t : L.tb.x;
      L.bs_s = 6u * bi;
      L.bs_e = 6u * bi + 6u < d.n ? 6u * bi + 6u : d.n;
      bj_stage(d.lu, L.tb.x, 0, L.tb.y < (uint32_t)kLuStage ? L.tb.y : (uint32_t)kLuStage, lu_buf);
    }
    L.blk = (!b6 && d.tile_blk) ? d.tile_blk[tile] : make_uint2(0u, 0u);
  };
  Ld L;
  // FP64 partials are reduced after the state is read: prefetched they
  // would hold 32 more VGPRs and cost the kernel its one-round fit
  const bool pre = !kDot64 && nt > 1u && nt <= kFetchMax;
  const TP* part_pap = kDot64 ? reinterpret_cast<const TP*>(d.part64)
                              : reinterpret_cast<const TP*>(d.part);
  TP vp[4][4];
  if (pre) level_fetch<TP>(part_pap, nt, vp);
  // ... except in a table, for a system of at most kLightMax tiles: two per
  // thread, which cost 4 registers
  constexpr bool kLightT = !kOne && kDot64;
  const bool light = kLightT && nt > 1u && nt <= kLightMax;
  double lvp[2];
  if constexpr (kLightT) light_fetch(d.part64, nt, light, lvp);
  load(tile0, L, s_lu);
  Ld Ln;  // the next tile's first load round
  const int32_t status = uni(st->status), phase = uni(st->phase);
  const float rz_old = st->rz_old;
  const double rz64_old = st->rz64_old, target2 = st->target2;
  const uint32_t outer = uni(st->outer), max_outer = uni(st->max_outer);
  if (status != kRunning) return;
  if (phase != kPhInit && phase != kPhIter && phase != kPhIrResid) return;
  const bool iter = phase == kPhIter, restart = phase == kPhIrResid;
  // the two-launch iteration: the residual left the folded x64 in x64b
  if (restart && (d.flags & kFlagTwoLaunch)) {
    for (uint32_t u = 0; u < kUT; ++u) {
      const uint32_t tile = tile0 + u;  // the workgroup's consecutive tiles
      const uint32_t row = tile * (uint32_t)kTile + t;
      if (tile < nt && row < d.n) bst(d.x64, row, bld(d.x64b, row));
    }
  }
  float alpha = 0.0f;
  double scale = 0.0;
  if (iter) {
    TP* lds = reinterpret_cast<TP*>(s_ra);
    TP tot;
    if (kLightT && light) {
      double unused = 0.0;
      if constexpr (kLightT) reduce_light<false>(lvp, lvp, nt, s_ra, s_rb, s_c, tot, unused);
    } else {
      tot = pre ? reduce_fetched<TP, true>(vp, nt, lds) : slot_total<TP, true>(d, 0u, part_pap, lds);
    }
    if constexpr (kDot64) {
      alpha = tot != 0.0 ? (float)(rz64_old / tot) : 0.0f;
      if (tile0 == 0 && t == 0) {
        st->p_ap = (float)tot;
        st->p_ap64 = tot;
        st->alpha = alpha;
      }
    } else {
      alpha = tot != 0.0f ? rz_old / tot : 0.0f;
      if (tile0 == 0 && t == 0) {
        st->p_ap = tot;
        st->alpha = alpha;
      }
    }
  } else if (!restart) {
    const float bn = slot_total<float, true>(d, 0u, d.part, reinterpret_cast<float*>(s_ra));
    if (tile0 == 0 && t == 0) st->b_norm2 = bn;
    if (bn == 0.0f) return;  // Ok(0), x untouched (compute.rs:252-254)
  } else {
    double rr, bb;
    if (ir_decide<true>(d, target2, outer, max_outer, s_ra, s_rb, rr, bb) != kRunning) return;
    scale = 1.0 / sqrt(rr);
  }
  __syncthreads();  // s_ra / s_rb are reused below
  const float na = -alpha;
  auto rnew = [&](float r, float q) -> float {
    if (iter) return r + na * q;
    if (!restart) return r + (-1.0f) * q;
    return r;
  };
  // The next tile's first load round (and, for 6-row blocks, its factor copy into
  // the other LDS buffer) is issued once this tile's factors have arrived,
  // so it is in flight during this tile's block solves and partials
  // (50^3: 10.9 -> 10.5 us per launch; 256-system batch 556 -> 568 solves/s).
#pragma unroll
  for (uint32_t u = 0; u < kUT; ++u) {
    float* const s_lu_u = (u & 1u) ? s_lu2 : s_lu;
    if (u > 0u) {
      if (tile0 + u >= nt) break;
      L = Ln;
    }
    float rs = L.ri_prev, rhs = L.rh_prev;
    if (restart) {
      rs = (float)(bld(d.r64, L.rowc) * scale);
      rhs = (float)(bld(d.r64, L.hrowc) * scale);
    } else if (!iter) {
      rs = bld(d.b, L.rowc);
      rhs = bld(d.b, L.hrowc);
    }
    uint2 tb = L.tb;
    uint32_t bs_s = L.bs_s, bs_e = L.bs_e;
    if (!b6) {
      const uint2 tb0 = d.tile_blk ? L.blk : tile_blocks(d.bstart, nullptr, L.tile, d.nb, L.row0,
                                                         L.nrows, s_blk);
      tb = make_uint2(uni(tb0.x), uni(tb0.y));
      const uint32_t bi = t < tb.y ? tb.x + t : tb.x;
      bs_s = bld(d.bstart, bi);
      bs_e = bld(d.bstart, bi + 1u);
      bj_stage(d.lu, tb.x, 0, tb.y < (uint32_t)kLuStage ? tb.y : (uint32_t)kLuStage, s_lu_u);
    }
    const float rn = L.own ? rnew(rs, L.qi) : 0.0f;
    if (L.own) {
      if (iter) {
        bst(d.x, L.row, L.xi + alpha * L.pi);  // compute.rs:386-394
      } else if (restart) {
        bst(d.b, L.row, rn);
        bst(d.x, L.row, 0.0f);
      }
      bst(r_out, L.row, rn);
      s_rh[t + kHalo] = rn;
      s_z[t] = L.zi_old;
    }
    if (L.halo) s_rh[L.hslot] = rnew(rhs, L.hq);
    lds_dma_wait();  // the factor copy into s_lu
    __syncthreads();
    if (u + 1u < kUT && tile0 + u + 1u < nt) load(tile0 + u + 1u, Ln, ((u + 1u) & 1u) ? s_lu2 : s_lu);
    bj_tile(d.lu, tb, d.n, L.row0, L.nrows, bs_s, bs_e, s_rh, s_z, s_lu_u);
    float zi = 0.0f;
    if (L.own) {
      zi = s_z[t];
      bst(d.z, L.row, zi);
    }
    if constexpr (kDot64) {
      double a = L.own ? (double)rn * (double)rn : 0.0;
      double b = L.own ? (double)rn * (double)zi : 0.0;
      tree256x2<double, true>(a, b, s_ra, s_rb);
      if (t == 0) {
        d.part64[nt + L.tile] = a;
        d.part64[2u * nt + L.tile] = b;
      }
    } else {
      float a = L.own ? rn * rn : 0.0f;
      float b = L.own ? rn * zi : 0.0f;
      tree256x2<float, true>(a, b, reinterpret_cast<float*>(s_ra), reinterpret_cast<float*>(s_rb));
      if (t == 0) {
        d.part[nt + L.tile] = a;
        d.part[2u * nt + L.tile] = b;
      }
    }
  }
}

// The state transition at the end of an iteration, in the reference's
// order: the late scalar update (pcg_update_scalars.wgsl:64-67), then the
// breakdown checks and the stop test (compute.rs:468-488); at init
// compute.rs:316-330; the refinement bookkeeping.  `ns` holds the
// iteration's state on entry (with alpha, p.Ap and ||b||^2 recorded by
// k_pcg_update) and the next iteration's on return.  rr, rz (rz64): this
// iteration's r.r and r.z; dec, rr64, bb64: the refinement decision (phase
// kPhIrResid; kRunning = restart) and the residual's norms.
__device__ __forceinline__ void pcg_transition(SysState& ns, uint32_t flags, float rr, float rz,
                                               double rz64, int32_t dec, double rr64, double bb64) {
  const int32_t phase = ns.phase;
  ns.sets += 1u;
  if (phase == kPhInit && ns.b_norm2 == 0.0f) {
    ns.status = 0;  // Ok(0), x untouched (compute.rs:252-254)
    return;
  }
  if (phase == kPhIrResid) {
    ns.rr64 = rr64;
    ns.bb64 = bb64;
    if (dec != kRunning) {
      ns.status = dec;
      return;
    }
    const double nr = sqrt(rr64);
    ns.scale = 1.0 / nr;
    ns.inv_scale = nr;
    ns.outer += 1u;
    // the inner tolerance of the step: the caller's, or -- when the target
    // needs less than that -- kIrMargin times the reduction still needed
    // (the residual falls step by step, so this only ever loosens it)
    const double need2 = kIrMargin * kIrMargin * ns.target2 * bb64 / rr64;
    if (need2 > (double)ns.rel_tol2) ns.rel_tol2 = (float)need2;
  }
  if (phase == kPhInit) {
    ns.rz_old = rz;  // compute.rs:316-330
    ns.rz64_old = rz64;
    ns.phase = kPhIter;
    ns.iter = 0;
    if (ns.max_iter == 0) ns.status = 5;  // loop body never runs (compute.rs:336,491)
    return;
  }
  if (phase == kPhIrResid) {
    // restart of the inner solve on b32 = r64 * scale (written by k_pcg_update)
    ns.rz_old = rz;
    ns.rz64_old = rz64;
    ns.b_norm2 = rr;  // ||b32||^2 of the restarted inner solve
    ns.phase = kPhIter;
    ns.iter = 0;
    return;
  }
  const float rz_old = ns.rz_old, p_ap = ns.p_ap;
  float beta = rz_old != 0.0f ? rz / rz_old : 0.0f;
  if (flags & kFlagDot64) beta = ns.rz64_old != 0.0 ? (float)(rz64 / ns.rz64_old) : 0.0f;
  ns.beta = beta;
  ns.rz_new = rz;
  ns.rr = rr;
  const uint32_t it = ns.iter + 1u;
  ns.iter = it;
  ns.total_iter += 1u;
  int32_t fin = kRunning;
  if (p_ap == 0.0f) fin = 3;            // HBJ_E_BREAKDOWN_PAP
  else if (rz_old == 0.0f) fin = 4;     // HBJ_E_BREAKDOWN_RZ
  else if (rr <= ns.abs_tol2 || rr <= ns.rel_tol2 * ns.b_norm2) fin = 0;
  else if (it >= ns.max_iter) fin = 5;  // HBJ_E_NOT_CONVERGED
  if (fin == kRunning) {
    ns.rz_old = rz;
    ns.rz64_old = rz64;
  } else if (flags & kFlagIr) {
    ns.phase = kPhIrResid;  // fold the correction (below), then a new FP64 residual
  } else {
    ns.status = fin;
  }
}


// After k_pcg_update.  Every workgroup reduces r.r and r.z (and, at a
// refinement decision, the residual's norms) from the partials -- the same
// trees, so the same bits -- and computes the state transition (without
// refinement the workgroups other than the state writer reduce r.z alone
// and take beta from it; below); the
// workgroup holding the system's tile 0 writes the next iteration's state
// (a final state into both copies: a finished system may get no more
// workgroups once its tiles are dropped from the table).  Then, over the
// workgroup's rows: p = (p * beta) + z, two roundings
// (scale_from_scalar_results.wgsl:37, axpy.wgsl:41 with alpha = 1), or p = z
// on the first iteration (compute.rs:313-314); after an inner solve of the
// refinement (phase just set to kPhIrResid) the fold x64 += x32 / scale.
// A workgroup covers kFinishTiles tiles of one system (so a system's
// reductions run in a quarter as many workgroups): kOne, tiles b + k *
// gridDim.x; a table (the engine's finish table, one entry per group),
// tiles tile0 + k of the entry's system.
template <bool kOne, bool kDot64>
__global__ __launch_bounds__(kTile) void k_pcg_finish(const TileArgs ta) {
  __shared__ double s_a[256];
  __shared__ double s_b[256];
  __shared__ double s_c[4];
  __shared__ SysState s_ns;
  constexpr uint32_t kWords = sizeof(SysState) / 4;
  static_assert(kWords <= (uint32_t)kTile && sizeof(SysState) % 4 == 0, "state copy by dwords");
  TileRef tr;
  uint32_t step = 1;
  if constexpr (kOne) {
    if (blockIdx.x >= ta.n_tiles) return;
    tr.d = ta.one;
    tr.tile = blockIdx.x;
    step = gridDim.x;
  } else {
    if (!get_tile<false>(ta, tr)) return;  // tr.tile: the group's first tile
  }
  const SysDesc& d = tr.d;
  SysState* in = d.st + ta.par * ta.st_stride;
  SysState* out = d.st + (ta.par ^ 1u) * ta.st_stride;
  const uint32_t t = threadIdx.x, nt = d.n_tiles;
  const bool writer = tr.tile == 0;
  // r.r and r.z partials fetched together with the state
  using TP = typename std::conditional<kDot64, double, float>::type;
  const TP* part = kDot64 ? reinterpret_cast<const TP*>(d.part64) : reinterpret_cast<const TP*>(d.part);
  // a table's FP64 partials are reduced after the state is read (prefetched
  // they hold 64 VGPRs, which costs a table's many rounds of workgroups more
  // than the extra load round; a single system's one round gains from them)
  const bool pre = (kOne || !kDot64) && nt > 1u && nt <= kFetchMax;
  TP va[4][4], vb[4][4];
  // ... except those of a system of at most kLightMax tiles (every system of
  // the batches and most served requests): two per thread, before the state
  constexpr bool kLightT = !kOne && kDot64;
  const bool light = kLightT && nt > 1u && nt <= kLightMax;
  double lva[2], lvb[2];
  if constexpr (kLightT) {
    light_fetch(d.part64 + nt, nt, light, lva);
    light_fetch(d.part64 + 2u * nt, nt, light, lvb);
  }
  // workgroups that will only need beta (below) do not fetch the r.r partials
  const bool maybe_beta_only = !kDot64 && !writer && !(d.flags & kFlagIr);
  const bool pre_a = pre && !maybe_beta_only;
  // the rows' z and p for an iteration's p-update (range-checked loads: rows
  // past the system read 0 and are not stored).  A single system's kernel
  // loads them once beta is known (issuing them before the state measured
  // no faster there, and costs registers); a table's FP64 kernel issues them
  // with its partials (kRowsEarly).
  constexpr uint32_t kT = kFinishTiles;
  float zv[kT], pv[kT];
  auto row_of = [&](uint32_t k) -> uint32_t {  // 0xFFFFFFFF: no row
    const uint32_t tile = tr.tile + k * step;
    const uint32_t row = tile * (uint32_t)kTile + t;
    return (tile < nt && row < d.n) ? row : 0xFFFFFFFFu;
  };
  auto load_rows = [&]() {
    const __amdgpu_buffer_rsrc_t rz = buf_rsrc(d.z, d.n * 4u), rp = buf_rsrc(d.p, d.n * 4u);
#pragma unroll
    for (uint32_t k = 0; k < kT; ++k) {
      const uint32_t row = (tr.tile + k * step) * (uint32_t)kTile + t;
      zv[k] = buf_gather<float>(rz, row);
      pv[k] = buf_gather<float>(rp, row);
    }
  };
  if (pre_a) level_fetch<TP>(part + nt, nt, va);
  if (pre) level_fetch<TP>(part + 2u * nt, nt, vb);
  // a table's FP64 kernel issues them with its partials (kRowsEarly)
  constexpr bool kRowsEarly = kLightT;
  if constexpr (kRowsEarly) load_rows();
  if (t < kWords) reinterpret_cast<uint32_t*>(&s_ns)[t] = reinterpret_cast<const uint32_t*>(in)[t];
  __syncthreads();
  const int32_t status = uni(s_ns.status), phase = uni(s_ns.phase);
  if (status != kRunning) {  // finished: carry the state to the other copy
    if (writer && t < kWords)
      reinterpret_cast<uint32_t*>(out)[t] = reinterpret_cast<const uint32_t*>(&s_ns)[t];
    return;
  }
  int32_t dec = kRunning;
  double rr64 = 0.0, bb64 = 0.0;
  if (phase == kPhIrResid) {
    dec = ir_decide(d, s_ns.target2, s_ns.outer, s_ns.max_outer, s_a, s_b, rr64, bb64);
    __syncthreads();  // s_a / s_b are reused below
  }
  float rr = 0.0f, rz = 0.0f;
  double rz64 = 0.0;
  const bool reduce = !(phase == kPhInit && s_ns.b_norm2 == 0.0f) &&
                      !(phase == kPhIrResid && dec != kRunning);
  // Without refinement only the state writer needs r.r (the stop test);
  // the other workgroups need beta, from r.z alone, and update p even in
  // the iteration that stops the system (harmless: nothing reads p then).
  const bool beta_only = maybe_beta_only && phase == kPhIter;
  if (reduce && beta_only) {
    TP* lb = reinterpret_cast<TP*>(s_b);
    const TP b = pre ? reduce_fetched<TP>(vb, nt, lb) : slot_total<TP>(d, 2u, part + 2u * nt, lb);
    const float rz_old = s_ns.rz_old;
    const float beta = rz_old != 0.0f ? (float)b / rz_old : 0.0f;  // pcg_transition's beta
    if constexpr (!kRowsEarly) load_rows();
#pragma unroll
    for (uint32_t k = 0; k < kT; ++k) {
      const uint32_t row = row_of(k);
      if (row == 0xFFFFFFFFu) continue;
      const float pb = pv[k] * beta;
      bst(d.p, row, pb + zv[k]);
    }
    return;
  }
  if (reduce) {
    TP* la = reinterpret_cast<TP*>(s_a);
    TP* lb = reinterpret_cast<TP*>(s_b);
    TP a, b;
    if (kLightT && light) {
      if constexpr (kLightT) reduce_light<true>(lva, lvb, nt, s_a, s_b, s_c, a, b);
    } else {
      a = pre_a ? reduce_fetched<TP>(va, nt, la) : slot_total<TP>(d, 1u, part + nt, la);
      b = pre ? reduce_fetched<TP>(vb, nt, lb) : slot_total<TP>(d, 2u, part + 2u * nt, lb);
    }
    rr = (float)a;
    rz = (float)b;
    if constexpr (kDot64) rz64 = b;
  }
  if (t == 0) pcg_transition(s_ns, d.flags, rr, rz, rz64, dec, rr64, bb64);
  __syncthreads();
  const int32_t nstatus = uni(s_ns.status), nphase = uni(s_ns.phase);
  const uint32_t niter = uni(s_ns.iter);
  const float beta = s_ns.beta;
  const double inv_scale = s_ns.inv_scale;
  if (writer && t < kWords) {
    const uint32_t w = reinterpret_cast<const uint32_t*>(&s_ns)[t];
    reinterpret_cast<uint32_t*>(out)[t] = w;
    if (nstatus != kRunning) reinterpret_cast<uint32_t*>(in)[t] = w;
  }
  if (nstatus != kRunning) return;
  if (nphase == kPhIter) {
    if constexpr (!kRowsEarly) load_rows();
#pragma unroll
    for (uint32_t k = 0; k < kT; ++k) {
      const uint32_t row = row_of(k);
      if (row == 0xFFFFFFFFu) continue;
      if (niter == 0u) {
        bst(d.p, row, zv[k]);
      } else {
        const float pb = pv[k] * beta;
        bst(d.p, row, pb + zv[k]);
      }
    }
  } else if (nphase == kPhIrResid) {
    for (uint32_t k = 0; k < kT; ++k) {
      const uint32_t tile = tr.tile + k * step;
      const uint32_t row = tile * (uint32_t)kTile + t;
      if (tile < nt && row < d.n)
        bst(d.x64, row, bld(d.x64, row) + (double)bld(d.x, row) * inv_scale);
    }
  }
}

// ----------------------------------------- the two-launch iteration
//
// k_pcg_fspmv folds k_pcg_finish into the next launch set's SpMV, so that
// an iteration is two launches, k_pcg_fspmv then k_pcg_update.  Per tile:
//   1. the previous set's finish, as k_pcg_finish does it: the state (copy
//      par ^ 1, the one the previous k_pcg_update used), the r.z reduction
//      (r.r too in the state writer and with refinement), beta, the
//      breakdown checks, the stop test and the transition; the system's
//      tile-0 workgroup writes the next state into copy par (a final state
//      into both copies);
//   2. this set's SpMV q = A p, with p = (p_prev * beta) + z (two
//      roundings; p = z on the first iteration) evaluated where it is used:
//      at the tile's listed columns when the vector is staged (or per
//      gathered entry), and for the tile's own rows, which the tile stores
//      into p[par].  p is double-buffered by launch-set parity (d.p for even
//      sets, d.pn for odd ones), since other tiles still read p_prev at
//      these rows; every evaluation applies the same operations to the same
//      values, so all copies of p_k carry the same bits.
// kPhStart / kPhIrStart mark a set whose SpMV has no finish before it (the
// first SpMV of a solve, q = A x0; a refinement residual).  When an inner
// solve of the refinement stops, the next set computes the residual from
// x64 + x32 * inv_scale, evaluated per gathered column (the three-launch
// form's fold, the same operations), and stores the tile's folded rows into
// x64b, which the following k_pcg_update copies to x64 (no tile rewrites a
// value another tile of the same launch reads).

// The thread's row of A v for the tile, v given by g(column): the sliced
// layout (its loads issued by the caller), or the tile-local / HBM-gather
// forms.  Every thread of the workgroup calls it (staging holds barriers).
template <int V, int kQ, typename G>
__device__ __forceinline__ float spmv_rows(const SysDesc& d, const TileRef& tr, bool sliced,
                                           const TileHdr& hdr, const SlicedPre& pre, G g,
                                           double* s_prod, float* s_v) {
  constexpr int kSliceF = SpmvLds<V>::kSliceF;
  constexpr bool kLocal = SpmvLds<V>::kLocal;
  // the header's fields as values (a lambda holding the header's address
  // would place it in scratch memory)
  const uint32_t h_off = hdr.off, h_cnt = hdr.cnt, h_goff = hdr.goff;
  const uint32_t* __restrict__ tgrp = d.tgrp;
  const uint32_t* __restrict__ tcols = d.tcols;
  if constexpr (V >= 4) {
    if (sliced)
      return sliced_rows<V == 4 ? kNt : 0, kQ>(pre, s_v, [=]() {
        if (h_goff) stage_groups(tgrp, h_goff - 1u, h_cnt, g, s_v);
        else stage_local(tcols, h_off, h_cnt, g, s_v);
      });
  }
  uint32_t mb, me;
  lane_bounds(d.row_ptr, tr.row0, tr.nrows, mb, me);
  float* slice = reinterpret_cast<float*>(s_prod) + (threadIdx.x >> 6) * kSliceF;
  if constexpr (kLocal) {
    if (h_cnt != kLocalOff) {
      // the wave's entry range: the bounds as separate scalars first (a
      // select over the header's array becomes an indexed scratch load)
      const uint32_t w = uni((uint32_t)(threadIdx.x >> 6));
      const uint32_t e0 = uni(hdr.eb[0]), e1 = uni(hdr.eb[1]), e2 = uni(hdr.eb[2]);
      const uint32_t e3 = uni(hdr.eb[3]), e4 = uni(hdr.eb[4]);
      const uint32_t wb = w == 0u ? e0 : w == 1u ? e1 : w == 2u ? e2 : e3;
      const uint32_t we = w == 0u ? e1 : w == 1u ? e2 : w == 2u ? e3 : e4;
      return spmv_wave_local<SpmvLds<V>::K>(wb, we, mb, me, d.lidx, d.val, s_v, slice, [=]() {
        stage_local(tcols, h_off, h_cnt, g, s_v);
      });
    }
    return spmv_wave<float, kPer>(mb, me, d.col, d.val, g, tr.nrows, slice);
  } else if constexpr (V == 1) {
    return spmv_wave_vec<float, 4>(mb, me, d.col, d.val, g, tr.nrows, slice);
  } else {
    return spmv_wave<float, kPer>(mb, me, d.col, d.val, g, tr.nrows, slice);
  }
}

// One tile of k_pcg_fspmv; kSliced: the system has the sliced layout (its
// loads are issued unconditionally, so their registers and resources stay
// out of the other systems' path).
template <int V, bool kOne, bool kDot64, bool kSliced>
__device__ __forceinline__ void fspmv_tile(const TileArgs& ta, const TileRef& tr, double* s_prod,
                                           float* s_v, SysState& s_ns) {
  constexpr uint32_t kWords = sizeof(SysState) / 4;
  const SysDesc& d = tr.d;
  const uint32_t t = threadIdx.x, nt = d.n_tiles;
  SysState* in = d.st + (ta.par ^ 1u) * ta.st_stride;
  SysState* out = d.st + ta.par * ta.st_stride;
  const bool writer = tr.tile == 0;
  // the matrix's loads that do not depend on the state go out first
  constexpr bool sliced = kSliced;
  TileHdr hdr{};
  SlicedPre pre;
  if constexpr (SpmvLds<V>::kLocal) hdr = uni_copy(d.thdr[tr.tile]);
  // a table's FP64-accumulated solves (batches, serving: their partials are
  // not prefetched) issue a row's first kPreQuads quads at once (256-system
  // batch 744 -> 776 solves/s); a single system keeps its registers for the
  // partials' prefetch (50^3: 21.9 against 23.4 us with both)
  constexpr int kQF = (!kOne && kDot64) ? kPreQuads : 1;
  if constexpr (kSliced) {
    const SliceHdr sh = sliced_hdr(d, tr.tile);
    sliced_issue<V == 4 ? kNt : 0, kQF>(d, sh, tr.row0 + 64u * uni((uint32_t)(t >> 6)), pre);
  }
  // the previous update's r.z partials, prefetched with the state (FP64
  // ones only for a single system: in a table their 32 registers cost more
  // rounds of workgroups than the load round they save)
  using TP = typename std::conditional<kDot64, double, float>::type;
  const TP* part = kDot64 ? reinterpret_cast<const TP*>(d.part64) : reinterpret_cast<const TP*>(d.part);
  const bool prefetch = (kOne || !kDot64) && nt > 1u && nt <= kFetchMax;
  TP vb[4][4];
  if (prefetch) level_fetch<TP>(part + 2u * nt, nt, vb);
  if (t < kWords) reinterpret_cast<uint32_t*>(&s_ns)[t] = reinterpret_cast<const uint32_t*>(in)[t];
  __syncthreads();
  const int32_t status = uni(s_ns.status), phase = uni(s_ns.phase);
  auto put_state = [&](bool both) {
    if (writer && t < kWords) {
      const uint32_t w = reinterpret_cast<const uint32_t*>(&s_ns)[t];
      reinterpret_cast<uint32_t*>(out)[t] = w;
      if (both) reinterpret_cast<uint32_t*>(in)[t] = w;
    }
  };
  if (status != kRunning) {  // finished: carry the state to the other copy
    put_state(false);
    return;
  }
  // 0: SpMV of x0 (init), 1: SpMV of p, 2: the FP64 residual
  int mode = -1;
  float beta = 0.0f;
  bool first = false;
  const bool fold = s_ns.outer >= 1u;  // residual: an inner solve's correction to fold in
  const double fold_scale = s_ns.inv_scale;
  double* s_a = s_prod;
  double* s_b = s_prod + kTile;
  if (phase == kPhStart || phase == kPhIrStart) {
    mode = phase == kPhStart ? 0 : 2;
    if (t == 0) {
      s_ns.phase = phase == kPhStart ? kPhInit : kPhIrResid;
      s_ns.sets += 1u;
    }
    __syncthreads();
    put_state(false);
  } else if (phase == kPhInit || phase == kPhIter || phase == kPhIrResid) {
    int32_t dec = kRunning;
    double rr64 = 0.0, bb64 = 0.0;
    if (phase == kPhIrResid) {
      dec = ir_decide(d, s_ns.target2, s_ns.outer, s_ns.max_outer, s_a, s_b, rr64, bb64);
      __syncthreads();  // s_a / s_b are reused below
    }
    const bool reduce = !(phase == kPhInit && s_ns.b_norm2 == 0.0f) &&
                        !(phase == kPhIrResid && dec != kRunning);
    // Within an iteration only the state writer needs r.r (the stop test):
    // the other workgroups take beta from r.z alone and run the SpMV even in
    // the set that stops the system or its inner solve (harmless: nothing
    // reads this q or p then; a stopped inner solve's correction is folded
    // in by the next residual).
    const bool beta_only = !writer && phase == kPhIter;
    if (reduce && beta_only) {
      TP* lb = reinterpret_cast<TP*>(s_b);
      const TP b = prefetch ? reduce_fetched<TP>(vb, nt, lb) : reduce_partials<TP>(part + 2u * nt, nt, lb);
      // pcg_transition's beta
      if constexpr (kDot64) {
        const double rz64_old = s_ns.rz64_old;
        beta = rz64_old != 0.0 ? (float)(b / rz64_old) : 0.0f;
      } else {
        const float rz_old = s_ns.rz_old;
        beta = rz_old != 0.0f ? b / rz_old : 0.0f;
      }
      mode = 1;
      __syncthreads();  // the slices alias the reduction's LDS
    } else {
      float rr = 0.0f, rz = 0.0f;
      double rz64 = 0.0;
      if (reduce) {
        TP* la = reinterpret_cast<TP*>(s_a);
        TP* lb = reinterpret_cast<TP*>(s_b);
        const TP a = reduce_partials<TP>(part + nt, nt, la);
        const TP b = prefetch ? reduce_fetched<TP>(vb, nt, lb) : reduce_partials<TP>(part + 2u * nt, nt, lb);
        rr = (float)a;
        rz = (float)b;
        if constexpr (kDot64) rz64 = b;
      }
      if (t == 0) {
        pcg_transition(s_ns, d.flags, rr, rz, rz64, dec, rr64, bb64);
        // an inner solve stopped: the next set computes the residual (with
        // the correction folded in)
        if (phase == kPhIter && s_ns.status == kRunning && s_ns.phase == kPhIrResid)
          s_ns.phase = kPhIrStart;
      }
      __syncthreads();
      const int32_t nstatus = uni(s_ns.status), nphase = uni(s_ns.phase);
      put_state(nstatus != kRunning);
      if (nstatus != kRunning) return;
      beta = s_ns.beta;
      if (nphase == kPhIter) {
        mode = 1;
        first = uni(s_ns.iter) == 0u;
      }
    }
  }
  if (mode < 0) return;
  const uint32_t row = tr.row0 + t, rowc = t < tr.nrows ? row : tr.row0;
  const bool own = t < tr.nrows;
  if (mode == 2) {
    if constexpr (kSliced) {
      spmv_ir_resid_sliced(d, tr, hdr, pre, s_prod, fold, fold_scale);
      return;
    }
    uint32_t mb, me;
    lane_bounds(d.row_ptr, tr.row0, tr.nrows, mb, me);
    spmv_ir_resid<V>(d, tr, mb, me, s_prod, fold, fold_scale);
    return;
  }
  if (mode == 0) {  // compute.rs:294 (q = A x0) and ||b||^2 (:235-242)
    const __amdgpu_buffer_rsrc_t rx = buf_rsrc(d.x, d.n * 4u);
    auto xg = [&](uint32_t c) -> float { return buf_gather<float>(rx, c); };
    const float a = bld(d.b, rowc);
    const float q = spmv_rows<V, kQF>(d, tr, sliced, hdr, pre, xg, s_prod, s_v);
    spmv_tail(d, tr, q, true, a, s_prod);
    return;
  }
  // p_k = (p_{k-1} * beta) + z (scale_from_scalar_results.wgsl:37, axpy.wgsl:41
  // with alpha = 1), or z on the first iteration (compute.rs:313-314)
  float* const pcur = p_of(d, ta.par);
  const __amdgpu_buffer_rsrc_t rpp = buf_rsrc(p_of(d, ta.par ^ 1u), d.n * 4u);
  const __amdgpu_buffer_rsrc_t rz = buf_rsrc(d.z, d.n * 4u);
  auto pg = [&](uint32_t c) -> float {
    const float zc = buf_gather<float>(rz, c);
    const float pc = buf_gather<float>(rpp, c);
    const float pb = pc * beta;
    return first ? zc : pb + zc;
  };
  const float a = pg(rowc);
  if (own) bst(pcur, row, a);
  const float q = spmv_rows<V, kQF>(d, tr, sliced, hdr, pre, pg, s_prod, s_v);
  spmv_tail(d, tr, q, false, a, s_prod);
}

template <int V, bool kOne, bool kDot64>
__global__ __launch_bounds__(kTile) void k_pcg_fspmv(const TileArgs ta) {
  __shared__ __align__(16) double s_prod[kWaves * SpmvLds<V>::kSliceF / 2];
  __shared__ float s_v[SpmvLds<V>::kLocal ? kLocalMax + kWaves : 1];  // tile-local copy of the vector
  __shared__ SysState s_ns;
  static_assert(sizeof(SysState) / 4 <= (uint32_t)kTile && sizeof(SysState) % 4 == 0,
                "state copy by dwords");
  static_assert(kWaves * SpmvLds<V>::kSliceF / 2 >= 2 * kTile, "finish reductions alias the slices");
  TileRef tr;
  if (!get_tile<kOne>(ta, tr)) return;
  if constexpr (V >= 4) {
    if (tr.d.flags & kFlagSliced) {
      fspmv_tile<V, kOne, kDot64, true>(ta, tr, s_prod, s_v, s_ns);
      return;
    }
  }
  fspmv_tile<V, kOne, kDot64, false>(ta, tr, s_prod, s_v, s_ns);
}

// ------------------------------------------ one-reduction PCG (optional)
//
// Chronopoulos-Gear PCG: the three dots of an iteration -- (r,u), (w,u) with
// w = A u, and (r,r) -- are reduced at one point, so an iteration is two
// launches where the reference's PCG needs three (and, across GPUs, one
// all-reduce where it needs two).  In exact arithmetic its iterates are the
// reference's; its operations are not, so its results differ in rounding
// and it is never the drop-in path (hbj_system_set_algorithm opts in).
//   k_ccg_spmv    w = A u and the tile's (r,u), (w,u), (r,r) partials
//                 [init: q = A x0 and b.b]
//   k_ccg_update  every workgroup reduces the three partials (same trees,
//                 same bits) and takes the same decision: stop test on
//                 (r,r) with the reference's thresholds, or
//                   beta = gamma / gamma_prev,
//                   alpha = gamma / (delta - beta * (gamma / alpha_prev))
//                   (first iteration: beta = 0, alpha = gamma / delta)
//                 then p = u + beta p, s = w + beta s, x = x + alpha p,
//                 r = r + (-alpha) s and u = M^-1 r; its tile-0 workgroup
//                 writes the next state  [init: r = b + (-1) q, u = M^-1 r,
//                 p = s = 0]
// Vectors: u in SysDesc.z, w in SysDesc.q; r and s double-buffered by
// parity (a tile reads its neighbours' rows for straddling blocks).

// w = A u of the tile's rows stored, and the tile's partials: b.b at init,
// else (r,u), (w,u) and (r,r).
__device__ __forceinline__ void ccg_spmv_tail(const SysDesc& d, const TileRef& tr, float q, bool init,
                                              float ri, float ui, double* s_prod) {
  float* s_ga = reinterpret_cast<float*>(s_prod);  // tile reductions reuse the slices
  float* s_gb = s_ga + kTile;
  float* s_gc = s_gb + kTile;
  const uint32_t t = threadIdx.x, row = tr.row0 + t;
  const bool own = t < tr.nrows;
  __syncthreads();  // all waves are done with their slices
  if (own) bst(d.q, row, q);
  if (init) {
    float bb = 0.0f;
    if (own) {
      const float bi = bld(d.b, row);
      bb = bi * bi;
    }
    const float pd = tree256<float>(bb, s_ga);
    if (t == 0) d.part[tr.tile] = pd;
    return;
  }
  float g = own ? ri * ui : 0.0f, dl = own ? q * ui : 0.0f;
  const float rr = tree256<float>(own ? ri * ri : 0.0f, s_gc);
  tree256x2<float>(g, dl, s_ga, s_gb);
  if (t == 0) {
    const uint32_t nt = d.n_tiles;
    d.part[tr.tile] = g;
    d.part[nt + tr.tile] = dl;
    d.part[2u * nt + tr.tile] = rr;
  }
}

// One tile of k_ccg_spmv on a system with the sliced layout: as
// spmv_tile_sliced, the headers go out with the state and the matrix's
// first loads before the state is tested, on a path of its own.
template <int kAux>
__device__ __forceinline__ void ccg_tile_sliced(const TileArgs& ta, const TileRef& tr, double* s_prod,
                                                float* s_v) {
  const SysDesc& d = tr.d;
  const SysState* st = d.st + ta.par * ta.st_stride;
  const uint32_t t = threadIdx.x;
  const uint32_t rowc = t < tr.nrows ? tr.row0 + t : tr.row0;
  const TileHdr hdr = uni_copy(d.thdr[tr.tile]);
  const SliceHdr sh = sliced_hdr(d, tr.tile);
  const int32_t status = uni(st->status), phase = uni(st->phase);
  SlicedPre pre;
  sliced_issue<kAux>(d, sh, tr.row0 + 64u * uni((uint32_t)(threadIdx.x >> 6)), pre);
  const float* __restrict__ r_cur = ta.par ? d.r[1] : d.r[0];
  const float ri = bld(r_cur, rowc), ui = bld(d.z, rowc);
  if (status != kRunning) return;
  if (phase != kPhCcgInit && phase != kPhCcg) return;
  const bool init = phase == kPhCcgInit;
  const __amdgpu_buffer_rsrc_t rv = buf_rsrc(init ? d.x : d.z, d.n * 4u);
  auto vec_at = [&](uint32_t c) -> float { return buf_gather<float>(rv, c); };
  const float q = sliced_rows<kAux>(pre, s_v, [&]() {
    if (hdr.goff) stage_groups(d.tgrp, hdr.goff - 1u, hdr.cnt, vec_at, s_v);
    else stage_local(d.tcols, hdr.off, hdr.cnt, vec_at, s_v);
  });
  ccg_spmv_tail(d, tr, q, init, ri, ui, s_prod);
}

template <int V, bool kOne>
__global__ __launch_bounds__(kTile) void k_ccg_spmv(const TileArgs ta) {
  __shared__ __align__(16) double s_prod[kWaves * SpmvLds<V>::kSliceF / 2];
  __shared__ float s_v[SpmvLds<V>::kLocal ? kLocalMax + kWaves : 1];  // tile-local copy of the vector
  constexpr int kSliceF = SpmvLds<V>::kSliceF;
  constexpr bool kLocal = SpmvLds<V>::kLocal;
  TileRef tr;
  if (!get_tile<kOne>(ta, tr)) return;
  const SysDesc& d = tr.d;
  if constexpr (V >= 4) {
    if (d.flags & kFlagSliced) {
      ccg_tile_sliced<V == 4 ? kNt : 0>(ta, tr, s_prod, s_v);
      return;
    }
  }
  const SysState* st = d.st + ta.par * ta.st_stride;
  const uint32_t t = threadIdx.x, row = tr.row0 + t, wave = t >> 6;
  const bool own = t < tr.nrows;
  const uint32_t rowc = own ? row : tr.row0;
  TileHdr hdr{};
  if constexpr (kLocal) hdr = uni_copy(d.thdr[tr.tile]);
  uint32_t mb, me;
  lane_bounds(d.row_ptr, tr.row0, tr.nrows, mb, me);
  const float* __restrict__ r_cur = ta.par ? d.r[1] : d.r[0];
  const float ri = bld(r_cur, rowc), ui = bld(d.z, rowc);
  const int32_t status = uni(st->status), phase = uni(st->phase);
  if (status != kRunning) return;
  if (phase != kPhCcgInit && phase != kPhCcg) return;
  const bool init = phase == kPhCcgInit;
  const __amdgpu_buffer_rsrc_t rv = buf_rsrc(init ? d.x : d.z, d.n * 4u);
  auto vec_at = [&](uint32_t c) -> float { return buf_gather<float>(rv, c); };
  float* slice = reinterpret_cast<float*>(s_prod) + wave * kSliceF;
  float q;
  if constexpr (kLocal) {
    if (hdr.cnt != kLocalOff) {
      const uint32_t w = uni(wave);
      const uint32_t wb = w == 0u ? hdr.eb[0] : w == 1u ? hdr.eb[1] : w == 2u ? hdr.eb[2] : hdr.eb[3];
      const uint32_t we = w == 0u ? hdr.eb[1] : w == 1u ? hdr.eb[2] : w == 2u ? hdr.eb[3] : hdr.eb[4];
      q = spmv_wave_local<SpmvLds<V>::K>(wb, we, mb, me, d.lidx, d.val, s_v, slice, [&]() {
        stage_local(d.tcols, hdr.off, hdr.cnt, vec_at, s_v);
      });
    } else {
      q = spmv_wave<float, kPer>(mb, me, d.col, d.val, vec_at, tr.nrows, slice);
    }
  } else if constexpr (V == 1) {
    q = spmv_wave_vec<float, 4>(mb, me, d.col, d.val, vec_at, tr.nrows, slice);
  } else {
    q = spmv_wave<float, kPer>(mb, me, d.col, d.val, vec_at, tr.nrows, slice);
  }
  ccg_spmv_tail(d, tr, q, init, ri, ui, s_prod);
}

// kUT consecutive tiles of the system per workgroup (single system: tiles
// kUT*b + u); the three reductions and the decision run once per workgroup.
// kPre: the partials are fetched before the state (48 more VGPRs; the
// two-tile form goes without: 53 VGPRs and one round of workgroups, 15.1 us
// at 50^3 against 15.5 with the prefetch and 16.5 for one tile each).
template <bool kOne, uint32_t kUT, bool kPre>
__global__ __launch_bounds__(kTile) void k_ccg_update(const TileArgs ta) {
  __shared__ float s_rh[kTile + 2 * kHalo];
  __shared__ float s_z[kTile];
  __shared__ float s_lu[kLuPer * kTile];
  __shared__ float s_g[3][kTile];
  __shared__ uint32_t s_blk[2];
  __shared__ SysState s_ns;
  TileRef tr;
  if constexpr (kOne) {
    tr.d = ta.one;
    tr.tile = kUT * (ta.remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x);
    if (tr.tile >= tr.d.n_tiles) return;
  } else {
    static_assert(kUT == 1, "tables take one tile per workgroup");
    if (!get_tile<false>(ta, tr)) return;
  }
  const SysDesc& d = tr.d;
  SysState* st = d.st + ta.par * ta.st_stride;
  SysState* st_next = d.st + (ta.par ^ 1u) * ta.st_stride;
  const uint32_t t = threadIdx.x, nt = d.n_tiles, tile0 = tr.tile;
  const float* __restrict__ r_in = ta.par ? d.r[1] : d.r[0];
  float* __restrict__ r_out = ta.par ? d.r[0] : d.r[1];
  const float* __restrict__ s_in = ta.par ? d.s[1] : d.s[0];
  float* __restrict__ s_out = ta.par ? d.s[0] : d.s[1];
  const bool b6 = (d.flags & kFlagBlocks6) != 0u;
  struct Ld {
    uint32_t tile, row0, nrows, row, rowc, hrow, hslot, hrowc, bs_s, bs_e;
    bool own, halo;
    float xi, pi, ui, wi, ri, si, hr, hs, hw;
    uint2 tb, blk;
  };
  // a tile's loads that do not depend on the state
  auto load = [&](uint32_t tile, Ld& L) {
    L.tile = tile;
    L.row0 = tile * (uint32_t)kTile;
    const uint32_t left = d.n - L.row0;
    L.nrows = left < (uint32_t)kTile ? left : (uint32_t)kTile;
    L.own = t < L.nrows;
    L.row = L.row0 + t;
    L.rowc = L.own ? L.row : L.row0;
    L.hrow = 0;
    L.hslot = 0;
    L.halo = halo_row(L.row0, L.nrows, d.n, L.hrow, L.hslot);
    L.hrowc = L.halo ? L.hrow : L.row0;
    L.xi = bld(d.x, L.rowc);
    L.pi = bld(d.p, L.rowc);
    L.ui = bld(d.z, L.rowc);
    L.wi = bld(d.q, L.rowc);
    L.ri = bld(r_in, L.rowc);
    L.si = bld(s_in, L.rowc);
    L.hr = bld(r_in, L.hrowc);
    L.hs = bld(s_in, L.hrowc);
    L.hw = bld(d.q, L.hrowc);
    L.tb = make_uint2(0u, 0u);
    L.bs_s = L.bs_e = 0;
    if (b6) {
      const uint32_t first = L.row0 / 6u, last = (L.row0 + L.nrows - 1u) / 6u;
      L.tb = make_uint2(first, last - first + 1u);
      const uint32_t bi = t < L.tb.y ? L.tb.x + t : L.tb.x;
      L.bs_s = 6u * bi;
      L.bs_e = 6u * bi + 6u < d.n ? 6u * bi + 6u : d.n;
      bj_stage(d.lu, L.tb.x, 0, L.tb.y < (uint32_t)kLuStage ? L.tb.y : (uint32_t)kLuStage, s_lu);
    }
    L.blk = (!b6 && d.tile_blk) ? d.tile_blk[tile] : make_uint2(0u, 0u);
  };
  const bool pre = kPre && nt > 1u && nt <= kFetchMax;
  float vg[4][4], vd[4][4], vr[4][4];
  if (pre) {
    level_fetch<float>(d.part, nt, vg);
    level_fetch<float>(d.part + nt, nt, vd);
    level_fetch<float>(d.part + 2u * nt, nt, vr);
  }
  Ld L;
  load(tile0, L);
  const int32_t status = uni(st->status), phase = uni(st->phase);
  const uint32_t it = uni(st->iter), max_iter = uni(st->max_iter);
  const float g_prev = st->rz_old, a_prev = st->alpha;
  const float rel2 = st->rel_tol2, abs2 = st->abs_tol2, bn_sys = st->b_norm2;
  if (status != kRunning) return;
  if (phase != kPhCcgInit && phase != kPhCcg) return;
  const bool init = phase == kPhCcgInit;

  int32_t fin = kRunning;
  float g = 0.0f, dl = 0.0f, rr = 0.0f, bn = 0.0f, alpha = 0.0f, beta = 0.0f;
  if (init) {
    bn = pre ? reduce_fetched<float, true>(vg, nt, s_g[0]) : reduce_partials<float, true>(d.part, nt, s_g[0]);
    if (bn == 0.0f) fin = 0;  // Ok(0), x untouched (compute.rs:252-254)
  } else {
    if (pre) {
      reduce_fetched3<float, true>(vg, vd, vr, nt, s_g[0], s_g[1], s_g[2], g, dl, rr);
    } else {
      g = reduce_partials<float, true>(d.part, nt, s_g[0]);
      dl = reduce_partials<float, true>(d.part + nt, nt, s_g[1]);
      rr = reduce_partials<float, true>(d.part + 2u * nt, nt, s_g[2]);
    }
    // r is r after `it` updates: the reference's stop test after an
    // iteration (compute.rs:481), never on r0
    if (it > 0u && (rr <= abs2 || rr <= rel2 * bn_sys)) {
      fin = 0;
    } else if (it >= max_iter) {
      fin = 5;  // HBJ_E_NOT_CONVERGED
    } else {
      float den = dl;
      if (it > 0u) {
        if (g_prev == 0.0f) {
          fin = 4;  // HBJ_E_BREAKDOWN_RZ
        } else {
          beta = g / g_prev;
          const float ga = g / a_prev;
          den = dl - beta * ga;
        }
      }
      if (fin == kRunning) {
        if (den == 0.0f) fin = 3;  // HBJ_E_BREAKDOWN_PAP: (p, A p) = 0
        else alpha = g / den;
      }
    }
  }
  if (tile0 == 0) {  // the next state, staged in LDS and copied out by dwords
    constexpr uint32_t kWords = sizeof(SysState) / 4;
    static_assert(kWords <= (uint32_t)kTile && sizeof(SysState) % 4 == 0, "state copy by dwords");
    if (t < kWords) reinterpret_cast<uint32_t*>(&s_ns)[t] = reinterpret_cast<const uint32_t*>(st)[t];
    __syncthreads();
    if (t == 0) {
      SysState& ns = s_ns;
      ns.sets += 1u;
      if (init) {
        ns.b_norm2 = bn;
        if (fin != kRunning) {
          ns.status = fin;
        } else {
          ns.phase = kPhCcg;
          ns.iter = 0;
        }
      } else {
        ns.rr = rr;
        ns.rz_new = g;
        if (fin != kRunning) {
          ns.status = fin;
        } else {
          ns.rz_old = g;
          ns.alpha = alpha;
          ns.neg_alpha = -alpha;
          ns.beta = beta;
          ns.iter = it + 1u;
          ns.total_iter = it + 1u;
        }
      }
    }
    __syncthreads();
    if (t < kWords) {
      const uint32_t w = reinterpret_cast<const uint32_t*>(&s_ns)[t];
      reinterpret_cast<uint32_t*>(st_next)[t] = w;
      if (fin != kRunning) reinterpret_cast<uint32_t*>(st)[t] = w;  // a final state into both copies
    }
  }
  if (fin != kRunning) return;
  const float na = -alpha;
  for (uint32_t u = 0; u < kUT; ++u) {
    if (u > 0u) {
      if (tile0 + u >= nt) break;
      load(tile0 + u, L);
    }
    // b at init; the block table when blocks are not 6 rows
    float rs = L.ri, rhs_ = L.hr;
    if (init) {
      rs = bld(d.b, L.rowc);
      rhs_ = bld(d.b, L.hrowc);
    }
    uint2 tb = L.tb;
    uint32_t bs_s = L.bs_s, bs_e = L.bs_e;
    if (!b6) {
      const uint2 tb0 = d.tile_blk ? L.blk : tile_blocks(d.bstart, nullptr, L.tile, d.nb, L.row0,
                                                         L.nrows, s_blk);
      tb = make_uint2(uni(tb0.x), uni(tb0.y));
      const uint32_t bi = t < tb.y ? tb.x + t : tb.x;
      bs_s = bld(d.bstart, bi);
      bs_e = bld(d.bstart, bi + 1u);
      bj_stage(d.lu, tb.x, 0, tb.y < (uint32_t)kLuStage ? tb.y : (uint32_t)kLuStage, s_lu);
    }
    float rn = 0.0f;
    if (L.own) {
      if (init) {
        rn = rs + (-1.0f) * L.wi;  // r0 = b - A x0 (q holds A x0)
        bst(d.p, L.row, 0.0f);
        bst(s_out, L.row, 0.0f);
      } else {
        const float pn = L.ui + beta * L.pi;
        const float sn = L.wi + beta * L.si;
        rn = rs + na * sn;
        bst(d.p, L.row, pn);
        bst(s_out, L.row, sn);
        bst(d.x, L.row, L.xi + alpha * pn);
      }
      bst(r_out, L.row, rn);
      s_rh[t + kHalo] = rn;
      s_z[t] = L.ui;  // rows no block writes keep u
    }
    if (L.halo) s_rh[L.hslot] = init ? rhs_ + (-1.0f) * L.hw : rhs_ + na * (L.hw + beta * L.hs);
    lds_dma_wait();  // the factor copy into s_lu
    __syncthreads();
    bj_tile(d.lu, tb, d.n, L.row0, L.nrows, bs_s, bs_e, s_rh, s_z, s_lu);
    if (L.own) bst(d.z, L.row, s_z[t]);
    __syncthreads();  // s_z and s_rh are written again by the next tile
  }
}

// FP64 check of an FP32 solution: per tile ||b - A x||^2 and ||b||^2 with
// FP32 inputs promoted exactly to FP64 (host sums the n_tiles partials).
__global__ __launch_bounds__(kTile) void k_true_resid(const SysDesc* __restrict__ dp,
                                                      const float* __restrict__ x,
                                                      double* __restrict__ part2,
                                                      uint32_t n_tiles) {
  __shared__ double s_prod[kWaves * 64 * kPer / 2];
  __shared__ double s_ra[kTile];
  __shared__ double s_rb[kTile];
  const SysDesc d = *dp;
  const uint32_t tile = blockIdx.x;
  const uint32_t row0 = tile * kTile;
  const uint32_t nrows = (d.n - row0) < (uint32_t)kTile ? d.n - row0 : (uint32_t)kTile;
  uint32_t mb, me;
  lane_bounds(d.row_ptr, row0, nrows, mb, me);
  const __amdgpu_buffer_rsrc_t rx = buf_rsrc(x, d.n * 4u);
  const double ax = spmv_wave<double, kPer / 2>(mb, me, d.col, d.val,
                                                [&](uint32_t c) { return buf_gather<float>(rx, c); }, nrows,
                                                       s_prod + (threadIdx.x >> 6) * 64 * (kPer / 2));
  double rr = 0.0, bb = 0.0;
  if (threadIdx.x < nrows) {
    const double bi = (double)d.b[row0 + threadIdx.x];
    const double ri = bi - ax;
    rr = ri * ri;
    bb = bi * bi;
  }
  tree256x2<double>(rr, bb, s_ra, s_rb);
  if (threadIdx.x == 0) {
    part2[tile] = rr;
    part2[n_tiles + tile] = bb;
  }
}

void launch_true_resid(const SysDesc* d, const float* x, double* part2, uint32_t n_tiles,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_true_resid, dim3(n_tiles), dim3(kTile), 0, s, d, x, part2, n_tiles);
}

// Plain streaming copy (16-byte loads and stores): the calibration kernel for
// the HBM traffic counters and a measured ceiling for the bandwidth figures.
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_stream_copy(const f32x4* __restrict__ src,
                                                     f32x4* __restrict__ dst, uint64_t n4) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
    dst[i] = __builtin_nontemporal_load(src + i);
}

void launch_stream_copy(const void* src, void* dst, uint64_t n4, hipStream_t s) {
  const uint64_t want = (n4 + 255) / 256;
  const uint32_t grid = (uint32_t)(want < 16384 ? (want ? want : 1) : 16384);
  hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, s, (const f32x4*)src, (f32x4*)dst, n4);
}

// ------------------------------------------------------ LU block build

// build_lu_blocks_from_csr_block_starts_6 (compute.rs:84-181) with
// lu_factor_inplace_6 (:37-71), one thread per block: extract the block's
// leading m x m (m = min(6, size)) from its rows' CSR entries in CSR order
// (a duplicate's last write wins) into an identity-filled 6x6, then the
// no-pivot LU in place, operation for operation as on the host.  Blocks
// with end > n or start >= end stay zero (:141-143).  The 6x6 lives in LDS
// (stride 37 words per thread) because the extraction indexes it at run
// time.  A zero pivot records (block << 32 | k) with an atomic minimum, so
// the host reports the first failing block as the host loop does; the
// other blocks are still written.  block_starts is validated by the caller.
__global__ __launch_bounds__(kTile) void k_build_lu6(uint32_t n, const uint32_t* __restrict__ row_ptr,
                                                     const uint32_t* __restrict__ col,
                                                     const float* __restrict__ val,
                                                     const uint32_t* __restrict__ bs, uint32_t nb,
                                                     float* __restrict__ lu,
                                                     unsigned long long* __restrict__ fail) {
  __shared__ float s_m[kTile * 37];
  const uint32_t blk = blockIdx.x * kTile + threadIdx.x;
  if (blk >= nb) return;
  float* mat = s_m + threadIdx.x * 37u;
  for (int i = 0; i < 36; ++i) mat[i] = 0.0f;
  const uint32_t off = bs[blk], end = bs[blk + 1];
  if (!(end > n || off >= end)) {
    const uint32_t m = end - off < 6u ? end - off : 6u;
    for (int i = 0; i < 6; ++i) mat[i * 6 + i] = 1.0f;
    for (uint32_t il = 0; il < m; ++il) {
      const uint32_t i = off + il;
      const uint32_t k1 = row_ptr[i + 1];
      for (uint32_t k = row_ptr[i]; k < k1; ++k) {
        const uint32_t j = col[k];
        if (j >= off && j < off + m) mat[il * 6u + (j - off)] = val[k];
      }
    }
    for (uint32_t k = 0; k < m; ++k) {
      const float pivot = mat[k * 6u + k];
      if (pivot == 0.0f) {
        atomicMin(fail, ((unsigned long long)blk << 32) | k);
        break;
      }
      for (uint32_t i = k + 1; i < m; ++i) mat[i * 6u + k] /= pivot;
      for (uint32_t i = k + 1; i < m; ++i) {
        const float l = mat[i * 6u + k];
        if (l == 0.0f) continue;  // compute.rs:62
        for (uint32_t j = k + 1; j < m; ++j) mat[i * 6u + j] -= l * mat[k * 6u + j];
      }
    }
  }
  float* dst = lu + 36ull * blk;
  for (int i = 0; i < 36; ++i) dst[i] = mat[i];
}

void launch_build_lu6(uint32_t n, const uint32_t* row_ptr, const uint32_t* col, const float* val,
                      const uint32_t* bs, uint32_t nb, float* lu, unsigned long long* fail,
                      hipStream_t s) {
  if (nb == 0) return;
  hipLaunchKernelGGL(k_build_lu6, dim3((nb + kTile - 1) / kTile), dim3(kTile), 0, s, n, row_ptr,
                     col, val, bs, nb, lu, fail);
}

// ------------------------------------------------- single-op kernels (KATs)

__global__ void k_op_axpy(uint32_t n, float alpha, const float* __restrict__ x, float* __restrict__ y) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = y[i] + alpha * x[i];
}

__global__ void k_op_scale(uint32_t n, float beta, float* __restrict__ x) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = x[i] * beta;
}

__global__ __launch_bounds__(kTile) void k_op_dot_partials(uint32_t n, const float* __restrict__ a,
                                                           const float* __restrict__ b, float* part) {
  __shared__ float sh[kTile];
  const uint32_t i = blockIdx.x * kTile + threadIdx.x;
  const float v = i < n ? a[i] * b[i] : 0.0f;  // dot_partials.wgsl:48-52
  const float p = tree256<float>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = p;
}

__global__ __launch_bounds__(kTile) void k_op_dot_reduce(const float* __restrict__ part, uint32_t len,
                                                         float* scratch, float* out) {
  __shared__ float sh[kTile];
  const float r = reduce_partials<float>(part, len, sh, scratch);
  if (threadIdx.x == 0) *out = r;
}

__global__ __launch_bounds__(kTile) void k_op_spmv(uint32_t n, const uint32_t* __restrict__ rp,
                                                   const uint32_t* __restrict__ col,
                                                   const float* __restrict__ val,
                                                   const float* __restrict__ x, float* __restrict__ y) {
  __shared__ float s_prod[kWaves * 64 * kPer];
  const uint32_t row0 = blockIdx.x * kTile;
  const uint32_t nrows = (n - row0) < (uint32_t)kTile ? n - row0 : (uint32_t)kTile;
  uint32_t mb, me;
  lane_bounds(rp, row0, nrows, mb, me);
  const __amdgpu_buffer_rsrc_t rx = buf_rsrc(x, n * 4u);
  const float q = spmv_wave<float, kPer>(mb, me, col, val,
                                         [&](uint32_t c) { return buf_gather<float>(rx, c); }, nrows,
                                                s_prod + (threadIdx.x >> 6) * 64 * kPer);
  if (threadIdx.x < nrows) y[row0 + threadIdx.x] = q;
}

__global__ __launch_bounds__(kTile) void k_op_block_jacobi(uint32_t n, uint32_t nb,
                                                           const float* __restrict__ lu,
                                                           const uint32_t* __restrict__ bs,
                                                           const float* __restrict__ r,
                                                           float* __restrict__ z) {
  __shared__ float s_rh[kTile + 2 * kHalo];
  __shared__ float s_z[kTile];
  __shared__ float s_lu[kLuPer * kTile];
  __shared__ uint32_t s_blk[2];
  const uint32_t row0 = blockIdx.x * kTile;
  const uint32_t nrows = (n - row0) < (uint32_t)kTile ? n - row0 : (uint32_t)kTile;
  const uint32_t t = threadIdx.x;
  if (t < nrows) {
    s_rh[t + kHalo] = r[row0 + t];
    s_z[t] = z[row0 + t];  // rows no block writes keep z (block_jacobi.wgsl:106-122)
  }
  uint32_t hrow = 0, hslot = 0;
  if (halo_row(row0, nrows, n, hrow, hslot)) s_rh[hslot] = r[hrow];
  const uint2 blk = tile_blocks(bs, nullptr, blockIdx.x, nb, row0, nrows, s_blk);
  uint32_t bs_s = 0, bs_e = 0;
  if (t < blk.y) {
    bs_s = bs[blk.x + t];
    bs_e = bs[blk.x + t + 1u];
  }
  bj_stage(lu, blk.x, 0, blk.y < (uint32_t)kLuStage ? blk.y : (uint32_t)kLuStage, s_lu);
  lds_dma_wait();
  __syncthreads();
  bj_tile(lu, blk, n, row0, nrows, bs_s, bs_e, s_rh, s_z, s_lu);
  if (t < nrows) z[row0 + t] = s_z[t];
}

__global__ void k_op_update_scalars(float* s, uint32_t i_pap, uint32_t i_rz_new, uint32_t i_rz_old,
                                    uint32_t i_alpha, uint32_t i_malpha, uint32_t i_beta) {
  // pcg_update_scalars.wgsl:47-73
  const float p_ap = s[i_pap], rz_new = s[i_rz_new], rz_old = s[i_rz_old];
  float alpha = 0.0f, beta = 0.0f;
  if (p_ap != 0.0f) alpha = rz_old / p_ap;
  if (rz_old != 0.0f) beta = rz_new / rz_old;
  s[i_alpha] = alpha;
  s[i_malpha] = -alpha;
  s[i_beta] = beta;
}

// ------------------------------------------------------- launch wrappers

// Launches a PCG kernel; when the launch set carries timing events they take
// the kernel's own start and end timestamps (hipExtLaunchKernelGGL), so the
// measured time excludes the dispatch gap before the kernel.
template <typename... A, typename K>
static void launch(const LaunchSet& ls, K k, dim3 g, dim3 b, hipStream_t s, A... args) {
  if (ls.ev_start || ls.ev_stop)
    hipExtLaunchKernelGGL(k, g, b, 0, s, ls.ev_start, ls.ev_stop, 0, args...);
  else
    hipLaunchKernelGGL(k, g, b, 0, s, args...);
}

static TileArgs tile_args(const LaunchSet& ls) {
  TileArgs a{};
  if (ls.one) a.one = *ls.one;
  a.descs = ls.descs;
  a.tiles = ls.tiles;
  a.n_tiles = ls.n_tiles;
  a.remap = ls.remap;
  a.par = ls.par;
  a.st_stride = ls.st_stride;
  return a;
}

template <int V>
static void launch_spmv_v(const LaunchSet& ls, hipStream_t s) {
  const TileArgs a = tile_args(ls);
  const dim3 g(ls.n_tiles), b(kTile);
  if (ls.one) launch(ls, k_pcg_spmv<V, true>, g, b, s, a);
  else launch(ls, k_pcg_spmv<V, false>, g, b, s, a);
}

void launch_pcg_spmv(const LaunchSet& ls, hipStream_t s) {
  switch (ls.spmv_variant) {
    case 1: launch_spmv_v<1>(ls, s); break;
    case 2: launch_spmv_v<2>(ls, s); break;
    case 3: launch_spmv_v<3>(ls, s); break;
    case 4: launch_spmv_v<4>(ls, s); break;
    case 5: launch_spmv_v<5>(ls, s); break;
    default: launch_spmv_v<0>(ls, s); break;
  }
}
// k_level1 ahead of the update (which 0) or the finish (which 1) of a
// kFlagL1 system; with kernel timing the pair is timed as one step (the
// start event on k_level1, the stop event on the kernel after it).
static bool launch_level1(LaunchSet& ls, uint32_t which, hipStream_t s) {
  if (!ls.one || !(ls.one->flags & kFlagL1)) return false;
  const SysDesc& d = *ls.one;
  const uint32_t g = (d.n_tiles + 255u) / 256u;
  const uint32_t jobs = which == 0u ? ((d.flags & kFlagDot64) ? 2u : 1u) : 2u;
  const dim3 grid((jobs * g + 3u) / 4u), b(kTile);
  if (ls.ev_start) {
    hipExtLaunchKernelGGL(k_level1, grid, b, 0, s, ls.ev_start, nullptr, 0, d, which);
    ls.ev_start = nullptr;
  } else {
    hipLaunchKernelGGL(k_level1, grid, b, 0, s, d, which);
  }
  return true;
}

void launch_pcg_update(const LaunchSet& ls0, hipStream_t s) {
  LaunchSet ls = ls0;
  launch_level1(ls, 0u, s);
  const TileArgs a = tile_args(ls);
  const dim3 g(ls.n_tiles), b(kTile);
  if (ls.update_groups && ls.one) {
    const dim3 g2((ls.n_tiles + 1) / 2);
    if (ls.dot64) launch(ls, k_pcg_update_grp<true, true, 2>, g2, b, s, a);
    else launch(ls, k_pcg_update_grp<true, false, 2>, g2, b, s, a);
    return;
  }
  if (ls.update_groups) {  // the finish table's groups of kFinishTiles tiles
    TileArgs ag = a;
    ag.tiles = ls.ftiles;
    ag.n_tiles = ls.n_ftiles;
    const dim3 g4(ls.n_ftiles ? ls.n_ftiles : 1);
    if (ls.dot64) launch(ls, k_pcg_update_grp<false, true, kFinishTiles>, g4, b, s, ag);
    else launch(ls, k_pcg_update_grp<false, false, kFinishTiles>, g4, b, s, ag);
    return;
  }
  if (ls.dot64) {
    if (ls.one) launch(ls, k_pcg_update<true, true>, g, b, s, a);
    else launch(ls, k_pcg_update<false, true>, g, b, s, a);
  } else {
    if (ls.one) launch(ls, k_pcg_update<true, false>, g, b, s, a);
    else launch(ls, k_pcg_update<false, false>, g, b, s, a);
  }
}
template <bool kDot64>
static void launch_finish_t(const LaunchSet& ls, const TileArgs& a, hipStream_t s) {
  if (ls.one) {
    const uint32_t g = (ls.n_tiles + kFinishTiles - 1) / kFinishTiles;
    launch(ls, k_pcg_finish<true, kDot64>, dim3(g ? g : 1), dim3(kTile), s, a);
  } else {
    TileArgs ag = a;
    ag.tiles = ls.ftiles;
    ag.n_tiles = ls.n_ftiles;
    launch(ls, k_pcg_finish<false, kDot64>, dim3(ls.n_ftiles ? ls.n_ftiles : 1), dim3(kTile), s, ag);
  }
}
void launch_pcg_finish(const LaunchSet& ls0, hipStream_t s) {
  LaunchSet ls = ls0;
  launch_level1(ls, 1u, s);
  const TileArgs a = tile_args(ls);
  if (ls.dot64) launch_finish_t<true>(ls, a, s);
  else launch_finish_t<false>(ls, a, s);
}

template <int V>
static void launch_fspmv_v(const LaunchSet& ls, hipStream_t s) {
  const TileArgs a = tile_args(ls);
  const dim3 g(ls.n_tiles), b(kTile);
  if (ls.one) {
    if (ls.dot64) launch(ls, k_pcg_fspmv<V, true, true>, g, b, s, a);
    else launch(ls, k_pcg_fspmv<V, true, false>, g, b, s, a);
  } else {
    if (ls.dot64) launch(ls, k_pcg_fspmv<V, false, true>, g, b, s, a);
    else launch(ls, k_pcg_fspmv<V, false, false>, g, b, s, a);
  }
}
void launch_pcg_fspmv(const LaunchSet& ls, hipStream_t s) {
  switch (ls.spmv_variant) {
    case 1: launch_fspmv_v<1>(ls, s); break;
    case 2: launch_fspmv_v<2>(ls, s); break;
    case 3: launch_fspmv_v<3>(ls, s); break;
    case 4: launch_fspmv_v<4>(ls, s); break;
    case 5: launch_fspmv_v<5>(ls, s); break;
    default: launch_fspmv_v<0>(ls, s); break;
  }
}

template <int V>
static void launch_ccg_spmv_v(const LaunchSet& ls, hipStream_t s) {
  const TileArgs a = tile_args(ls);
  const dim3 g(ls.n_tiles), b(kTile);
  if (ls.one) launch(ls, k_ccg_spmv<V, true>, g, b, s, a);
  else launch(ls, k_ccg_spmv<V, false>, g, b, s, a);
}
void launch_ccg_spmv(const LaunchSet& ls, hipStream_t s) {
  switch (ls.spmv_variant) {
    case 1: launch_ccg_spmv_v<1>(ls, s); break;
    case 2: launch_ccg_spmv_v<2>(ls, s); break;
    case 3: launch_ccg_spmv_v<3>(ls, s); break;
    case 4: launch_ccg_spmv_v<4>(ls, s); break;
    case 5: launch_ccg_spmv_v<5>(ls, s); break;
    default: launch_ccg_spmv_v<0>(ls, s); break;
  }
}
void launch_ccg_update(const LaunchSet& ls, hipStream_t s) {
  const TileArgs a = tile_args(ls);
  const dim3 b(kTile);
  if (ls.one && ls.update_groups) {
    launch(ls, k_ccg_update<true, 2, false>, dim3((ls.n_tiles + 1) / 2), b, s, a);
  } else if (ls.one) {
    launch(ls, k_ccg_update<true, 1, true>, dim3(ls.n_tiles), b, s, a);
  } else {
    launch(ls, k_ccg_update<false, 1, true>, dim3(ls.n_tiles), b, s, a);
  }
}

void launch_op_axpy(uint32_t n, float alpha, const float* x, float* y, hipStream_t s) {
  hipLaunchKernelGGL(k_op_axpy, dim3((n + 255) / 256), dim3(256), 0, s, n, alpha, x, y);
}
void launch_op_scale(uint32_t n, float beta, float* x, hipStream_t s) {
  hipLaunchKernelGGL(k_op_scale, dim3((n + 255) / 256), dim3(256), 0, s, n, beta, x);
}
void launch_op_dot(uint32_t n, const float* a, const float* b, float* part, float* scratch,
                   float* out, hipStream_t s) {
  const uint32_t groups = (n + kTile - 1) / kTile;
  hipLaunchKernelGGL(k_op_dot_partials, dim3(groups), dim3(kTile), 0, s, n, a, b, part);
  hipLaunchKernelGGL(k_op_dot_reduce, dim3(1), dim3(kTile), 0, s, part, groups, scratch, out);
}
void launch_op_spmv(uint32_t n, const uint32_t* rp, const uint32_t* col, const float* val,
                    const float* x, float* y, hipStream_t s) {
  hipLaunchKernelGGL(k_op_spmv, dim3((n + kTile - 1) / kTile), dim3(kTile), 0, s, n, rp, col,
                     val, x, y);
}
void launch_op_block_jacobi(uint32_t n, uint32_t nb, const float* lu, const uint32_t* bs,
                            const float* r, float* z, hipStream_t s) {
  hipLaunchKernelGGL(k_op_block_jacobi, dim3((n + kTile - 1) / kTile), dim3(kTile), 0, s, n, nb,
                     lu, bs, r, z);
}
void launch_op_update_scalars(float* slots, uint32_t a, uint32_t b, uint32_t c, uint32_t d,
                              uint32_t e, uint32_t f, hipStream_t s) {
  hipLaunchKernelGGL(k_op_update_scalars, dim3(1), dim3(1), 0, s, slots, a, b, c, d, e, f);
}

}  // namespace hbj
