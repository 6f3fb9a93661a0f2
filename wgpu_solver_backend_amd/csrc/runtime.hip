// runtime.hip -- host runtime behind the C ABI: context, device-resident
// systems, the iteration engine that drives the PCG kernels, single-op
// executors, iterative refinement and batches.
//
// The reference drives one PCG iteration per submit and blocks on a scalar
// read-back every iteration (compute.rs:336-489).  Here every decision of an
// iteration (alpha, beta, breakdown, stop test) is taken on the device, and
// the host only enqueues chunks of iterations and polls a status word.  The
// poll of chunk c-1 overlaps the execution of chunk c, so the GPU never
// waits on the host; iterations enqueued after a system stops exit at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "hbj_internal.hpp"

#include "engine.hpp"

using namespace hbj;

static_assert(HBJ_SPMV_VARIANTS == kSpmvVariantCount, "public and internal SpMV variant counts");
static_assert(HBJ_ITER_KERNELS == Engine::kKernels, "public and internal kernels per iteration");

namespace {

void status_message(int32_t st, uint32_t max_iter) {
  switch (st) {
    case HBJ_E_BREAKDOWN_PAP:
      set_error("PCG(BlockJacobiGpu): dot(p,Ap) is zero (breakdown)");  // compute.rs:474
      break;
    case HBJ_E_BREAKDOWN_RZ:
      set_error("PCG(BlockJacobiGpu): rz_old is zero (breakdown)");  // compute.rs:477
      break;
    case HBJ_E_NOT_CONVERGED:
      set_error("PCG(BlockJacobiGpu): did not converge in %u iterations", max_iter);  // :492
      break;
    default:
      break;
  }
}

void fill_result(const SysState& s, bool ir, double ms, hbj_result* r) {
  if (!r) return;
  std::memset(r, 0, sizeof *r);
  r->status = s.status;
  r->iterations = ir ? s.total_iter : s.iter;
  r->outer_iterations = ir ? s.outer : 0;
  r->rr = s.rr;
  r->b_norm2 = s.b_norm2;
  r->rel_res_recurrence = s.b_norm2 > 0 ? std::sqrt((double)s.rr / (double)s.b_norm2) : 0.0;
  r->rel_res_true_fp64 = ir ? (s.bb64 > 0 ? std::sqrt(s.rr64 / s.bb64) : 0.0) : NAN;
  r->solve_ms = ms;
}

}  // namespace

extern "C" {

const char* hbj_version(void) { return "hbj 0.1 (gfx950)"; }

hbj_status hbj_ctx_create(int device, hbj_ctx** out) {
  clear_error();
  if (!out) {
    set_error("hbj_ctx_create: out is null");
    return HBJ_E_INVALID;
  }
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    (void)hipGetLastError();
    set_error("no HIP device available (hipGetDeviceCount: %s, count %d)", hipGetErrorString(e),
              count);
    return HBJ_E_NO_DEVICE;
  }
  if (device < 0 || device >= count) {
    set_error("device %d out of range (%d devices)", device, count);
    return HBJ_E_INVALID;
  }
  auto* c = new hbj_ctx();
  c->device = device;
  e = hipSetDevice(device);
  if (e == hipSuccess) e = hipGetDeviceProperties(&c->prop, device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error("device init failed: %s", hipGetErrorString(e));
    delete c;
    return HBJ_E_DEVICE;
  }
  // the device address of the no-alloc dot's scratch array (once per device)
  const hbj_status ns = noalloc_init(device);
  if (ns != HBJ_OK) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return ns;
  }
  *out = c;
  return HBJ_OK;
}

void hbj_ctx_destroy(hbj_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

hbj_status hbj_ctx_describe(hbj_ctx* ctx, char* buf, size_t len) {
  if (!ctx || !buf || len == 0) {
    set_error("hbj_ctx_describe: bad argument");
    return HBJ_E_INVALID;
  }
  const hipDeviceProp_t& p = ctx->prop;
  snprintf(buf, len,
           "{\"adapter_name\": \"%s\", \"backend\": \"HIP\", \"device_type\": \"DiscreteGpu\", "
           "\"arch\": \"%s\", \"device\": %d, \"compute_units\": %d, \"hbm_bytes\": %zu, "
           "\"l2_bytes\": %d, \"clock_khz\": %d}",
           p.name, p.gcnArchName, ctx->device, p.multiProcessorCount, (size_t)p.totalGlobalMem,
           p.l2CacheSize, p.clockRate);
  return HBJ_OK;
}

// ------------------------------------------------------------ calibration

hbj_status hbj_stream_copy(hbj_ctx* ctx, uint64_t bytes, uint32_t reps, double* avg_us) {
  clear_error();
  if (!ctx || !avg_us || bytes < 16 || reps == 0) {
    set_error("hbj_stream_copy: bad argument");
    return HBJ_E_INVALID;
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  const uint64_t n4 = bytes / 16;
  float *a = nullptr, *b = nullptr;
  hbj_status st = dalloc(&a, n4 * 4);
  if (st == HBJ_OK) st = dalloc(&b, n4 * 4);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (st == HBJ_OK) {
    HIPCHK(hipMemset(a, 0, n4 * 16));
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    launch_stream_copy(a, b, n4, ctx->stream);  // warm
    HIPCHK(hipEventRecord(e0, ctx->stream));
    for (uint32_t r = 0; r < reps; ++r) launch_stream_copy(a, b, n4, ctx->stream);
    HIPCHK(hipEventRecord(e1, ctx->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *avg_us = ms * 1000.0 / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  dfree(a);
  dfree(b);
  return st;
}

// ----------------------------------------------------------- single-op API

hbj_status hbj_op_axpy(hbj_ctx* ctx, uint32_t n, float alpha, const float* x, float* y) {
  clear_error();
  if (!ctx || (n && (!x || !y))) { set_error("hbj_op_axpy: bad argument"); return HBJ_E_INVALID; }
  if (n == 0) return HBJ_OK;
  std::lock_guard<std::mutex> g(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  float *dx = nullptr, *dy = nullptr;
  hbj_status st = dalloc(&dx, n);
  if (st == HBJ_OK) st = dalloc(&dy, n);
  if (st == HBJ_OK) {
    HIPCHK(hipMemcpy(dx, x, 4ull * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dy, y, 4ull * n, hipMemcpyHostToDevice));
    launch_op_axpy(n, alpha, dx, dy, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(y, dy, 4ull * n, hipMemcpyDeviceToHost));
  }
  dfree(dx);
  dfree(dy);
  return st;
}

hbj_status hbj_op_scale(hbj_ctx* ctx, uint32_t n, float beta, float* x) {
  clear_error();
  if (!ctx || (n && !x)) { set_error("hbj_op_scale: bad argument"); return HBJ_E_INVALID; }
  if (n == 0) return HBJ_OK;
  std::lock_guard<std::mutex> g(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  float* dx = nullptr;
  CHK(dalloc(&dx, n));
  HIPCHK(hipMemcpy(dx, x, 4ull * n, hipMemcpyHostToDevice));
  launch_op_scale(n, beta, dx, ctx->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(x, dx, 4ull * n, hipMemcpyDeviceToHost));
  dfree(dx);
  return HBJ_OK;
}

hbj_status hbj_op_dot(hbj_ctx* ctx, uint32_t n, const float* a, const float* b, float* out) {
  clear_error();
  if (!ctx || !out || (n && (!a || !b))) { set_error("hbj_op_dot: bad argument"); return HBJ_E_INVALID; }
  if (n == 0) {  // dot_scalar_exec.rs:164-174
    *out = 0.0f;
    return HBJ_OK;
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  const uint32_t groups = (n + kTile - 1) / kTile;
  float *da = nullptr, *db = nullptr, *dout = nullptr, *scratch = nullptr, *part = nullptr;
  hbj_status st = dalloc(&da, n);
  if (st == HBJ_OK) st = dalloc(&db, n);
  if (st == HBJ_OK) st = dalloc(&part, groups);
  if (st == HBJ_OK) st = dalloc(&scratch, (groups + 255) / 256);
  if (st == HBJ_OK) st = dalloc(&dout, 1);
  if (st == HBJ_OK) {
    HIPCHK(hipMemcpy(da, a, 4ull * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(db, b, 4ull * n, hipMemcpyHostToDevice));
    launch_op_dot(n, da, db, part, scratch, dout, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(out, dout, 4, hipMemcpyDeviceToHost));
  }
  dfree(da); dfree(db); dfree(part); dfree(scratch); dfree(dout);
  return st;
}

hbj_status hbj_op_spmv(hbj_ctx* ctx, uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                       const uint32_t* col_idx, const float* values, const float* x, float* y) {
  clear_error();
  if (!ctx || !row_ptr || (n && (!x || !y)) || (nnz && (!col_idx || !values))) {
    set_error("hbj_op_spmv: bad argument");
    return HBJ_E_INVALID;
  }
  if (row_ptr[n] != nnz) {
    set_error("CSR row_ptr must end with nnz (%u), got Some(%u)", nnz, row_ptr[n]);
    return HBJ_E_INVALID;
  }
  for (uint32_t k = 0; k < nnz; ++k)
    if (col_idx[k] >= n) {
      set_error("CSR col_idx contains out-of-range column %u for n_cols=%u", col_idx[k], n);
      return HBJ_E_INVALID;
    }
  if (n == 0) return HBJ_OK;
  std::lock_guard<std::mutex> g(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  uint32_t *drp = nullptr, *dcol = nullptr;
  float *dval = nullptr, *dx = nullptr, *dy = nullptr;
  hbj_status st = dalloc(&drp, n + 1);
  if (st == HBJ_OK) st = dalloc(&dcol, nnz + 4ull);
  if (st == HBJ_OK) st = dalloc(&dval, nnz + 4ull);
  if (st == HBJ_OK) st = dalloc(&dx, n);
  if (st == HBJ_OK) st = dalloc(&dy, n);
  if (st == HBJ_OK) {
    HIPCHK(hipMemcpy(drp, row_ptr, 4ull * (n + 1), hipMemcpyHostToDevice));
    if (nnz) {
      HIPCHK(hipMemcpy(dcol, col_idx, 4ull * nnz, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(dval, values, 4ull * nnz, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(dx, x, 4ull * n, hipMemcpyHostToDevice));
    launch_op_spmv(n, drp, dcol, dval, dx, dy, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(y, dy, 4ull * n, hipMemcpyDeviceToHost));
  }
  dfree(drp); dfree(dcol); dfree(dval); dfree(dx); dfree(dy);
  return st;
}

hbj_status hbj_op_block_jacobi(hbj_ctx* ctx, uint32_t n, const float* lu, const uint32_t* bs,
                               uint32_t n_starts, const float* r, float* z) {
  clear_error();
  if (!ctx || !bs || (n && (!r || !z)) || (n_starts > 1 && !lu)) {
    set_error("hbj_op_block_jacobi: bad argument");
    return HBJ_E_INVALID;
  }
  if (n == 0 || n_starts < 2) return HBJ_OK;
  // the tile search needs a partition of [0, n]
  if (bs[0] != 0 || bs[n_starts - 1] != n) {
    set_error("block_starts must start with 0 and end with n (%u)", n);
    return HBJ_E_INVALID;
  }
  for (uint32_t w = 0; w + 1 < n_starts; ++w)
    if (bs[w + 1] <= bs[w]) {
      set_error("block_starts must be strictly increasing");
      return HBJ_E_INVALID;
    }
  const uint32_t nb = n_starts - 1;
  std::lock_guard<std::mutex> g(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  float *dlu = nullptr, *dr = nullptr, *dz = nullptr;
  uint32_t* dbs = nullptr;
  hbj_status st = dalloc(&dlu, 36ull * nb);
  if (st == HBJ_OK) st = dalloc(&dbs, n_starts);
  if (st == HBJ_OK) st = dalloc(&dr, n);
  if (st == HBJ_OK) st = dalloc(&dz, n);
  if (st == HBJ_OK) {
    HIPCHK(hipMemcpy(dlu, lu, 144ull * nb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dbs, bs, 4ull * n_starts, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dr, r, 4ull * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dz, z, 4ull * n, hipMemcpyHostToDevice));
    launch_op_block_jacobi(n, nb, dlu, dbs, dr, dz, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(z, dz, 4ull * n, hipMemcpyDeviceToHost));
  }
  dfree(dlu); dfree(dbs); dfree(dr); dfree(dz);
  return st;
}

hbj_status hbj_op_update_scalars(hbj_ctx* ctx, float* slots, uint32_t n_slots, uint32_t a,
                                 uint32_t b, uint32_t c, uint32_t d, uint32_t e, uint32_t f) {
  clear_error();
  if (!ctx || !slots) { set_error("hbj_op_update_scalars: bad argument"); return HBJ_E_INVALID; }
  if (a >= n_slots || b >= n_slots || c >= n_slots || d >= n_slots || e >= n_slots || f >= n_slots) {
    set_error("PcgUpdateScalarsExecutor: slot index out of range");
    return HBJ_E_INVALID;
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  float* ds = nullptr;
  CHK(dalloc(&ds, n_slots));
  HIPCHK(hipMemcpy(ds, slots, 4ull * n_slots, hipMemcpyHostToDevice));
  launch_op_update_scalars(ds, a, b, c, d, e, f, ctx->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(slots, ds, 4ull * n_slots, hipMemcpyDeviceToHost));
  dfree(ds);
  return HBJ_OK;
}

// --------------------------------------------------------------- systems

hbj_status hbj_system_create(hbj_ctx* ctx, uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                             const uint32_t* col_idx, const float* values,
                             const uint32_t* block_starts, uint32_t n_starts, hbj_system** out) {
  return hbj::system_create(ctx, n, nnz, row_ptr, col_idx, values, block_starts, n_starts, false,
                            out);
}

}  // extern "C"

namespace hbj {

hbj_status prepare_image(uint32_t n, uint32_t nnz, const uint32_t* row_ptr, const uint32_t* col_idx,
                         const float* values, const uint32_t* block_starts, uint32_t n_starts,
                         bool lean, SystemImage& im) {
  if ((uint64_t)n > (uint64_t)kMaxTiles * kTile) {
    set_error("hbj_system_create: n = %u exceeds the per-system limit of %llu rows", n,
              (unsigned long long)kMaxTiles * kTile);
    return HBJ_E_INVALID;
  }
  im = SystemImage{};
  im.n = n;
  im.nnz = nnz;
  im.n_starts = n_starts;
  im.nb = n_starts - 1;
  im.n_tiles = (n + kTile - 1) / kTile;
  im.lean = lean;
  const uint32_t nb = im.nb, nt = im.n_tiles ? im.n_tiles : 1u;
  build_tile_local(n, row_ptr, col_idx, im.lidx, im.tloc, im.tcols);
  const bool sliced = build_sliced(n, row_ptr, values, im.lidx, im.tloc,
                                   env_u32("HBJ_SLICED_NZ", 1) != 0, im.sval, im.runs, im.shdr);
  // column lists made of whole groups of 6 columns, also kept as their groups
  // (staged from by the sliced SpMV): goff[t] = 1 + offset in tgrp, or 0
  std::vector<uint32_t> goff(im.tloc.size(), 0u);
  uint64_t staged_columns = 0;  // list entries the sliced SpMV reads
  if (sliced) {
    for (uint32_t t = 0; t < (uint32_t)im.tloc.size(); ++t) {
      const uint32_t off = im.tloc[t].x, cnt = im.tloc[t].y;
      bool whole = cnt > 0 && cnt % 6u == 0;
      for (uint32_t i = 0; whole && i < cnt; i += 6)
        whole = im.tcols[off + i] % 6u == 0 && im.tcols[off + i + 5] == im.tcols[off + i] + 5u;
      if (whole) {
        goff[t] = (uint32_t)im.tgrp.size() + 1u;
        for (uint32_t i = 0; i < cnt; i += 6) im.tgrp.push_back(im.tcols[off + i] / 6u);
        staged_columns += cnt / 6u;
      } else {
        staged_columns += cnt;
      }
    }
  }
  {
    uint64_t csr_entries = 0;
    for (uint32_t t = 0; t < (uint32_t)im.tloc.size(); ++t) {
      const uint32_t r0 = t * kTile, r1 = std::min<uint32_t>(r0 + kTile, n);
      const uint64_t e = row_ptr[r1] - row_ptr[r0];
      if (im.tloc[t].y == kLocalOff) {
        csr_entries += e;
      } else {
        im.local_tiles += 1;
        im.local_columns += im.tloc[t].y;
        im.spmv_layout_bytes += 6 * e + 4ull * im.tloc[t].y + sizeof(TileHdr);
      }
    }
    im.spmv_layout_bytes += 8 * csr_entries + 12ull * n + 4;
  }
  bool blocks6 = nb > 0 && block_starts[nb] == n && n <= 6ull * nb && n > 6ull * (nb - 1);
  for (uint32_t b = 0; b < nb; ++b) {
    if (block_starts[b + 1] - block_starts[b] > (uint32_t)kBlock) im.flags |= kFlagOversize;
    if (block_starts[b] != 6ull * b) blocks6 = false;
  }
  if (blocks6) im.flags |= kFlagBlocks6;
  if (sliced) {
    im.flags |= kFlagSliced;
    for (const SliceHdr& h : im.shdr) im.sliced_nz_waves += (h.nr & kSliceNz) ? 1u : 0u;
    im.spmv_sliced_bytes = 4ull * im.sval.size() + 4ull * im.runs.size() +
                           (sizeof(SliceHdr) * kWaves + sizeof(TileHdr)) * im.n_tiles +
                           4ull * staged_columns + 8ull * n;
  }
  im.thdr.resize(im.tloc.size());
  for (uint32_t t = 0; t < (uint32_t)im.tloc.size(); ++t) {
    const uint32_t r0 = t * kTile, r1 = std::min<uint32_t>(r0 + kTile, n);
    TileHdr& h = im.thdr[t];
    h.off = im.tloc[t].x;
    h.cnt = im.tloc[t].y;
    for (uint32_t k = 0; k < 4; ++k) h.eb[k] = row_ptr[std::min<uint32_t>(r0 + 64u * k, r1)];
    h.eb[4] = row_ptr[r1];
    h.goff = goff[t];
  }
  // per tile: the first block touching rows [256t, 256t+256) and how many do
  im.tile_blk.assign(nt, make_uint2(0u, 0u));
  for (uint32_t t = 0; t < im.n_tiles; ++t) {
    const uint32_t r0 = t * kTile, r1 = std::min<uint32_t>(r0 + kTile, n) - 1;
    const uint32_t bf = (uint32_t)(std::upper_bound(block_starts, block_starts + n_starts, r0) - block_starts) - 1;
    const uint32_t bl = (uint32_t)(std::upper_bound(block_starts, block_starts + n_starts, r1) - block_starts) - 1;
    im.tile_blk[t] = make_uint2(bf, bl - bf + 1);
  }
  if (lean) {
    // the LU blocks on the host (compute.rs:84-181; the device builder's
    // bits): a lean system may keep no CSR on the device to build them from
    im.lu.assign(36ull * (nb ? nb : 1u), 0.0f);
    const hbj_status st = hbj_build_lu_blocks_6(n, row_ptr, n + 1, col_idx, nnz, values, nnz,
                                                block_starts, n_starts, im.lu.data());
    if (st != HBJ_OK) return st;
    if (im.no_csr()) std::vector<uint16_t>().swap(im.lidx);  // host-only from here on
  }
  return HBJ_OK;
}

uint64_t lean_block_bytes(const SystemImage& im) {
  hbj_system probe;
  BlockLayout L;
  if (im.no_csr()) {
    uint64_t data_end = 0;
    probe.layout_lean(L, im, &data_end);
  } else {
    probe.layout(L, im.n, im.nnz, im.n_starts, im.tcols.size(), im.sliced() ? im.sval.size() : 0u,
                 im.sliced() ? im.runs.size() : 0u, im.tgrp.size(), true);
  }
  return (L.size + 255) & ~255ull;
}

// The fields of a new system from its image (no device memory yet).
static hbj_system* new_system(hbj_ctx* ctx, const SystemImage& im) {
  auto* s = new hbj_system();
  s->ctx = ctx;
  s->n = im.n;
  s->nnz = im.nnz;
  s->nb = im.nb;
  s->n_tiles = im.n_tiles;
  s->flags = im.flags;
  s->local_tiles = im.local_tiles;
  s->local_columns = im.local_columns;
  s->spmv_layout_bytes = im.spmv_layout_bytes;
  s->spmv_sliced_bytes = im.spmv_sliced_bytes;
  s->sliced_nz_waves = im.sliced_nz_waves;
  s->lean = im.no_csr();
  return s;
}

hbj_status system_create(hbj_ctx* ctx, uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                         const uint32_t* col_idx, const float* values, const uint32_t* block_starts,
                         uint32_t n_starts, bool with_ir, hbj_system** out) {
  clear_error();
  if (!ctx || !out || (nnz && !values)) {
    set_error("hbj_system_create: bad argument");
    return HBJ_E_INVALID;
  }
  *out = nullptr;
  CHK(validate_case(n, nnz, row_ptr, col_idx, block_starts, n_starts));
  SystemImage im;
  CHK(prepare_image(n, nnz, row_ptr, col_idx, values, block_starts, n_starts, false, im));
  const uint32_t nb = im.nb;
  hbj_system* s = new_system(ctx, im);
  auto fail = [&](hbj_status st) {
    s->release();
    delete s;
    return st;
  };
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) {
    set_error("hipSetDevice failed: %s", hipGetErrorString(e));
    return fail(HBJ_E_DEVICE);
  }
  const uint32_t nt = s->n_tiles ? s->n_tiles : 1;
  BlockLayout L;
  s->layout(L, n, nnz, n_starts, im.tcols.size(), im.sliced() ? im.sval.size() : 0u,
            im.sliced() ? im.runs.size() : 0u, im.tgrp.size(), with_ir);
  {
    hbj_status st = dalloc(&s->block, L.size);
    if (st != HBJ_OK) return fail(st);
  }
  s->block_bytes = L.size;
  s->bytes = L.size;
  L.assign(s->block);
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (bytes == 0 || e != hipSuccess) return;
    e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  };
  up(s->rp, row_ptr, 4ull * (n + 1));
  up(s->col, col_idx, 4ull * nnz);
  up(s->val, values, 4ull * nnz);
  up(s->bs, block_starts, 4ull * n_starts);
  up(s->lidx, im.lidx.data(), sizeof(uint16_t) * im.lidx.size());
  up(s->thdr, im.thdr.data(), sizeof(TileHdr) * im.thdr.size());
  up(s->tcols, im.tcols.data(), sizeof(uint32_t) * im.tcols.size());
  if (im.sliced()) {
    up(s->sval, im.sval.data(), sizeof(float) * im.sval.size());
    up(s->runs, im.runs.data(), sizeof(uint32_t) * im.runs.size());
    up(s->shdr, im.shdr.data(), sizeof(SliceHdr) * im.shdr.size());
    up(s->tgrp, im.tgrp.data(), sizeof(uint32_t) * im.tgrp.size());
  }
  up(s->tile_blk, im.tile_blk.data(), sizeof(uint2) * nt);
  // LU blocks on the device from the uploaded CSR (compute.rs:84-181; the
  // block starts were validated above, so the only error left is a zero
  // pivot, reported for the first failing block as the host builder does)
  if (e == hipSuccess && nb) {
    // the residual check's partials (>= 16 bytes) are free during setup
    unsigned long long* d_fail = reinterpret_cast<unsigned long long*>(s->resid2);
    unsigned long long h_fail = 0;
    e = hipMemset(d_fail, 0xFF, sizeof h_fail);
    if (e == hipSuccess) {
      launch_build_lu6(n, s->rp, s->col, s->val, s->bs, nb, s->lu, d_fail, nullptr);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(&h_fail, d_fail, sizeof h_fail, hipMemcpyDeviceToHost);
    if (e == hipSuccess && h_fail != ~0ull) {
      set_error("LU zero pivot at k = %u", (uint32_t)(h_fail & 0xFFFFFFFFu));
      return fail(HBJ_E_ZERO_PIVOT);
    }
  }
  // the padding past nnz is read by 16-byte loads: keep it zero (column 0)
  if (e == hipSuccess) e = hipMemset(s->col + nnz, 0, 16);
  if (e == hipSuccess) e = hipMemset(s->val + nnz, 0, 16);
  if (e == hipSuccess && n) e = hipMemset(s->z, 0, 4ull * n);  // wgpu zero-initialises buffers
  if (e == hipSuccess && n) e = hipMemset(s->x0, 0, 4ull * n);
  if (e == hipSuccess && n) e = hipMemset(s->b, 0, 4ull * n);
  if (with_ir && n)  // x64 starts at x0_64 = 0 and b64 = 0 until a right-hand side is set
    for (double* v : {s->x64, s->x0_64, s->b64})
      if (e == hipSuccess) e = hipMemset(v, 0, 8ull * n);
  if (e == hipSuccess) {
    SysDesc d1 = s->desc(false, false);
    e = hipMemcpy(s->d_desc1, &d1, sizeof d1, hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    set_error("upload failed: %s", hipGetErrorString(e));
    return fail(HBJ_E_DEVICE);
  }
  *out = s;
  return HBJ_OK;
}

hbj_status system_place(hbj_ctx* ctx, const SystemImage& im, const uint32_t* row_ptr,
                        const uint32_t* col_idx, const float* values, const uint32_t* block_starts,
                        const double* b64, DeviceArena* arena, uint64_t off, char* staging,
                        hipStream_t stream, hipEvent_t done, hbj_system** out, uint64_t* copied) {
  *out = nullptr;
  *copied = 0;
  hbj_system* s = new_system(ctx, im);
  const uint32_t n = im.n, nnz = im.nnz, nt = im.n_tiles ? im.n_tiles : 1u;
  BlockLayout L;
  uint64_t data_end = 0;
  if (im.no_csr()) {
    s->layout_lean(L, im, &data_end);
  } else {
    s->layout(L, n, nnz, im.n_starts, im.tcols.size(), im.sliced() ? im.sval.size() : 0u,
              im.sliced() ? im.runs.size() : 0u, im.tgrp.size(), true);
    data_end = L.size;  // the whole block goes through the staging image
    std::memset(staging, 0, L.size);
  }
  s->arena = arena;
  s->arena_off = off;
  s->block = arena->base + off;
  s->block_bytes = L.size;
  s->bytes = L.size;
  // the pieces' offsets in the staging image: the block's own layout
  L.assign(staging);
  auto put = [&](void* dst, const void* src, size_t bytes) {
    if (dst && bytes) std::memcpy(dst, src, bytes);
  };
  put(s->bs, block_starts, 4ull * im.n_starts);
  put(s->lu, im.lu.data(), 4ull * 36u * im.nb);
  put(s->tile_blk, im.tile_blk.data(), sizeof(uint2) * nt);
  put(s->thdr, im.thdr.data(), sizeof(TileHdr) * im.thdr.size());
  put(s->tcols, im.tcols.data(), sizeof(uint32_t) * im.tcols.size());
  if (im.sliced()) {
    put(s->sval, im.sval.data(), sizeof(float) * im.sval.size());
    put(s->runs, im.runs.data(), sizeof(uint32_t) * im.runs.size());
    put(s->shdr, im.shdr.data(), sizeof(SliceHdr) * im.shdr.size());
    put(s->tgrp, im.tgrp.data(), sizeof(uint32_t) * im.tgrp.size());
  }
  put(s->b64, b64, 8ull * n);
  if (!im.no_csr()) {  // the CSR path keeps the arrays (and their zero padding)
    put(s->rp, row_ptr, 4ull * (n + 1));
    put(s->col, col_idx, 4ull * nnz);
    put(s->val, values, 4ull * nnz);
    put(s->lidx, im.lidx.data(), sizeof(uint16_t) * im.lidx.size());
  }
  // the device block: the same offsets from its base
  L.assign(s->block);
  if (!im.no_csr()) {  // the FP64 check's descriptor (device pointers), in the image
    const SysDesc d1 = s->desc(false, false);
    std::memcpy(staging + (reinterpret_cast<char*>(s->d_desc1) - s->block), &d1, sizeof d1);
  }
  hipError_t e = hipMemcpyAsync(s->block, staging, data_end, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess && L.size > data_end)
    e = hipMemsetAsync(s->block + data_end, 0, L.size - data_end, stream);
  if (e == hipSuccess) e = hipEventRecord(done, stream);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error("upload failed: %s", hipGetErrorString(e));
    s->arena = nullptr;  // the caller returns the block
    s->block = nullptr;
    delete s;
    return HBJ_E_DEVICE;
  }
  *out = s;
  *copied = data_end;
  return HBJ_OK;
}

}  // namespace hbj

extern "C" {

void hbj_system_destroy(hbj_system* s) {
  if (!s) return;
  (void)hipSetDevice(s->ctx->device);
  s->release();
  delete s;
}

uint64_t hbj_system_device_bytes(const hbj_system* s) { return s ? s->bytes : 0; }

hbj_status hbj_system_get_lu(const hbj_system* s, float* lu_out) {
  clear_error();
  if (!s || (s->nb && !lu_out)) { set_error("hbj_system_get_lu: bad argument"); return HBJ_E_INVALID; }
  if (s->nb == 0) return HBJ_OK;
  HIPCHK(hipSetDevice(s->ctx->device));
  HIPCHK(hipMemcpy(lu_out, s->lu, 144ull * s->nb, hipMemcpyDeviceToHost));
  return HBJ_OK;
}

hbj_status hbj_system_get_info(const hbj_system* s, hbj_system_info* out) {
  clear_error();
  if (!s || !out) { set_error("hbj_system_get_info: bad argument"); return HBJ_E_INVALID; }
  std::memset(out, 0, sizeof *out);
  out->n = s->n;
  out->nnz = s->nnz;
  out->n_blocks = s->nb;
  out->n_tiles = s->n_tiles;
  out->local_tiles = s->local_tiles;
  out->spmv_variant = s->spmv_variant >= 0 ? (uint32_t)s->spmv_variant
                                           : (s->eng_ready ? s->eng.spmv_variant : kSpmvDefault);
  out->local_columns = s->local_columns;
  out->device_bytes = s->bytes;
  out->spmv_csr_bytes = 8ull * s->nnz + 12ull * s->n + 4;
  out->spmv_layout_bytes = s->spmv_layout_bytes;
  out->spmv_sliced_bytes = s->spmv_sliced_bytes;
  const bool two = s->eng_ready ? s->eng.fused : s->two_launch();
  out->launches_per_iteration = (s->alg == 1 || two) ? 2u : 3u;
  out->sliced_nz_waves = s->sliced_nz_waves;
  return HBJ_OK;
}

hbj_status hbj_system_set_spmv_variant(hbj_system* s, uint32_t v) {
  clear_error();
  if (!s || v >= kSpmvVariantCount) {
    set_error("hbj_system_set_spmv_variant: need a system and a variant < %u", kSpmvVariantCount);
    return HBJ_E_INVALID;
  }
  s->spmv_variant = (int)v;
  if (s->eng_ready) s->eng.set_spmv_variant(v);
  return HBJ_OK;
}

hbj_status hbj_system_set_algorithm(hbj_system* s, uint32_t alg) {
  clear_error();
  if (!s || alg > HBJ_ALG_PCG_ONE_REDUCTION) {
    set_error("hbj_system_set_algorithm: need a system and an algorithm <= %d",
              HBJ_ALG_PCG_ONE_REDUCTION);
    return HBJ_E_INVALID;
  }
  if (alg == HBJ_ALG_PCG_ONE_REDUCTION) CHK(s->ensure_ccg_buffers());
  if (alg != s->alg) {
    s->alg = alg;
    s->eng_mode = -1;  // the device descriptor is written again (s vectors)
  }
  return HBJ_OK;
}

hbj_status hbj_system_autotune(hbj_system* s, uint32_t iters, uint32_t* best, double* us) {
  clear_error();
  if (!s || iters == 0) { set_error("hbj_system_autotune: need a system and iters > 0"); return HBJ_E_INVALID; }
  if (s->n == 0) {
    if (best) *best = kSpmvDefault;
    return HBJ_OK;
  }
  HIPCHK(hipSetDevice(s->ctx->device));
  CHK(s->ensure_engine(0));
  const uint32_t timing = s->eng.timing_every;
  s->eng.timing_every = 0;
  double best_us = 0;
  uint32_t bv = kSpmvDefault;
  for (uint32_t v = 0; v < kSpmvVariantCount; ++v) {
    s->eng.set_spmv_variant(v);
    double t = 0;
    for (int rep = 0; rep < 4; ++rep) {  // the first run captures the graph; best of 3
      HIPCHK(hipMemcpyAsync(s->x, s->x0, 4ull * s->n, hipMemcpyDeviceToDevice, s->eng.stream));
      if (s->flags & kFlagOversize) HIPCHK(hipMemsetAsync(s->z, 0, 4ull * s->n, s->eng.stream));
      // tolerances 0: runs `iters` iterations unless the residual is exactly 0
      SysState init = s->init_state(false, iters, 0.0f, 0.0f, nullptr, s->alg == 1, s->eng.fused), fin;
      double ms = 0;
      const hbj_status st = s->eng.run(&init, &fin, &ms);
      if (st != HBJ_OK) {
        s->eng.timing_every = timing;
        return st;
      }
      const double u = ms * 1e3 / (double)(fin.iter ? fin.iter : 1);
      if (rep == 1 || (rep > 1 && u < t)) t = u;
    }
    if (us) us[v] = t;
    if (v == 0 || t < best_us) {
      best_us = t;
      bv = v;
    }
  }
  s->eng.timing_every = timing;
  s->spmv_variant = (int)bv;
  s->eng.set_spmv_variant(bv);
  if (best) *best = bv;
  return HBJ_OK;
}

hbj_status hbj_system_set_rhs(hbj_system* s, const float* b, const float* x0) {
  clear_error();
  if (!s || (s->n && !b)) { set_error("hbj_system_set_rhs: bad argument"); return HBJ_E_INVALID; }
  if (s->n == 0) return HBJ_OK;
  HIPCHK(hipSetDevice(s->ctx->device));
  HIPCHK(hipMemcpy(s->b, b, 4ull * s->n, hipMemcpyHostToDevice));
  if (x0) HIPCHK(hipMemcpy(s->x0, x0, 4ull * s->n, hipMemcpyHostToDevice));
  else HIPCHK(hipMemset(s->x0, 0, 4ull * s->n));
  return HBJ_OK;
}

hbj_status hbj_system_get_x(hbj_system* s, float* x) {
  clear_error();
  if (!s || (s->n && !x)) { set_error("hbj_system_get_x: bad argument"); return HBJ_E_INVALID; }
  if (s->n == 0) return HBJ_OK;
  HIPCHK(hipSetDevice(s->ctx->device));
  HIPCHK(hipMemcpy(x, s->x, 4ull * s->n, hipMemcpyDeviceToHost));
  return HBJ_OK;
}

// spmv_exec.rs:122-147 on the system's resident matrix: the iteration's SpMV
// kernel (three-launch form) in its initial phase, q = A x with x as given.
hbj_status hbj_system_spmv(hbj_system* s, const float* x, float* y) {
  clear_error();
  if (!s || (s->n && (!x || !y))) { set_error("hbj_system_spmv: bad argument"); return HBJ_E_INVALID; }
  if (s->n == 0) return HBJ_OK;
  HIPCHK(hipSetDevice(s->ctx->device));
  CHK(s->ensure_engine(0));
  Engine& E = s->eng;
  hipStream_t st = E.stream;
  HIPCHK(hipMemcpyAsync(s->x, x, 4ull * s->n, hipMemcpyHostToDevice, st));
  std::memset(E.h_init, 0, sizeof(SysState));
  E.h_init[0].status = kRunning;
  E.h_init[0].phase = kPhInit;
  HIPCHK(hipMemcpyAsync(E.d_states, E.h_init, sizeof(SysState), hipMemcpyHostToDevice, st));
  LaunchSet l = E.ls();
  l.par = 0;
  launch_pcg_spmv(l, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(y, s->q, 4ull * s->n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return HBJ_OK;
}

hbj_status hbj_system_set_kernel_timing(hbj_system* s, uint32_t every) {
  if (!s) { set_error("hbj_system_set_kernel_timing: null system"); return HBJ_E_INVALID; }
  CHK(s->ensure_engine(s->eng_mode < 0 ? 0 : s->eng_mode));
  s->eng.timing_every = every;
  return HBJ_OK;
}

hbj_status hbj_system_kernel_timing(hbj_system* s, double* avg_us, uint64_t* counts) {
  if (!s || !avg_us || !counts) { set_error("hbj_system_kernel_timing: bad argument"); return HBJ_E_INVALID; }
  for (int k = 0; k < Engine::kKernels; ++k) {
    counts[k] = s->eng_ready ? s->eng.t_cnt[k] : 0;
    avg_us[k] = counts[k] ? s->eng.t_sum[k] / (double)counts[k] : 0.0;
    if (s->eng_ready) {
      s->eng.t_cnt[k] = 0;
      s->eng.t_sum[k] = 0;
    }
  }
  return HBJ_OK;
}

hbj_status hbj_system_solve_resident(hbj_system* s, uint32_t max_iter, float rel_tol,
                                     float abs_tol, hbj_result* res) {
  clear_error();
  if (!s) { set_error("hbj_system_solve_resident: null system"); return HBJ_E_INVALID; }
  if (s->n == 0) {  // ||b||^2 of an empty vector is 0 -> Ok(0)
    if (res) {
      std::memset(res, 0, sizeof *res);
      res->rel_res_true_fp64 = 0.0;
    }
    return HBJ_OK;
  }
  HIPCHK(hipSetDevice(s->ctx->device));
  CHK(s->ensure_engine(0));
  hipStream_t st = s->eng.stream;
  HIPCHK(hipMemcpyAsync(s->x, s->x0, 4ull * s->n, hipMemcpyDeviceToDevice, st));
  if (s->flags & kFlagOversize) HIPCHK(hipMemsetAsync(s->z, 0, 4ull * s->n, st));
  SysState init = s->init_state(false, max_iter, rel_tol, abs_tol, nullptr, s->alg == 1, s->eng.fused), fin;
  double ms = 0;
  CHK(s->eng.run(&init, &fin, &ms));
  fill_result(fin, false, ms, res);
  if (fin.status == HBJ_OK && res) {
    if (fin.b_norm2 == 0.0f) res->rel_res_true_fp64 = 0.0;
    else CHK(s->true_residual(st, &res->rel_res_true_fp64));
  }
  if (fin.status != HBJ_OK) status_message(fin.status, max_iter);
  return (hbj_status)fin.status;
}

hbj_status hbj_solve(hbj_system* s, const float* b, float* x_inout, uint32_t max_iter,
                     float rel_tol, float abs_tol, hbj_result* res) {
  clear_error();
  if (!s || (s->n && (!b || !x_inout))) {
    set_error("hbj_solve: bad argument");
    return HBJ_E_INVALID;
  }
  CHK(hbj_system_set_rhs(s, b, x_inout));
  hbj_status st = hbj_system_solve_resident(s, max_iter, rel_tol, abs_tol, res);
  if (st == HBJ_OK) {
    CHK(hbj_system_get_x(s, x_inout));
  }
  return st;
}

// ----------------------------------------------------------- mixed precision

hbj_status hbj_system_set_rhs64(hbj_system* s, const double* b, const double* x0) {
  clear_error();
  if (!s || (s->n && !b)) { set_error("hbj_system_set_rhs64: bad argument"); return HBJ_E_INVALID; }
  CHK(s->ensure_ir_buffers());
  if (s->n == 0) return HBJ_OK;
  HIPCHK(hipMemcpy(s->b64, b, 8ull * s->n, hipMemcpyHostToDevice));
  if (x0) HIPCHK(hipMemcpy(s->x0_64, x0, 8ull * s->n, hipMemcpyHostToDevice));
  else HIPCHK(hipMemset(s->x0_64, 0, 8ull * s->n));
  return HBJ_OK;
}

hbj_status hbj_system_get_x64(hbj_system* s, double* x) {
  clear_error();
  if (!s || (s->n && !x)) { set_error("hbj_system_get_x64: bad argument"); return HBJ_E_INVALID; }
  CHK(s->ensure_ir_buffers());
  if (s->n == 0) return HBJ_OK;
  HIPCHK(hipMemcpy(x, s->x64, 8ull * s->n, hipMemcpyDeviceToHost));
  return HBJ_OK;
}

static hbj_status check_ir(const hbj_ir_params* p) {
  if (!p || p->max_inner == 0 || !(p->inner_rel_tol > 0.0f) || !(p->target_rel_res > 0.0)) {
    set_error("hbj_ir_params: max_inner > 0, inner_rel_tol > 0 and target_rel_res > 0 required");
    return HBJ_E_INVALID;
  }
  return HBJ_OK;
}

static hbj_status solve_ir_resident(hbj_system* s, const hbj_ir_params* p, hbj_result* res) {
  CHK(check_ir(p));
  CHK(s->ensure_ir_buffers());
  if (s->n == 0) {
    if (res) std::memset(res, 0, sizeof *res);
    return HBJ_OK;
  }
  HIPCHK(hipSetDevice(s->ctx->device));
  CHK(s->ensure_engine(1));
  hipStream_t st = s->eng.stream;
  HIPCHK(hipMemcpyAsync(s->x64, s->x0_64, 8ull * s->n, hipMemcpyDeviceToDevice, st));
  if (s->flags & kFlagOversize) HIPCHK(hipMemsetAsync(s->z, 0, 4ull * s->n, st));
  SysState init = s->init_state(true, 0, 0, 0, p, false, s->eng.fused), fin;
  double ms = 0;
  CHK(s->eng.run(&init, &fin, &ms));
  fill_result(fin, true, ms, res);
  if (fin.status != HBJ_OK) {
    if (fin.status == HBJ_E_NOT_CONVERGED)
      set_error("IR: relative residual %.3e above target %.3e after %u refinement steps",
                fin.bb64 > 0 ? std::sqrt(fin.rr64 / fin.bb64) : 0.0, p->target_rel_res, fin.outer);
    else
      status_message(fin.status, p->max_inner);
  }
  return (hbj_status)fin.status;
}

hbj_status hbj_solve_ir(hbj_system* s, const double* b, double* x_inout, const hbj_ir_params* p,
                        hbj_result* res) {
  clear_error();
  if (!s || (s->n && (!b || !x_inout))) {
    set_error("hbj_solve_ir: bad argument");
    return HBJ_E_INVALID;
  }
  CHK(check_ir(p));
  CHK(hbj_system_set_rhs64(s, b, x_inout));
  hbj_status st = solve_ir_resident(s, p, res);
  if (st == HBJ_OK) CHK(hbj_system_get_x64(s, x_inout));
  return st;
}

}  // extern "C"

// ------------------------------------------------------------------ batch

struct hbj_batch {
  hbj_ctx* ctx = nullptr;
  std::vector<hbj_system*> sys;
  // The systems in up to two tables, each with its own engine and stream,
  // run side by side (the second on a helper thread): every kernel launch
  // of one table ends in a tail of few workgroups and a launch gap, which
  // the other table's launches fill.  Measured on the 256-system batch
  // (sides 4-24, refinement to 1e-8): 906 solves/s in one table, 998 in
  // two, 852 in three.  Each system's results do not depend on its table.
  // A batch with kernel timing on runs one table, so that each sampled
  // launch has the GPU to itself.
  static constexpr int kMaxTables = 2;
  std::unique_ptr<Engine> eng[kMaxTables];
  std::vector<uint32_t> members[kMaxTables];  // each table's systems (batch indices, ascending)
  std::vector<int> table_of;                  // system -> table
  std::vector<uint32_t> slot_of;              // system -> its slot in its table
  int n_tab = 0;
  int built_for = 0;  // the tables_wanted() the tables were built for
  int mode = -1;
  uint32_t n_tiles = 0;
  uint32_t timing_every = 0;
  uint32_t variant = kSpmvDefault;
  bool variant_set = false;

  int tables_wanted() const {
    if (timing_every || sys.size() < 2) return 1;
    const uint32_t t = env_u32("HBJ_BATCH_TABLES", 2);
    return t < 1 ? 1 : (t > (uint32_t)kMaxTables ? kMaxTables : (int)t);
  }
  Engine& engine_of(uint32_t i) { return *eng[table_of[i]]; }

  void release() {
    for (int e = 0; e < kMaxTables; ++e) {
      if (eng[e]) eng[e]->destroy();
      eng[e].reset();
    }
    n_tab = 0;
    mode = -1;
  }

  // Tables balanced by tiles: systems by decreasing size, each to the table
  // with fewer tiles so far.
  void partition(int k) {
    for (int e = 0; e < kMaxTables; ++e) members[e].clear();
    table_of.assign(sys.size(), 0);
    slot_of.assign(sys.size(), 0);
    std::vector<uint32_t> order(sys.size());
    for (uint32_t i = 0; i < sys.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return sys[x]->n_tiles > sys[y]->n_tiles; });
    uint64_t load[kMaxTables] = {0, 0};
    for (uint32_t i : order) {
      int e = 0;
      for (int f = 1; f < k; ++f)
        if (load[f] < load[e]) e = f;
      table_of[i] = e;
      load[e] += sys[i]->n_tiles;
    }
    for (uint32_t i = 0; i < sys.size(); ++i) {
      slot_of[i] = (uint32_t)members[table_of[i]].size();
      members[table_of[i]].push_back(i);
    }
  }

  hbj_status ensure(int m) {
    const int asked = tables_wanted();
    if (n_tab == 0 || built_for != asked) {
      release();
      int want = asked;
      partition(want);
      if (want > 1 && members[want - 1].empty()) {  // systems without rows only
        want = 1;
        partition(1);
      }
      built_for = asked;
      for (int e = 0; e < want; ++e) {
        eng[e].reset(new Engine());
        uint32_t tiles = 0;
        for (uint32_t i : members[e]) tiles += sys[i]->n_tiles;
        CHK(eng[e]->init(ctx->device, (uint32_t)members[e].size(), tiles ? tiles : 1));
        // the two-launch iteration when every system's partials fit the
        // prefetch, for a table of one system or when HBJ_FUSED=1 asks for it
        bool fused = members[e].size() == 1 ? fused_mode() : fused_tables();
        for (uint32_t i : members[e]) fused = fused && sys[i]->two_launch();
        eng[e]->fused = fused;
        if (variant_set) eng[e]->set_spmv_variant(variant);
      }
      n_tab = want;
      mode = -1;
    }
    if (mode != m) {
      for (int e = 0; e < n_tab; ++e) {
        std::vector<SysDesc> descs;
        std::vector<uint2> tiles;
        for (uint32_t j = 0; j < members[e].size(); ++j) {
          const hbj_system* s = sys[members[e][j]];
          descs.push_back(s->desc(m == 1, eng[e]->fused));
          for (uint32_t t = 0; t < s->n_tiles; ++t) tiles.push_back(make_uint2(j, t));
        }
        const bool empty = tiles.empty();
        if (empty) tiles.push_back(make_uint2(0, 0));
        CHK(eng[e]->set_table(descs, tiles));
        if (empty) eng[e]->n_tiles = eng[e]->n_ftiles = 0;
      }
      mode = m;
    }
    for (int e = 0; e < n_tab; ++e) eng[e]->timing_every = timing_every;
    return HBJ_OK;
  }

  void set_variant(uint32_t v) {
    variant = v;
    variant_set = true;
    for (int e = 0; e < n_tab; ++e) eng[e]->set_spmv_variant(v);
  }

  // Runs every table until all its systems have final states: init[i] and
  // fin[i] are system i's, done[i] its time to solution; *ms the longest
  // table's device time.  The tables run side by side.
  hbj_status run_all(const std::vector<SysState>& init, std::vector<SysState>& fin,
                     std::vector<double>& done, double* ms) {
    std::vector<SysState> in[kMaxTables], out[kMaxTables];
    for (int e = 0; e < n_tab; ++e) {
      for (uint32_t i : members[e]) in[e].push_back(init[i]);
      out[e].resize(members[e].size());
    }
    double tms[kMaxTables] = {0, 0};
    hbj_status st[kMaxTables] = {HBJ_OK, HBJ_OK};
    std::string err1;
    std::thread helper;
    if (n_tab > 1 && !members[1].empty())
      helper = std::thread([&]() {
        st[1] = eng[1]->run(in[1].data(), out[1].data(), &tms[1]);
        if (st[1] != HBJ_OK) err1 = hbj_last_error();
      });
    if (!members[0].empty()) st[0] = eng[0]->run(in[0].data(), out[0].data(), &tms[0]);
    if (helper.joinable()) helper.join();
    if (st[0] != HBJ_OK) return st[0];
    if (st[1] != HBJ_OK) {
      set_error("%s", err1.c_str());
      return st[1];
    }
    fin.resize(sys.size());
    done.assign(sys.size(), 0.0);
    for (int e = 0; e < n_tab; ++e)
      for (uint32_t j = 0; j < members[e].size(); ++j) {
        fin[members[e][j]] = out[e][j];
        done[members[e][j]] = eng[e]->done_ms[j];
      }
    if (ms) *ms = tms[0] > tms[1] ? tms[0] : tms[1];
    return HBJ_OK;
  }
};

extern "C" {

hbj_status hbj_batch_create(hbj_ctx* ctx, hbj_system* const* systems, uint32_t count,
                            hbj_batch** out) {
  clear_error();
  if (!ctx || !systems || !out || count == 0) {
    set_error("hbj_batch_create: bad argument");
    return HBJ_E_INVALID;
  }
  auto* b = new hbj_batch();
  b->ctx = ctx;
  for (uint32_t i = 0; i < count; ++i) {
    if (!systems[i] || systems[i]->ctx != ctx) {
      delete b;
      set_error("hbj_batch_create: system %u is null or on another context", i);
      return HBJ_E_INVALID;
    }
    b->sys.push_back(systems[i]);
    b->n_tiles += systems[i]->n_tiles;
  }
  *out = b;
  return HBJ_OK;
}

void hbj_batch_destroy(hbj_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  b->release();
  delete b;
}

hbj_status hbj_batch_solve(hbj_batch* b, uint32_t max_iter, float rel_tol, float abs_tol,
                           hbj_result* results) {
  clear_error();
  if (!b || !results) { set_error("hbj_batch_solve: bad argument"); return HBJ_E_INVALID; }
  HIPCHK(hipSetDevice(b->ctx->device));
  CHK(b->ensure(0));
  const uint32_t S = (uint32_t)b->sys.size();
  std::vector<SysState> init(S), fin;
  std::vector<double> done;
  for (uint32_t i = 0; i < S; ++i) {
    hbj_system* s = b->sys[i];
    Engine& E = b->engine_of(i);
    init[i] = s->init_state(false, max_iter, rel_tol, abs_tol, nullptr, false, E.fused);
    if (s->n == 0) {
      init[i].status = HBJ_OK;
      continue;
    }
    HIPCHK(hipMemcpyAsync(s->x, s->x0, 4ull * s->n, hipMemcpyDeviceToDevice, E.stream));
    if (s->flags & kFlagOversize) HIPCHK(hipMemsetAsync(s->z, 0, 4ull * s->n, E.stream));
  }
  double ms = 0;
  CHK(b->run_all(init, fin, done, &ms));
  hbj_status worst = HBJ_OK;
  for (uint32_t i = 0; i < S; ++i) {
    fill_result(fin[i], false, done[i], &results[i]);
    hbj_system* s = b->sys[i];
    if (fin[i].status == HBJ_OK && s->n) {
      if (fin[i].b_norm2 == 0.0f) results[i].rel_res_true_fp64 = 0.0;
      else CHK(s->true_residual(b->engine_of(i).stream, &results[i].rel_res_true_fp64));
    }
    if (fin[i].status != HBJ_OK && worst == HBJ_OK) {
      worst = (hbj_status)fin[i].status;
      status_message(fin[i].status, max_iter);
    }
  }
  return worst;
}

hbj_status hbj_batch_solve_ir(hbj_batch* b, const hbj_ir_params* p, hbj_result* results) {
  clear_error();
  if (!b || !results) { set_error("hbj_batch_solve_ir: bad argument"); return HBJ_E_INVALID; }
  CHK(check_ir(p));
  HIPCHK(hipSetDevice(b->ctx->device));
  for (auto* s : b->sys) CHK(s->ensure_ir_buffers());
  if (b->mode != 1) b->mode = -1;
  CHK(b->ensure(1));
  const uint32_t S = (uint32_t)b->sys.size();
  std::vector<SysState> init(S), fin;
  std::vector<double> done;
  for (uint32_t i = 0; i < S; ++i) {
    hbj_system* s = b->sys[i];
    Engine& E = b->engine_of(i);
    init[i] = s->init_state(true, 0, 0, 0, p, false, E.fused);
    if (s->n == 0) {
      init[i].status = HBJ_OK;
      continue;
    }
    HIPCHK(hipMemcpyAsync(s->x64, s->x0_64, 8ull * s->n, hipMemcpyDeviceToDevice, E.stream));
    if (s->flags & kFlagOversize) HIPCHK(hipMemsetAsync(s->z, 0, 4ull * s->n, E.stream));
  }
  double ms = 0;
  CHK(b->run_all(init, fin, done, &ms));
  hbj_status worst = HBJ_OK;
  for (uint32_t i = 0; i < S; ++i) {
    fill_result(fin[i], true, done[i], &results[i]);
    if (fin[i].status != HBJ_OK && worst == HBJ_OK) {
      worst = (hbj_status)fin[i].status;
      status_message(fin[i].status, p->max_inner);
    }
  }
  return worst;
}

hbj_status hbj_batch_set_kernel_timing(hbj_batch* b, uint32_t every) {
  clear_error();
  if (!b) { set_error("hbj_batch_set_kernel_timing: null batch"); return HBJ_E_INVALID; }
  HIPCHK(hipSetDevice(b->ctx->device));
  b->timing_every = every;  // the next solve runs one table while this is on
  return HBJ_OK;
}

hbj_status hbj_batch_kernel_samples(hbj_batch* b, double* rec, uint64_t cap, uint64_t* n_out,
                                    uint32_t* sets) {
  clear_error();
  if (!b || !n_out || (cap && !rec)) {
    set_error("hbj_batch_kernel_samples: bad argument");
    return HBJ_E_INVALID;
  }
  // samples come from a solve with timing on, which ran one table
  const bool one = b->n_tab == 1 && b->eng[0];
  const Engine* E = one ? b->eng[0].get() : nullptr;
  const uint64_t n = one ? (uint64_t)E->t_last.size() : 0u;
  *n_out = n;
  constexpr uint32_t kRec = Engine::kKernels + 1;
  for (uint64_t i = 0; i < n && i < cap; ++i)
    for (uint32_t k = 0; k < kRec; ++k) rec[i * kRec + k] = E->t_last[i][k];
  if (sets)
    for (uint32_t i = 0; i < (uint32_t)b->sys.size(); ++i) {
      const uint32_t j = one && i < b->slot_of.size() ? b->slot_of[i] : 0u;
      sets[i] = one && j < E->sets_last.size() ? E->sets_last[j] : 0u;
    }
  return HBJ_OK;
}

hbj_status hbj_batch_autotune(hbj_batch* b, uint32_t iters, int ir, uint32_t* best, double* us) {
  clear_error();
  if (!b || iters == 0) { set_error("hbj_batch_autotune: need a batch and iters > 0"); return HBJ_E_INVALID; }
  HIPCHK(hipSetDevice(b->ctx->device));
  if (ir) {
    for (auto* s : b->sys) CHK(s->ensure_ir_buffers());
    if (b->mode != 1) b->mode = -1;
  }
  // timed in the form the solves run: without kernel timing
  const uint32_t timing = b->timing_every;
  b->timing_every = 0;
  const hbj_status es = b->ensure(ir ? 1 : 0);
  if (es != HBJ_OK) {
    b->timing_every = timing;
    return es;
  }
  const uint32_t S = (uint32_t)b->sys.size();
  // every system runs exactly `iters` inner iterations: tolerances that
  // square to 0 (refinement: one outer step, then out of steps)
  hbj_ir_params p{};
  p.max_outer = 1;
  p.max_inner = iters;
  p.inner_rel_tol = 1e-30f;
  p.target_rel_res = 1e-300;
  double best_us = 0;
  uint32_t bv = kSpmvDefault;
  std::vector<SysState> init(S), fin;
  std::vector<double> done;
  for (uint32_t v = 0; v < kSpmvVariantCount; ++v) {
    b->set_variant(v);
    double tv = 0;
    for (int rep = 0; rep < 4; ++rep) {  // the first run captures the graph; best of 3
      for (uint32_t i = 0; i < S; ++i) {
        hbj_system* s = b->sys[i];
        Engine& E = b->engine_of(i);
        init[i] = ir ? s->init_state(true, 0, 0, 0, &p, false, E.fused)
                     : s->init_state(false, iters, 0.0f, 0.0f, nullptr, false, E.fused);
        if (s->n == 0) {
          init[i].status = HBJ_OK;
          continue;
        }
        if (ir) HIPCHK(hipMemcpyAsync(s->x64, s->x0_64, 8ull * s->n, hipMemcpyDeviceToDevice, E.stream));
        else HIPCHK(hipMemcpyAsync(s->x, s->x0, 4ull * s->n, hipMemcpyDeviceToDevice, E.stream));
        if (s->flags & kFlagOversize) HIPCHK(hipMemsetAsync(s->z, 0, 4ull * s->n, E.stream));
      }
      double ms = 0;
      const hbj_status rs = b->run_all(init, fin, done, &ms);
      if (rs != HBJ_OK) {
        b->timing_every = timing;
        return rs;
      }
      uint32_t its = 1;
      for (uint32_t i = 0; i < S; ++i) {
        const uint32_t k = ir ? fin[i].total_iter : fin[i].iter;
        its = k > its ? k : its;
      }
      const double u = ms * 1e3 / (double)its;
      if (rep == 1 || (rep > 1 && u < tv)) tv = u;
    }
    if (us) us[v] = tv;
    if (v == 0 || tv < best_us) {
      best_us = tv;
      bv = v;
    }
  }
  b->timing_every = timing;
  b->set_variant(bv);
  if (best) *best = bv;
  return HBJ_OK;
}

}  // extern "C"
