/*
 * hbj.h -- C ABI of the MI355X-native block-Jacobi PCG solver.
 *
 * Drop-in for the solver API of RomanShushakov/wgpu_solver_backend
 * (crates/wgpu_solver_backend, Rust + WGSL).  Each entry point names the
 * reference interface it replaces (paths relative to the reference checkout).
 * All arguments are plain pointers and sizes; host buffers unless stated.
 *
 * Conventions (same as the reference, compute.rs:189-495):
 *   - inputs are borrowed and copied; the solution buffer is written only
 *     when the call returns HBJ_OK (x keeps x0 on every error);
 *   - errors are status codes plus hbj_last_error(), a thread-local text
 *     with the reference's own wording ("PCG(BlockJacobiGpu): dot(p,Ap) is
 *     zero (breakdown)", "LU zero pivot at k = 2", ...);
 *   - one in-flight solve per hbj_system / hbj_batch; a context may be
 *     shared by threads.
 */
#ifndef HBJ_H
#define HBJ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hbj_status {
  HBJ_OK = 0,
  HBJ_E_INVALID = 1,        /* bad sizes / CSR / block_starts (loaders.rs:48-184, compute.rs:94-131) */
  HBJ_E_ZERO_PIVOT = 2,     /* compute.rs:50-52 */
  HBJ_E_BREAKDOWN_PAP = 3,  /* compute.rs:473-475 */
  HBJ_E_BREAKDOWN_RZ = 4,   /* compute.rs:476-478 */
  HBJ_E_NOT_CONVERGED = 5,  /* compute.rs:491-494 */
  HBJ_E_DEVICE = 6,         /* HIP runtime error (the reference panics in readback.rs:42-46) */
  HBJ_E_OOM = 7,            /* device allocation failed / over the HBM budget */
  HBJ_E_NO_DEVICE = 8,      /* no HIP device (context.rs:66-73 request_adapter failure) */
  HBJ_E_REJECTED = 9        /* serving: request larger than the admission budget */
} hbj_status;

typedef struct hbj_ctx hbj_ctx;
typedef struct hbj_system hbj_system;
typedef struct hbj_batch hbj_batch;
typedef struct hbj_server hbj_server;

/* Result of one solve.  iterations follows the reference: Ok(k+1) when the
 * stop test passes after iteration k, Ok(0) when ||b||^2 == 0. */
typedef struct hbj_result {
  int32_t status;              /* hbj_status of this system */
  uint32_t iterations;         /* PCG iterations (IR: inner iterations summed) */
  uint32_t outer_iterations;   /* IR refinement steps (0 for plain FP32 PCG) */
  float rr;                    /* last recurrence r.r in FP32 */
  float b_norm2;               /* ||b||^2 in FP32 as the reference computes it */
  float reserved0;
  double rel_res_recurrence;   /* sqrt(rr / b_norm2) */
  double rel_res_true_fp64;    /* ||b - A x||_2 / ||b||_2 evaluated in FP64 (NaN if not evaluated) */
  double solve_ms;             /* host wall time of the call's device work; for a
                                  batch's systems: from the start of the batch's
                                  device work to the status poll (every chunk of
                                  iterations) that first showed this system done */
} hbj_result;

/* Mixed-precision (iterative refinement) parameters. */
typedef struct hbj_ir_params {
  uint32_t max_outer;          /* refinement steps, e.g. 20 */
  uint32_t max_inner;          /* FP32 PCG iterations per step, e.g. 5000 */
  float inner_rel_tol;         /* FP32 PCG relative tolerance per step, e.g. 1e-4; a step
                                  that needs less reduction to reach the target solves to
                                  0.8 times the reduction it needs instead */
  float reserved0;
  double target_rel_res;       /* stop when ||b - A x||/||b|| (FP64) <= this, e.g. 1e-10 */
} hbj_ir_params;

const char* hbj_version(void);
/* Thread-local text of the last error on this thread ("" if none). */
const char* hbj_last_error(void);

/* ---------------------------------------------------------------- context */

/* GpuContext::create (src/gpu/context.rs:60-109): select HIP device `device`. */
hbj_status hbj_ctx_create(int device, hbj_ctx** out);
void hbj_ctx_destroy(hbj_ctx* ctx);
/* GpuContext::describe (context.rs:111-120) / the `info` subcommand. */
hbj_status hbj_ctx_describe(hbj_ctx* ctx, char* buf, size_t buf_len);

/* ---------------------------------------------------- CPU setup (host) */

/* lu_factor_inplace_6 (compute.rs:37-71): no-pivot LU of the leading n x n of a
 * row-major 6x6 buffer; strict lower = L (unit diagonal implied), rest = U. */
hbj_status hbj_lu_factor_inplace_6(float* mat36, uint32_t n);

/* build_lu_blocks_from_csr_block_starts_6 (compute.rs:84-181).  lu_out holds
 * 36*(n_starts-1) floats.  Slice lengths are passed as the Rust slices carry them. */
hbj_status hbj_build_lu_blocks_6(uint32_t n, const uint32_t* row_ptr, uint32_t row_ptr_len,
                                 const uint32_t* col_idx, uint32_t col_idx_len,
                                 const float* values, uint32_t values_len,
                                 const uint32_t* block_starts, uint32_t n_starts, float* lu_out);

/* Tile-local SpMV tables of a CSR matrix, as hbj_system_create builds them:
 * per tile of 256 rows the sorted distinct columns its entries reference
 * (tile_cnt[t] of them at tcols + tile_off[t]; tile_cnt[t] = 0xFFFFFFFF and no
 * list when there are more than 2048) and per entry its index in its tile's
 * list (lidx, nnz entries; 0 in tiles without a list).  tile_off / tile_cnt
 * hold ceil(n / 256) entries; *n_tcols receives the lists' total length, and
 * a tcols_cap below it is an error (query with tcols_cap = 0).  No reference
 * counterpart: the setup of this SpMV's layout, exposed for tests. */
hbj_status hbj_spmv_tile_tables(uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                                const uint32_t* col_idx, uint32_t* tile_off, uint32_t* tile_cnt,
                                uint16_t* lidx, uint32_t* tcols, uint64_t tcols_cap,
                                uint64_t* n_tcols);

/* The sliced SpMV layout hbj_system_create would build for this matrix
 * (host only, nothing on the device): whether the system qualifies for it,
 * its 64-row waves, how many of them store only their rows' nonzero values
 * (allow_nz != 0; 0 keeps every entry, as HBJ_SLICED_NZ=0 does), and the
 * layout's value words (values, the nonzero-only waves' masks and padding)
 * and run words.  No reference counterpart: the setup of this SpMV's layout,
 * exposed for tests. */
typedef struct {
  uint32_t sliced, waves, nz_waves, reserved0;
  uint64_t value_words, run_words;
} hbj_sliced_info;
hbj_status hbj_sliced_layout_info(uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                                  const uint32_t* col_idx, const float* values, int allow_nz,
                                  hbj_sliced_info* out);

/* ------------------------------------- single-op executors (device) */
/* Each runs the same device code as the fused solver kernels on host buffers. */

/* VecOpsExecutor::encode_axpy_inplace (vec_ops_exec.rs:130-160): y = y + alpha*x */
hbj_status hbj_op_axpy(hbj_ctx* ctx, uint32_t n, float alpha, const float* x, float* y);
/* VecOpsExecutor::encode_scale_inplace_from_scalar_results (vec_ops_exec.rs:200-232): x = x*beta */
hbj_status hbj_op_scale(hbj_ctx* ctx, uint32_t n, float beta, float* x);
/* DotScalarExecutor::encode_dot_scalar_into (dot_scalar_exec.rs:151-247): *out = a.b */
hbj_status hbj_op_dot(hbj_ctx* ctx, uint32_t n, const float* a, const float* b, float* out);
/* SpmvExecutor::create + encode_copy_x_from + encode_spmv (spmv_exec.rs:38-147): y = A x */
hbj_status hbj_op_spmv(hbj_ctx* ctx, uint32_t n_rows, uint32_t nnz, const uint32_t* row_ptr,
                       const uint32_t* col_idx, const float* values, const float* x, float* y);
/* No counterpart in the reference: hbj_op_dot for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  a, b: n floats and out: one float, all
 * in device memory.  Enqueued on `stream` (a hipStream_t; NULL = the
 * context's stream) without host synchronisation and without any allocation:
 * the per-256-element partials live in a fixed device array of the library
 * (1 MiB per GPU, held until the process ends), so n is limited to
 * HBJ_DOT_NOALLOC_MAX_N (larger n: HBJ_E_INVALID).  All calls on one GPU
 * share that array: calls that overlap on one GPU race, the caller
 * serialises them.  *out receives the same bits as hbj_op_dot. */
#define HBJ_DOT_NOALLOC_MAX_N (1u << 26)
hbj_status hbj_op_dot_noalloc(hbj_ctx* ctx, uint32_t n, const float* a, const float* b,
                              float* out, void* stream);
/* BlockJacobiExecutor::create + encode_apply (block_jacobi_exec.rs:48-124): z = M^-1 r.
 * z is in/out: rows no block writes (past the 6th row of an oversize block) keep their value. */
hbj_status hbj_op_block_jacobi(hbj_ctx* ctx, uint32_t n, const float* lu,
                               const uint32_t* block_starts, uint32_t n_starts, const float* r,
                               float* z);
/* PcgUpdateScalarsExecutor::encode_update_scalars (pcg_update_scalars_exec.rs:72-118)
 * on a slot array of n_slots floats (in/out). */
hbj_status hbj_op_update_scalars(hbj_ctx* ctx, float* slots, uint32_t n_slots, uint32_t i_pap,
                                 uint32_t i_rz_new, uint32_t i_rz_old, uint32_t i_alpha,
                                 uint32_t i_minus_alpha, uint32_t i_beta);

/* ------------------------------------------- device-resident system */

/* SpmvExecutor::create + build_lu_blocks_from_csr_block_starts_6 +
 * BlockJacobiExecutor::create (spmv_exec.rs:38-118, compute.rs:84-181,
 * block_jacobi_exec.rs:48-91): validates like loaders.rs:84-105,164-184,
 * builds the LU blocks, and keeps CSR, LU and all PCG vectors in HBM. */
hbj_status hbj_system_create(hbj_ctx* ctx, uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                             const uint32_t* col_idx, const float* values,
                             const uint32_t* block_starts, uint32_t n_starts, hbj_system** out);
void hbj_system_destroy(hbj_system* sys);
/* Device bytes held by the system (CSR + LU + vectors). */
uint64_t hbj_system_device_bytes(const hbj_system* sys);

/* pcg_block_jacobi_csr_wgpu (compute.rs:189-495): the drop-in solve.
 * b: n floats; x_inout: x0 in, solution out (only on HBJ_OK). */
hbj_status hbj_solve(hbj_system* sys, const float* b, float* x_inout, uint32_t max_iter,
                     float rel_tol, float abs_tol, hbj_result* result);

/* Same solve with b and x0 already resident in HBM (uploaded once by
 * hbj_system_set_rhs); the solution stays on the device until hbj_system_get_x. */
hbj_status hbj_system_set_rhs(hbj_system* sys, const float* b, const float* x0);
hbj_status hbj_system_solve_resident(hbj_system* sys, uint32_t max_iter, float rel_tol,
                                     float abs_tol, hbj_result* result);
hbj_status hbj_system_get_x(hbj_system* sys, float* x);

/* SpmvExecutor::encode_copy_x_from + encode_spmv (spmv_exec.rs:122-147) on a
 * created system: y = A x (x, y: n floats on the host) through the SpMV of
 * the system's PCG iteration -- its resident layout and chosen variant --, so
 * y has the bits every iteration's A p has (each row summed in entry order
 * from 0.0).  Uses the system's x and q vectors: call between solves, not
 * during one. */
hbj_status hbj_system_spmv(hbj_system* sys, const float* x, float* y);

/* Launches per PCG iteration. */
#define HBJ_ITER_KERNELS 3
/* Bit-identical SpMV streaming variants a system can use: 0 = dword loads,
 * gathers from HBM; 1 = 16-byte loads, gathers from HBM; 2 = tile-local
 * column lists in LDS, 512 entries per wave pass; 3 = tile-local, 1024
 * entries per pass; 4 = sliced (default): each lane streams its own row's
 * values step by step and takes the column positions from the row's runs of
 * consecutive columns, no per-entry index (systems whose rows have at most 8
 * such runs, e.g. FEA node blocks; for other systems the same as 2), values
 * read non-temporally; 5 = 4 with the values read through the caches, so a
 * layout that fits the GPU's last-level cache stays there between SpMVs
 * (faster for such systems, slower for larger ones: the autotuner decides). */
#define HBJ_SPMV_VARIANTS 6

typedef struct {
  uint32_t n, nnz, n_blocks, n_tiles;
  uint32_t local_tiles;    /* tiles whose SpMV gathers from an LDS copy of x */
  uint32_t spmv_variant;   /* SpMV variant of the next solve */
  uint64_t local_columns;  /* total length of the tiles' column lists */
  uint64_t device_bytes;   /* device memory held by the system */
  uint64_t spmv_csr_bytes; /* one SpMV in CSR accounting: 8 nnz + 12 n + 4 */
  uint64_t spmv_layout_bytes; /* one SpMV in the tile-local layout: 6 bytes per
                                 entry of a local tile, 8 per other entry, the
                                 column lists, row_ptr, x once and q */
  uint64_t spmv_sliced_bytes; /* one SpMV in the sliced layout (variant 4): the
                                 padded values, the rows' runs (once per group of
                                 6 rows where the group shares them), the column
                                 lists (as groups of 6 columns where they are made
                                 of whole groups), the tile and slice headers, the
                                 rows' own x and q; 0 when the system has no
                                 sliced layout */
  uint32_t launches_per_iteration; /* kernel launches per PCG iteration of its
                                      solves: 2 (the previous iteration's finish
                                      folded into the SpMV; the default), 3 (SpMV,
                                      update, finish: HBJ_FUSED=0), 2 for the
                                      one-reduction PCG */
  uint32_t sliced_nz_waves; /* 64-row waves of the sliced layout that store
                               only their rows' nonzero values (rows of at most
                               64 entries with many explicit zeros, e.g. dense
                               FEA node blocks; HBJ_SLICED_NZ=0 at creation
                               keeps every entry) */
} hbj_system_info;

/* Algorithm of a system's plain solves (hbj_solve, hbj_system_solve_resident,
 * hbj_system_autotune).  HBJ_ALG_PCG is pcg_block_jacobi_csr_wgpu
 * (compute.rs:189-495) and the default.  HBJ_ALG_PCG_ONE_REDUCTION is the
 * Chronopoulos-Gear form of the same preconditioned CG: the iteration's
 * dots (r,u), (w,u) with w = A u, and (r,r) are reduced at one point, so an
 * iteration is two launches instead of three (and one all-reduce instead of
 * two when a system spans GPUs).  Same iterates in exact arithmetic, same
 * stop test, iteration count and errors; its rounding differs from the
 * reference's, so its x is not bit-identical to HBJ_ALG_PCG's.  Refinement,
 * batches and the server always use HBJ_ALG_PCG.  No reference counterpart. */
#define HBJ_ALG_PCG 0
#define HBJ_ALG_PCG_ONE_REDUCTION 1
hbj_status hbj_system_set_algorithm(hbj_system* sys, uint32_t alg);

/* No reference counterpart (introspection for measurement). */
hbj_status hbj_system_get_info(const hbj_system* sys, hbj_system_info* out);
/* Copies the system's LU blocks (36 * n_blocks floats, the layout of
 * hbj_build_lu_blocks_6; built on the device from the CSR when the system
 * was created) to lu_out.  No reference counterpart (introspection; the
 * reference's builder is build_lu_blocks_from_csr_block_starts_6,
 * compute.rs:84-181, which hbj_build_lu_blocks_6 restates on the host). */
hbj_status hbj_system_get_lu(const hbj_system* sys, float* lu_out);
/* Selects the SpMV variant (< HBJ_SPMV_VARIANTS) of later solves; results do
 * not depend on it.  No reference counterpart. */
hbj_status hbj_system_set_spmv_variant(hbj_system* sys, uint32_t variant);
/* Autotuner: runs `iters` PCG iterations from the current right-hand side
 * with each SpMV variant (one untimed run, then the best of three), keeps the
 * fastest for later solves and returns it in *best; us_per_iter
 * [HBJ_SPMV_VARIANTS] (may be null) receives each variant's time per
 * iteration.  Overwrites the
 * system's solution vector, not its right-hand side or x0.  No reference
 * counterpart. */
hbj_status hbj_system_autotune(hbj_system* sys, uint32_t iters, uint32_t* best,
                               double* us_per_iter);

/* Sample HIP-event timings of the per-iteration kernels during subsequent
 * solves of this system: every `every`-th iteration (0 = off). */
hbj_status hbj_system_set_kernel_timing(hbj_system* sys, uint32_t every);
/* avg_us[HBJ_ITER_KERNELS], counts[HBJ_ITER_KERNELS]: mean duration and
 * samples of the launches of an iteration -- SpMV + p.Ap partials; alpha +
 * x/r update + block solve + r.r/r.z partials; beta, stop test and state
 * transition + p-update.  Clears the accumulators. */
hbj_status hbj_system_kernel_timing(hbj_system* sys, double* avg_us, uint64_t* counts);

/* ------------------------------------------------ mixed precision (IR) */

/* FP64 outer iterative refinement around the FP32 PCG: r = b - A x in FP64
 * on the device, FP32 PCG on the scaled residual, x += correction in FP64.
 * No reference counterpart (extends compute.rs:189 for 1e-10 residuals). */
hbj_status hbj_solve_ir(hbj_system* sys, const double* b, double* x_inout,
                        const hbj_ir_params* params, hbj_result* result);
hbj_status hbj_system_set_rhs64(hbj_system* sys, const double* b, const double* x0);
hbj_status hbj_system_get_x64(hbj_system* sys, double* x);

/* ----------------------------------------------------------- batching */

/* Many independent systems advanced together: every PCG iteration of every
 * system in one set of kernel launches (graph-replayed), per-system
 * alpha/beta/status on the device.  A batch of several systems splits them
 * into two tables balanced by size, each with its own stream, run side by
 * side (HBJ_BATCH_TABLES=1 in the environment keeps one table).  Systems
 * keep their own resident b/x0 (hbj_system_set_rhs / _set_rhs64).  Results
 * are identical to solving each system alone.  No reference counterpart
 * (the reference solves one system). */
hbj_status hbj_batch_create(hbj_ctx* ctx, hbj_system* const* systems, uint32_t count,
                            hbj_batch** out);
void hbj_batch_destroy(hbj_batch* batch);
hbj_status hbj_batch_solve(hbj_batch* batch, uint32_t max_iter, float rel_tol, float abs_tol,
                           hbj_result* results /* count */);
hbj_status hbj_batch_solve_ir(hbj_batch* batch, const hbj_ir_params* params,
                              hbj_result* results /* count */);
/* Times each of the HBJ_SPMV_VARIANTS bit-identical SpMV variants on the
 * batch's own tile table (one untimed run, then the best of three runs of
 * `iters` iterations of every system; ir != 0: inside the refinement loop on
 * the systems' FP64 right-hand sides, else the plain PCG on their FP32 ones)
 * and keeps the fastest for later solves; us (HBJ_SPMV_VARIANTS doubles, may
 * be NULL) receives microseconds per iteration.  Overwrites the systems'
 * solution vectors; results of later solves do not change.  No reference
 * counterpart. */
hbj_status hbj_batch_autotune(hbj_batch* batch, uint32_t iters, int ir, uint32_t* best,
                              double* us);
/* Sample HIP-event timings of the per-iteration kernels during subsequent
 * solves of this batch: every `every`-th launch set (0 = off); the samples
 * then launch directly instead of replaying the captured graphs.  No
 * reference counterpart. */
hbj_status hbj_batch_set_kernel_timing(hbj_batch* batch, uint32_t every);
/* (While timing is on, the batch runs as one table, so that each sampled
 * launch has the GPU to itself.) */
/* The last timed solve's samples: record i is rec[i * (1 + HBJ_ITER_KERNELS)
 * ...]: the launch set's index within the solve, then each kernel's
 * duration in microseconds (-1: the iteration form has no such launch).
 * Only sets in which some system still ran are kept.  *n_out receives the
 * number of records (at most `cap` are written); sets (count entries, may be
 * null) receives, per system, the launch sets that found it running: system
 * i was part of sets [0, sets[i]) of the solve, so the work of a sampled set
 * follows from the systems with sets[i] above its index.  No reference
 * counterpart. */
hbj_status hbj_batch_kernel_samples(hbj_batch* batch, double* rec, uint64_t cap, uint64_t* n_out,
                                    uint32_t* sets);

/* ------------------------------------------------------------ serving */

typedef struct hbj_server_config {
  uint64_t hbm_budget_bytes;   /* admission limit on device bytes of resident requests;
                                  the server allocates it at creation and places
                                  each resident system in a block of it */
  uint32_t max_active;         /* most systems advanced together */
  uint32_t chunk_iters;        /* iterations between host status polls (0: 64) */
} hbj_server_config;

typedef struct hbj_server_stats {
  uint64_t submitted, completed, rejected, active, queued;
  uint64_t device_bytes_in_use;
  uint64_t peak_device_bytes;
  uint64_t peak_active;
  uint64_t chunks;         /* chunks of iterations enqueued */
  uint64_t table_changes;  /* slot-table rebuilds (arrivals / retirements folded in) */
  uint64_t arena_waits;    /* admissions that waited for a block to come back
                              (the budget allowed them, no free range was long enough) */
  double admit_ms;         /* host time spent making requests resident (summed) */
  double table_ms;         /* worker time spent on table rebuilds (summed) */
  uint64_t upload_bytes;   /* bytes copied host -> device for admitted requests */
  uint64_t chunks_cached_spmv; /* chunks whose SpMV read the values through the caches
                                  (variant 5: the resident layouts fit the last-level
                                  cache); the others read them non-temporally (4) */
} hbj_server_stats;

/* Continuous-batching request loop: a host thread admits queued requests as
 * HBM budget and slots free up, uploads them while others iterate, advances
 * all resident requests together, and retires each as it converges; the
 * slot table changes between chunks of iterations without waiting for the
 * device.  A request's system is built on the host (tables, SpMV layout, LU
 * blocks) and copied in one asynchronous upload from pinned staging; a
 * system with the sliced SpMV layout keeps only that layout on the device
 * (no CSR arrays).  The resident requests run in two lanes (each its own
 * table, stream and worker thread, side by side; HBJ_SERVER_LANES=1 in the
 * environment keeps one); a request joins the lane with fewer resident
 * tiles. */
hbj_status hbj_server_create(hbj_ctx* ctx, const hbj_server_config* cfg, hbj_server** out);
void hbj_server_destroy(hbj_server* srv);
/* Copies the request; returns a ticket.  b64 in FP64; solved with IR (params).
 * HBJ_E_REJECTED at once when no block for these sizes could fit the budget;
 * a request whose exact block (hbj_server_request_bytes) exceeds the budget
 * completes with HBJ_E_REJECTED. */
hbj_status hbj_server_submit(hbj_server* srv, uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                             const uint32_t* col_idx, const float* values,
                             const uint32_t* block_starts, uint32_t n_starts, const double* b64,
                             const hbj_ir_params* params, uint64_t* ticket);
/* Blocks until the request is done; x (n doubles, may be NULL) receives the
 * solution; latency_ms = completion - submission. */
hbj_status hbj_server_wait(hbj_server* srv, uint64_t ticket, double* x, hbj_result* result,
                           double* latency_ms);
hbj_status hbj_server_stats_get(hbj_server* srv, hbj_server_stats* stats);
/* Device bytes a request reserves from the budget: the exact block its
 * system takes in the server's arena (the host tables are built to find it;
 * 0 and an error text for an invalid matrix or a zero pivot).  No reference
 * counterpart: sizing of budgets. */
uint64_t hbj_server_request_bytes(uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                                  const uint32_t* col_idx, const float* values,
                                  const uint32_t* block_starts, uint32_t n_starts);

/* Calibration: `reps` launches of a plain streaming copy of `bytes` (read
 * bytes + write bytes per launch); *avg_us = mean launch duration (HIP events).
 * No reference counterpart: the known-size kernel against which the HBM
 * traffic counters are scaled, and a measured bandwidth ceiling. */
hbj_status hbj_stream_copy(hbj_ctx* ctx, uint64_t bytes, uint32_t reps, double* avg_us);

/* -------------------------------------------- synthetic workload (host) */

/* 3D space-frame lattice of N^3 nodes, 6 DOF/node, Euler-Bernoulli beams
 * between axis neighbours (spacing 1 m, E=2.1e11, G=8.1e10, A~U[1e-3,1e-2],
 * Iy,Iz,J~U[1e-6,1e-5]); bottom layer clamped as identity rows; dense 6x6
 * node blocks kept; RHS ~U[-1e3,1e3] on free DOFs; block_starts = 0,6,..,n.
 * e_log10_half_range > 0 scales E and G per element by 10^U[-h,h].
 * Query sizes with hbj_lattice_size, then fill caller-allocated arrays.
 * No reference counterpart: the benchmark's input generator. */
hbj_status hbj_lattice_size(uint32_t N, uint32_t* n, uint64_t* nnz);
hbj_status hbj_lattice_generate(uint32_t N, uint64_t seed, double e_log10_half_range,
                                uint32_t* row_ptr, uint32_t* col_idx, float* values, float* rhs,
                                uint32_t* block_starts);
/* 12x12 global stiffness of one beam element along axis dir (0=x,1=y,2=z). */
hbj_status hbj_beam_element_stiffness(double E, double G, double A, double Iy, double Iz,
                                      double J, double L, int dir, double* k144);

#ifdef __cplusplus
}
#endif
#endif /* HBJ_H */
