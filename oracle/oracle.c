/*
 * oracle.c -- CPU restatement of the reference block-Jacobi PCG path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP product
 * path, never part of it.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off -fPIC -shared).
 *
 * How each function is pinned (tests/test_oracle.py), i.e. checked against
 * something other than its own formula:
 *   or_lu_factor_inplace_6   L*U reproduces the block (numpy, FP64 bound);
 *                            zero-pivot text and index; unit-L/upper-U layout
 *                            recovered from numpy's own unpivoted elimination
 *                            on diagonally dominant blocks (exact on small
 *                            integer blocks).
 *   or_build_lu_blocks_6     identity fill for m<6, oversize blocks (leading
 *                            6x6 only), duplicate entries (last wins), the
 *                            reference's validation texts.
 *   or_spmv                  reference KAT spmv-test (cli main.rs:250-315);
 *                            exact equality with scipy on integer data;
 *                            FP64 error bound on random data.
 *   or_dot                   reference KAT dot-test (cli main.rs:213-248);
 *                            hand-derived cases whose result differs between
 *                            the stride-halving tree, an adjacent-pair tree and
 *                            a sequential sum (first level and partial level).
 *   or_axpy / or_scale       reference KAT vec-test (cli main.rs:171-211);
 *                            elementwise numpy float32 (same single rounding).
 *   or_block_jacobi          reference KAT block-jacobi-test (cli main.rs:
 *                            317-368); numpy FP64 solve of the unfactored
 *                            block within a conditioning bound; stale rows.
 *   or_update_scalars        reference KAT pcg-update-scalars-test (cli
 *                            main.rs:377-443); zero-denominator guards.
 *   or_pcg                   x_ref of datasets 1-3 with the commands.txt
 *                            tolerances; scipy FP64 direct solve; iteration
 *                            counts of an independent numpy FP64 PCG on a
 *                            well-conditioned system; breakdown / b=0 /
 *                            max_iter error paths.
 *   or_ccg                   not a reference function (the library's optional
 *                            one-reduction PCG): an independent numpy FP64
 *                            restatement of the same recurrences takes the
 *                            FP64 PCG's iteration count; FP32 runs converge to
 *                            x_ref / the scipy solution like or_pcg; error
 *                            paths.
 *   or_pcg64, or_spmv64,
 *   or_true_rel_residual64   scipy spsolve / numpy residuals.
 */
#include "oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;
static __thread char g_err[512];

static void set_err(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

void or_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int or_get_threads(void) { return g_threads; }
const char* or_last_error(void) { return g_err; }

/* compute.rs:37-71 -- in-place no-pivot LU of the leading n x n of a 6x6
 * row-major buffer; strict lower = L (unit diagonal implied), rest = U. */
int or_lu_factor_inplace_6(float* mat, uint32_t n) {
  const uint32_t s = 6;
  if (n > 6) {
    set_err("lu_factor_inplace_6: n must be <= 6, got %u", n);
    return OR_E_INVALID;
  }
  for (uint32_t k = 0; k < n; ++k) {
    float akk = mat[k * s + k];
    if (akk == 0.0f) {
      set_err("LU zero pivot at k = %u", k);
      return OR_E_ZERO_PIVOT;
    }
    for (uint32_t i = k + 1; i < n; ++i) mat[i * s + k] /= akk;
    for (uint32_t i = k + 1; i < n; ++i) {
      float lik = mat[i * s + k];
      if (lik != 0.0f) {
        for (uint32_t j = k + 1; j < n; ++j) mat[i * s + j] -= lik * mat[k * s + j];
      }
    }
  }
  return OR_OK;
}

/* compute.rs:84-181 */
int or_build_lu_blocks_6(uint32_t n, const uint32_t* row_ptr, uint32_t row_ptr_len,
                         const uint32_t* col_idx, const float* values, uint32_t nnz,
                         const uint32_t* bs, uint32_t n_starts, float* out) {
  if (row_ptr_len != n + 1) {
    set_err("build_lu_blocks_from_csr_block_starts_6: row_ptr len must be n+1 (%u), got %u", n + 1,
            row_ptr_len);
    return OR_E_INVALID;
  }
  uint32_t rp_nnz = row_ptr[n];
  if (rp_nnz != nnz) {
    set_err("build_lu_blocks_from_csr_block_starts_6: nnz mismatch: row_ptr says %u, col_idx %u, values %u",
            rp_nnz, nnz, nnz);
    return OR_E_INVALID;
  }
  if (n_starts < 2) return OR_OK;
  if (bs[0] != 0) {
    set_err("build_lu_blocks_from_csr_block_starts_6: block_starts must start at 0");
    return OR_E_INVALID;
  }
  if (bs[n_starts - 1] != n) {
    set_err("build_lu_blocks_from_csr_block_starts_6: block_starts last must equal n (%u), got %u", n,
            bs[n_starts - 1]);
    return OR_E_INVALID;
  }
  for (uint32_t w = 0; w + 1 < n_starts; ++w) {
    if (bs[w + 1] <= bs[w]) {
      set_err("build_lu_blocks_from_csr_block_starts_6: block_starts must be strictly increasing");
      return OR_E_INVALID;
    }
  }
  uint32_t nb = n_starts - 1;
  memset(out, 0, sizeof(float) * 36u * nb);
  for (uint32_t blk = 0; blk < nb; ++blk) {
    uint32_t off = bs[blk], end = bs[blk + 1];
    if (end > n || off >= end) continue;
    uint32_t m = end - off < 6 ? end - off : 6;
    float mat[36];
    memset(mat, 0, sizeof mat);
    for (uint32_t i = 0; i < 6; ++i) mat[i * 6 + i] = 1.0f;
    for (uint32_t il = 0; il < m; ++il) {
      uint32_t i = off + il;
      for (uint32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
        uint32_t j = col_idx[k];
        if (j >= off && j < off + m) mat[il * 6 + (j - off)] = values[k];
      }
    }
    int st = or_lu_factor_inplace_6(mat, m);
    if (st != OR_OK) return st;
    memcpy(out + 36u * blk, mat, sizeof mat);
  }
  return OR_OK;
}

/* spmv.wgsl:43-64: one sequential FP32 sum per row, starting at 0.0 */
void or_spmv(uint32_t n_rows, const uint32_t* row_ptr, const uint32_t* col_idx,
             const float* values, const float* x, float* y) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
  for (int64_t i = 0; i < (int64_t)n_rows; ++i) {
    float sum = 0.0f;
    for (uint32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) sum = sum + values[k] * x[col_idx[k]];
    y[i] = sum;
  }
}

/* dot_partials.wgsl:57-69 / dot_reduce.wgsl:56-68: stride-halving tree over
 * 256 slots, s[t] = s[t] + s[t+offset] for offset = 128, 64, ..., 1. */
static float tree256(float* s) {
  for (uint32_t off = 128; off >= 1; off >>= 1)
    for (uint32_t t = 0; t < off; ++t) s[t] = s[t] + s[t + off];
  return s[0];
}

/* dot_scalar_exec.rs:151-247: partials over 256-chunks (zero padded), then
 * repeated 256-way tree passes until one value is left; n == 0 gives 0. */
float or_dot(uint32_t n, const float* a, const float* b) {
  if (n == 0) return 0.0f;
  uint32_t len = (n + 255u) / 256u;
  float* part = (float*)malloc(sizeof(float) * len);
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1 && len > 64)
  for (int64_t g = 0; g < (int64_t)len; ++g) {
    float s[256];
    for (uint32_t t = 0; t < 256; ++t) {
      uint64_t i = (uint64_t)g * 256u + t;
      s[t] = i < n ? a[i] * b[i] : 0.0f;
    }
    part[g] = tree256(s);
  }
  while (len > 1) {
    uint32_t out_len = (len + 255u) / 256u;
    for (uint32_t g = 0; g < out_len; ++g) {
      float s[256];
      for (uint32_t t = 0; t < 256; ++t) {
        uint64_t i = (uint64_t)g * 256u + t;
        s[t] = i < len ? part[i] : 0.0f;
      }
      part[g] = tree256(s); /* g <= g*256: only consumed slots are overwritten */
    }
    len = out_len;
  }
  float r = part[0];
  free(part);
  return r;
}

void or_axpy(uint32_t n, float alpha, const float* x, float* y) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
  for (int64_t i = 0; i < (int64_t)n; ++i) y[i] = y[i] + alpha * x[i];
}

void or_scale(uint32_t n, float beta, float* x) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
  for (int64_t i = 0; i < (int64_t)n; ++i) x[i] = x[i] * beta;
}

/* block_jacobi.wgsl:63-127 */
void or_block_jacobi(uint32_t n, const float* lu, const uint32_t* bs, uint32_t n_starts,
                     const float* r, float* z) {
  if (n_starts < 2) return;
  uint32_t nb = n_starts - 1;
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
  for (int64_t blk = 0; blk < (int64_t)nb; ++blk) {
    uint32_t off = bs[blk], next = bs[blk + 1];
    if (off >= n || next <= off) continue;
    uint32_t m = next - off < 6 ? next - off : 6;
    const float* L = lu + 36u * (uint64_t)blk;
    float y[6] = {0}, xx[6] = {0};
    for (uint32_t i = 0; i < m; ++i) {
      float sum = r[off + i];
      for (uint32_t j = 0; j < i; ++j) sum = sum - L[i * 6 + j] * y[j];
      y[i] = sum;
    }
    for (int ii = (int)m - 1; ii >= 0; --ii) {
      uint32_t i = (uint32_t)ii;
      float sum = y[i];
      for (uint32_t j = i + 1; j < m; ++j) sum = sum - L[i * 6 + j] * xx[j];
      xx[i] = sum / L[i * 6 + i];
    }
    for (uint32_t i = 0; i < m; ++i) z[off + i] = xx[i];
  }
}

/* pcg_update_scalars.wgsl:47-73 */
void or_update_scalars(float* s, uint32_t i_pap, uint32_t i_rz_new, uint32_t i_rz_old,
                       uint32_t i_alpha, uint32_t i_minus_alpha, uint32_t i_beta) {
  float p_ap = s[i_pap], rz_new = s[i_rz_new], rz_old = s[i_rz_old];
  float alpha = 0.0f, beta = 0.0f;
  if (p_ap != 0.0f) alpha = rz_old / p_ap;
  if (rz_old != 0.0f) beta = rz_new / rz_old;
  s[i_alpha] = alpha;
  s[i_minus_alpha] = -alpha;
  s[i_beta] = beta;
}

/* compute.rs:189-495: init sequence (:283-331), loop steps A-K (:336-460),
 * host checks in the reference's order (:468-488). */
int or_pcg(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const float* values,
           const float* lu, const uint32_t* bs, uint32_t n_starts, const float* b, float* x_io,
           uint32_t max_iter, float rel_tol, float abs_tol, uint32_t* iterations, float* rr_hist) {
  enum { S_PAP = 0, S_RR = 1, S_RZ_NEW = 2, S_RZ_OLD = 3, S_ALPHA = 4, S_MALPHA = 5, S_BETA = 6 };
  float slots[7] = {0};
  *iterations = 0;
  float b_norm2 = or_dot(n, b, b); /* :228-250 */
  if (b_norm2 == 0.0f) return OR_OK; /* :252-254 */
  float rel_tol2 = rel_tol * rel_tol, abs_tol2 = abs_tol * abs_tol; /* :256-257 */

  size_t nb = sizeof(float) * (n ? n : 1);
  float* x = (float*)malloc(nb);
  float* r = (float*)calloc(n ? n : 1, sizeof(float));
  float* p = (float*)calloc(n ? n : 1, sizeof(float));
  float* z = (float*)calloc(n ? n : 1, sizeof(float)); /* wgpu buffers start zeroed */
  float* y = (float*)calloc(n ? n : 1, sizeof(float));
  memcpy(x, x_io, nb);

  /* init :293-330 */
  memcpy(r, b, nb);
  or_spmv(n, row_ptr, col_idx, values, x, y);
  or_axpy(n, -1.0f, y, r);
  or_block_jacobi(n, lu, bs, n_starts, r, z);
  memcpy(p, z, nb);
  float rz_old = or_dot(n, r, z);

  int status = OR_E_NOT_CONVERGED;
  uint32_t k;
  for (k = 0; k < max_iter; ++k) {
    or_spmv(n, row_ptr, col_idx, values, p, y);                                    /* A */
    slots[S_PAP] = or_dot(n, p, y);                                                /* B */
    slots[S_RZ_OLD] = rz_old;                                                      /* C */
    or_update_scalars(slots, S_PAP, S_RZ_NEW, S_RZ_OLD, S_ALPHA, S_MALPHA, S_BETA); /* D */
    or_axpy(n, slots[S_ALPHA], p, x);                                              /* E */
    or_axpy(n, slots[S_MALPHA], y, r);
    slots[S_RR] = or_dot(n, r, r);                                                 /* F */
    or_block_jacobi(n, lu, bs, n_starts, r, z);                                    /* G */
    slots[S_RZ_NEW] = or_dot(n, r, z);                                             /* H */
    or_update_scalars(slots, S_PAP, S_RZ_NEW, S_RZ_OLD, S_ALPHA, S_MALPHA, S_BETA); /* I */
    or_scale(n, slots[S_BETA], p);                                                 /* J */
    or_axpy(n, 1.0f, z, p);
    float p_ap = slots[S_PAP], r_norm2 = slots[S_RR], rz_new = slots[S_RZ_NEW];
    if (rr_hist) rr_hist[k] = r_norm2;
    *iterations = k + 1;
    if (p_ap == 0.0f) {
      set_err("PCG(BlockJacobiGpu): dot(p,Ap) is zero (breakdown)");
      status = OR_E_BREAKDOWN_PAP;
      break;
    }
    if (rz_old == 0.0f) {
      set_err("PCG(BlockJacobiGpu): rz_old is zero (breakdown)");
      status = OR_E_BREAKDOWN_RZ;
      break;
    }
    if (r_norm2 <= abs_tol2 || r_norm2 <= rel_tol2 * b_norm2) {
      memcpy(x_io, x, nb);
      status = OR_OK;
      break;
    }
    rz_old = rz_new;
  }
  if (status == OR_E_NOT_CONVERGED)
    set_err("PCG(BlockJacobiGpu): did not converge in %u iterations", max_iter);
  free(x);
  free(r);
  free(p);
  free(z);
  free(y);
  return status;
}

/* One-reduction PCG (Chronopoulos-Gear), the HIP library's optional
 * algorithm (hbj_system_set_algorithm), not a reference function: the same
 * FP32 building blocks as or_pcg (or_spmv, or_dot, or_block_jacobi), in the
 * library kernels' order.  u = M^-1 r, w = A u, gamma = (r,u), delta =
 * (w,u); beta = gamma / gamma_prev, alpha = gamma / (delta - beta * (gamma /
 * alpha_prev)) (first iteration: beta = 0, alpha = gamma / delta); p = u +
 * beta p, s = w + beta s, x = x + alpha p, r = r + (-alpha) s.  Stop test on
 * (r,r) after each update (never on r0) with or_pcg's thresholds; the same
 * error statuses and texts. */
int or_ccg(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const float* values,
           const float* lu, const uint32_t* bs, uint32_t n_starts, const float* b, float* x_io,
           uint32_t max_iter, float rel_tol, float abs_tol, uint32_t* iterations) {
  *iterations = 0;
  float b_norm2 = or_dot(n, b, b);
  if (b_norm2 == 0.0f) return OR_OK;
  float rel_tol2 = rel_tol * rel_tol, abs_tol2 = abs_tol * abs_tol;
  size_t nn = n ? n : 1, nb = sizeof(float) * nn;
  float* x = (float*)malloc(nb);
  float* r = (float*)calloc(nn, sizeof(float));
  float* u = (float*)calloc(nn, sizeof(float));
  float* w = (float*)calloc(nn, sizeof(float));
  float* p = (float*)calloc(nn, sizeof(float));
  float* s = (float*)calloc(nn, sizeof(float));
  memcpy(x, x_io, nb);
  or_spmv(n, row_ptr, col_idx, values, x, w);
  for (uint32_t i = 0; i < n; ++i) r[i] = b[i] + (-1.0f) * w[i];
  or_block_jacobi(n, lu, bs, n_starts, r, u);
  float g_prev = 0.0f, a_prev = 0.0f;
  int status = OR_E_NOT_CONVERGED;
  for (uint32_t it = 0;; ++it) {
    or_spmv(n, row_ptr, col_idx, values, u, w);
    float g = or_dot(n, r, u), dl = or_dot(n, w, u), rr = or_dot(n, r, r);
    *iterations = it;
    if (it > 0 && (rr <= abs_tol2 || rr <= rel_tol2 * b_norm2)) {
      memcpy(x_io, x, nb);
      status = OR_OK;
      break;
    }
    if (it >= max_iter) break;
    float beta = 0.0f, den = dl;
    if (it > 0) {
      if (g_prev == 0.0f) {
        set_err("PCG(BlockJacobiGpu): rz_old is zero (breakdown)");
        status = OR_E_BREAKDOWN_RZ;
        break;
      }
      beta = g / g_prev;
      float ga = g / a_prev;
      den = dl - beta * ga;
    }
    if (den == 0.0f) {
      set_err("PCG(BlockJacobiGpu): dot(p,Ap) is zero (breakdown)");
      status = OR_E_BREAKDOWN_PAP;
      break;
    }
    float alpha = g / den, na = -alpha;
    for (uint32_t i = 0; i < n; ++i) {
      p[i] = u[i] + beta * p[i];
      s[i] = w[i] + beta * s[i];
      x[i] = x[i] + alpha * p[i];
      r[i] = r[i] + na * s[i];
    }
    or_block_jacobi(n, lu, bs, n_starts, r, u);
    g_prev = g;
    a_prev = alpha;
  }
  if (status == OR_E_NOT_CONVERGED)
    set_err("PCG(BlockJacobiGpu): did not converge in %u iterations", max_iter);
  free(x);
  free(r);
  free(u);
  free(w);
  free(p);
  free(s);
  return status;
}

/* ------------------------------------------------------------------ */
/* FP64 oracle                                                         */
/* ------------------------------------------------------------------ */

void or_spmv64(uint32_t n_rows, const uint32_t* row_ptr, const uint32_t* col_idx,
               const float* values, const double* x, double* y) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
  for (int64_t i = 0; i < (int64_t)n_rows; ++i) {
    double sum = 0.0;
    for (uint32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) sum += (double)values[k] * x[col_idx[k]];
    y[i] = sum;
  }
}

/* FP64 dot with a fixed summation order whatever the thread count: blocks
 * of 4096 summed sequentially, then the block sums in order. */
static double dot64(uint32_t n, const double* a, const double* b) {
  enum { kBlk = 4096 };
  const int64_t nblk = ((int64_t)n + kBlk - 1) / kBlk;
  double part_stack[64];
  double* part = nblk <= 64 ? part_stack : (double*)malloc(sizeof(double) * (size_t)nblk);
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1 && nblk > 1)
  for (int64_t k = 0; k < nblk; ++k) {
    const int64_t i1 = (k + 1) * kBlk < (int64_t)n ? (k + 1) * kBlk : (int64_t)n;
    double s = 0.0;
    for (int64_t i = k * kBlk; i < i1; ++i) s += a[i] * b[i];
    part[k] = s;
  }
  double s = 0.0;
  for (int64_t k = 0; k < nblk; ++k) s += part[k];
  if (part != part_stack) free(part);
  return s;
}

double or_true_rel_residual64(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
                              const float* values, const double* b, const double* x) {
  double* ax = (double*)malloc(sizeof(double) * (n ? n : 1));
  or_spmv64(n, row_ptr, col_idx, values, x, ax);
  double rr = 0.0, bb = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    double d = b[i] - ax[i];
    rr += d * d;
    bb += b[i] * b[i];
  }
  free(ax);
  return bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
}

/* block LU in double: leading m x m of each block, identity elsewhere */
static int lu64_blocks(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
                       const float* values, const uint32_t* bs, uint32_t nb, double* lu) {
  for (uint32_t blk = 0; blk < nb; ++blk) {
    uint32_t off = bs[blk], end = bs[blk + 1];
    uint32_t m = end - off < 6 ? end - off : 6;
    double* a = lu + 36u * (uint64_t)blk;
    memset(a, 0, sizeof(double) * 36);
    for (uint32_t i = 0; i < 6; ++i) a[i * 6 + i] = 1.0;
    for (uint32_t il = 0; il < m; ++il) {
      uint32_t i = off + il;
      for (uint32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
        uint32_t j = col_idx[k];
        if (j >= off && j < off + m) a[il * 6 + (j - off)] = values[k];
      }
    }
    for (uint32_t k = 0; k < m; ++k) {
      if (a[k * 6 + k] == 0.0) return OR_E_ZERO_PIVOT;
      for (uint32_t i = k + 1; i < m; ++i) {
        double l = a[i * 6 + k] / a[k * 6 + k];
        a[i * 6 + k] = l;
        for (uint32_t j = k + 1; j < m; ++j) a[i * 6 + j] -= l * a[k * 6 + j];
      }
    }
  }
  (void)n;
  return OR_OK;
}

static void bj64(uint32_t n, const double* lu, const uint32_t* bs, uint32_t nb, const double* r,
                 double* z) {
  memcpy(z, r, sizeof(double) * n); /* rows past the leading 6 of a block: identity */
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
  for (int64_t blk = 0; blk < (int64_t)nb; ++blk) {
    uint32_t off = bs[blk], end = bs[blk + 1];
    uint32_t m = end - off < 6 ? end - off : 6;
    const double* L = lu + 36u * (uint64_t)blk;
    double y[6], xx[6];
    for (uint32_t i = 0; i < m; ++i) {
      double s = r[off + i];
      for (uint32_t j = 0; j < i; ++j) s -= L[i * 6 + j] * y[j];
      y[i] = s;
    }
    for (int ii = (int)m - 1; ii >= 0; --ii) {
      double s = y[ii];
      for (uint32_t j = (uint32_t)ii + 1; j < m; ++j) s -= L[ii * 6 + j] * xx[j];
      xx[ii] = s / L[ii * 6 + ii];
    }
    for (uint32_t i = 0; i < m; ++i) z[off + i] = xx[i];
  }
}

int or_pcg64(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const float* values,
             const uint32_t* bs, uint32_t n_starts, const double* b, double* x, uint32_t max_iter,
             double rel_tol, uint32_t* iterations) {
  uint32_t nb = n_starts - 1;
  *iterations = 0;
  double* lu = (double*)malloc(sizeof(double) * 36u * (nb ? nb : 1));
  if (lu64_blocks(n, row_ptr, col_idx, values, bs, nb, lu) != OR_OK) {
    free(lu);
    set_err("LU zero pivot (fp64 oracle)");
    return OR_E_ZERO_PIVOT;
  }
  size_t sz = sizeof(double) * (n ? n : 1);
  double *r = (double*)malloc(sz), *z = (double*)malloc(sz), *p = (double*)malloc(sz),
         *q = (double*)malloc(sz), *xbest = (double*)malloc(sz);
  double best = INFINITY; /* lowest true relative residual seen (at replacements) */
  double bb = dot64(n, b, b);
  int status = OR_E_NOT_CONVERGED;
  if (bb == 0.0) {
    memset(x, 0, sz);
    status = OR_OK;
    goto done;
  }
  or_spmv64(n, row_ptr, col_idx, values, x, q);
  for (uint32_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
  bj64(n, lu, bs, nb, r, z);
  memcpy(p, z, sz);
  double rz = dot64(n, r, z);
  for (uint32_t k = 0; k < max_iter; ++k) {
    or_spmv64(n, row_ptr, col_idx, values, p, q);
    double pq = dot64(n, p, q);
    if (pq == 0.0) break;
    double alpha = rz / pq;
    for (uint32_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    if ((k + 1) % 50 == 0) { /* residual replacement keeps r honest */
      or_spmv64(n, row_ptr, col_idx, values, x, q);
      for (uint32_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
      const double t = sqrt(dot64(n, r, r) / bb);
      if (t < best) { /* past the attainable accuracy the iteration only drifts */
        best = t;
        memcpy(xbest, x, sz);
      }
      if (!isfinite(t)) break;
    }
    double rr = dot64(n, r, r);
    *iterations = k + 1;
    if (sqrt(rr / bb) <= rel_tol) {
      or_spmv64(n, row_ptr, col_idx, values, x, q);
      double tr = 0.0;
      for (uint32_t i = 0; i < n; ++i) tr += (b[i] - q[i]) * (b[i] - q[i]);
      if (sqrt(tr / bb) <= rel_tol) {
        status = OR_OK;
        break;
      }
      for (uint32_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
    }
    bj64(n, lu, bs, nb, r, z);
    double rz_new = dot64(n, r, z);
    double beta = rz_new / rz;
    rz = rz_new;
    for (uint32_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  if (status != OR_OK && best < INFINITY) {
    /* not converged: return the best iterate if the last one is worse */
    or_spmv64(n, row_ptr, col_idx, values, x, q);
    double tr = 0.0;
    for (uint32_t i = 0; i < n; ++i) tr += (b[i] - q[i]) * (b[i] - q[i]);
    if (!(sqrt(tr / bb) <= best)) memcpy(x, xbest, sz);
  }
done:
  free(lu);
  free(r);
  free(z);
  free(p);
  free(q);
  free(xbest);
  return status;
}
