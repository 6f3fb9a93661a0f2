/*
 * oracle.h -- CPU restatement of the reference block-Jacobi PCG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: it may be loaded by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, and by nothing else.  The product library
 * (wgpu_solver_backend_amd/libhbj.so) never links or calls it.
 *
 * Every function restates one piece of RomanShushakov/wgpu_solver_backend
 * (paths relative to the reference checkout) in plain C, following the
 * reference's operation order so that FP32 results can be compared bit for
 * bit.  Arithmetic rules shared by all functions:
 *   - built with -ffp-contract=off: a*b+c is two roundings, as written;
 *   - IEEE single precision, round-to-nearest, correctly rounded division;
 *   - dot products use the reference's 256-wide stride-halving tree per
 *     workgroup, then 256-way tree passes over the partials.
 *
 * Pinning (what checks each function independently of its own formula) is
 * listed in the header comment of oracle.c.
 */
#ifndef HBJ_ORACLE_H
#define HBJ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: same values as hbj_status in include/hbj.h */
enum {
  OR_OK = 0,
  OR_E_INVALID = 1,
  OR_E_ZERO_PIVOT = 2,
  OR_E_BREAKDOWN_PAP = 3,
  OR_E_BREAKDOWN_RZ = 4,
  OR_E_NOT_CONVERGED = 5
};

/* Number of OpenMP threads the oracle uses (1 = scalar).  Results do not
 * depend on it: every reduction keeps the reference's fixed order. */
void or_set_threads(int n);
int or_get_threads(void);

/* Last error text (reference wording), thread-local. */
const char* or_last_error(void);

/* compute.rs:37-71 */
int or_lu_factor_inplace_6(float* mat, uint32_t n);

/* compute.rs:84-181; out has 36*(n_starts-1) floats */
int or_build_lu_blocks_6(uint32_t n, const uint32_t* row_ptr, uint32_t row_ptr_len,
                         const uint32_t* col_idx, const float* values, uint32_t nnz,
                         const uint32_t* block_starts, uint32_t n_starts, float* out);

/* spmv.wgsl:43-64 */
void or_spmv(uint32_t n_rows, const uint32_t* row_ptr, const uint32_t* col_idx,
             const float* values, const float* x, float* y);

/* dot_partials.wgsl:37-75 + dot_reduce.wgsl:36-74 + dot_scalar_exec.rs:151-247 */
float or_dot(uint32_t n, const float* a, const float* b);

/* axpy.wgsl:41 and axpy_from_scalar_results.wgsl:41: y = y + alpha*x */
void or_axpy(uint32_t n, float alpha, const float* x, float* y);

/* scale_from_scalar_results.wgsl:37: x = x*beta */
void or_scale(uint32_t n, float beta, float* x);

/* block_jacobi.wgsl:63-127 (z rows not covered by a block are left as they are) */
void or_block_jacobi(uint32_t n, const float* lu, const uint32_t* block_starts,
                     uint32_t n_starts, const float* r, float* z);

/* pcg_update_scalars.wgsl:47-73 on a slot array */
void or_update_scalars(float* slots, uint32_t i_pap, uint32_t i_rz_new, uint32_t i_rz_old,
                       uint32_t i_alpha, uint32_t i_minus_alpha, uint32_t i_beta);

/* compute.rs:189-495.  x holds x0 on entry and the solution on OR_OK only.
 * *iterations receives the reference's Ok(k+1) / Ok(0) value (or the number
 * of iterations run on error).  rr_hist (optional, max_iter floats) receives
 * r.r after every iteration. */
int or_pcg(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const float* values,
           const float* lu, const uint32_t* block_starts, uint32_t n_starts,
           const float* b, float* x, uint32_t max_iter, float rel_tol, float abs_tol,
           uint32_t* iterations, float* rr_hist);

/* One-reduction (Chronopoulos-Gear) PCG of the HIP library's optional
 * algorithm; x is written only on OR_OK. */
int or_ccg(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const float* values,
           const float* lu, const uint32_t* block_starts, uint32_t n_starts,
           const float* b, float* x, uint32_t max_iter, float rel_tol, float abs_tol,
           uint32_t* iterations);

/* ---- FP64 oracle (error reporting; not a restatement) ---- */

/* y = A x in double, row order. */
void or_spmv64(uint32_t n_rows, const uint32_t* row_ptr, const uint32_t* col_idx,
               const float* values, const double* x, double* y);

/* ||b - A x||_2 / ||b||_2 in double for FP64 b and x. */
double or_true_rel_residual64(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
                              const float* values, const double* b, const double* x);

/* Block-Jacobi PCG entirely in double (LU of the same blocks in double),
 * run until ||r||/||b|| <= rel_tol (recomputing r = b - A x every 50
 * iterations).  Returns OR_OK or OR_E_NOT_CONVERGED; x in/out (when not
 * converged: the iterate with the lowest true residual among those checked).
 * FP64 dots use a fixed summation order, so results do not depend on the
 * thread count. */
int or_pcg64(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const float* values,
             const uint32_t* block_starts, uint32_t n_starts, const double* b, double* x,
             uint32_t max_iter, double rel_tol, uint32_t* iterations);

#ifdef __cplusplus
}
#endif
#endif
