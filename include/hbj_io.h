/*
 * hbj_io.h -- case-directory IO and the compare-x metric (host only).
 *
 * Same file formats and validation messages as the reference, so the
 * reference's commands.txt recipe runs unchanged against this backend
 * (text-to-bin -> run-pcg-case -> compare-x, see the hbj_cli tool):
 *   matrix.csr.bin   u32 magic 0x43535231 "CSR1", u32 version 1, u32 n_rows,
 *                    u32 n_cols, u32 nnz, 3 x u32 reserved, u32 row_ptr[n+1],
 *                    u32 col_idx[nnz], f32 values[nnz]      (little endian)
 *   rhs.bin, x0.bin, x_ref.bin, x.bin   u32 n, f32[n]
 *   block_starts.bin u32 len, u32[len]
 */
#ifndef HBJ_IO_H
#define HBJ_IO_H

#include <stdint.h>

#include "hbj.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hbj_case hbj_case;

/* load_case_dir (crates/wgpu_solver_backend/src/io/loaders.rs:8-42) with the
 * validations of load_csr_matrix_bin (:44-124), load_vector_bin (:126-144),
 * load_block_starts_bin (:146-162) and validate_block_starts (:164-184). */
hbj_status hbj_case_load(const char* case_dir, hbj_case** out);
void hbj_case_free(hbj_case* c);
/* Borrowed views of the loaded arrays (valid until hbj_case_free). */
hbj_status hbj_case_view(const hbj_case* c, uint32_t* n, uint32_t* nnz, uint32_t* n_starts,
                         const uint32_t** row_ptr, const uint32_t** col_idx,
                         const float** values, const float** b, const float** x0,
                         const uint32_t** block_starts);

/* text-to-bin (crates/wgpu_solver_backend_convert_cli/src/main.rs:104-494):
 * parse a dataset text file and write the .bin files (+ x_ref.bin when the
 * dataset has x_ref_f32, + meta.json) into out_dir. */
hbj_status hbj_text_to_bin(const char* input_txt, const char* out_dir, int force);

/* write_x_bin (cli main.rs:445-464) / read_f32_vec_bin (cli main.rs:560-578).
 * hbj_vec_bin_read with x == NULL only reports the length. */
hbj_status hbj_vec_bin_write(const char* path, const float* x, uint32_t n);
hbj_status hbj_vec_bin_read(const char* path, float* x, uint32_t capacity, uint32_t* n);

typedef struct hbj_compare_stats {
  uint32_t n;
  uint32_t n_worst;   /* entries written to worst_idx (<= top_k) */
  double l2_ref;      /* ||x_ref||_2 (FP64 accumulation) */
  double l2_err;      /* ||x - x_ref||_2 */
  double rel_l2;      /* l2_err / max(l2_ref, 1e-30) */
  double rmse;        /* sqrt(sum err^2 / max(n, 1)) */
  float max_abs_err;
  float reserved0;
} hbj_compare_stats;

/* compare_x_vectors (cli main.rs:603-662).  worst_idx (top_k entries, may be
 * NULL when top_k == 0) receives the indices of the largest |x - x_ref|,
 * largest first.  PASS rule of run_compare_x (:684-689):
 * rel_l2 <= rel_tol || rmse <= abs_tol. */
hbj_status hbj_compare_x(const float* x_ref, const float* x, uint32_t n, uint32_t top_k,
                         hbj_compare_stats* out, uint32_t* worst_idx);

#ifdef __cplusplus
}
#endif
#endif /* HBJ_IO_H */
