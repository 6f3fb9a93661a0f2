"""ctypes binding of libhbj.so (the C ABI declared in include/hbj.h).

Loading never falls back to anything: if the library is missing the import of
the entry points raises, so no solve can silently run elsewhere.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# HBJ_LIB_PATH selects another build of the same library (e.g. a debug build)
LIB_PATH = os.environ.get("HBJ_LIB_PATH") or os.path.join(PKG, "libhbj.so")

c_u32, c_u64, c_i32, c_f32, c_f64 = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32,
                                     ctypes.c_float, ctypes.c_double)
c_void_p, c_char_p = ctypes.c_void_p, ctypes.c_char_p
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class hbj_result(ctypes.Structure):
    _fields_ = [("status", c_i32), ("iterations", c_u32), ("outer_iterations", c_u32),
                ("rr", c_f32), ("b_norm2", c_f32), ("reserved0", c_f32),
                ("rel_res_recurrence", c_f64), ("rel_res_true_fp64", c_f64), ("solve_ms", c_f64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


class hbj_ir_params(ctypes.Structure):
    _fields_ = [("max_outer", c_u32), ("max_inner", c_u32), ("inner_rel_tol", c_f32),
                ("reserved0", c_f32), ("target_rel_res", c_f64)]


class hbj_server_config(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", c_u64), ("max_active", c_u32), ("chunk_iters", c_u32)]


class hbj_server_stats(ctypes.Structure):
    _fields_ = [("submitted", c_u64), ("completed", c_u64), ("rejected", c_u64), ("active", c_u64),
                ("queued", c_u64), ("device_bytes_in_use", c_u64), ("peak_device_bytes", c_u64),
                ("peak_active", c_u64), ("chunks", c_u64), ("table_changes", c_u64),
                ("arena_waits", c_u64), ("admit_ms", c_f64), ("table_ms", c_f64),
                ("upload_bytes", c_u64), ("chunks_cached_spmv", c_u64)]


class hbj_system_info(ctypes.Structure):
    _fields_ = [("n", c_u32), ("nnz", c_u32), ("n_blocks", c_u32), ("n_tiles", c_u32),
                ("local_tiles", c_u32), ("spmv_variant", c_u32), ("local_columns", c_u64),
                ("device_bytes", c_u64), ("spmv_csr_bytes", c_u64), ("spmv_layout_bytes", c_u64),
                ("spmv_sliced_bytes", c_u64), ("launches_per_iteration", c_u32),
                ("sliced_nz_waves", c_u32)]


class hbj_sliced_info(ctypes.Structure):
    _fields_ = [("sliced", c_u32), ("waves", c_u32), ("nz_waves", c_u32), ("reserved0", c_u32),
                ("value_words", c_u64), ("run_words", c_u64)]


class hbj_compare_stats(ctypes.Structure):
    _fields_ = [("n", c_u32), ("n_worst", c_u32), ("l2_ref", c_f64), ("l2_err", c_f64),
                ("rel_l2", c_f64), ("rmse", c_f64), ("max_abs_err", c_f32), ("reserved0", c_f32)]


P = ctypes.POINTER
_SIGS = {
    "hbj_version": (c_char_p, []),
    "hbj_last_error": (c_char_p, []),
    "hbj_ctx_create": (c_i32, [ctypes.c_int, P(c_void_p)]),
    "hbj_ctx_destroy": (None, [c_void_p]),
    "hbj_ctx_describe": (c_i32, [c_void_p, ctypes.c_char_p, ctypes.c_size_t]),
    "hbj_lu_factor_inplace_6": (c_i32, [f32p, c_u32]),
    "hbj_build_lu_blocks_6": (c_i32, [c_u32, u32p, c_u32, u32p, c_u32, f32p, c_u32, u32p, c_u32, f32p]),
    "hbj_sliced_layout_info": (c_i32, [c_u32, c_u32, u32p, c_void_p, c_void_p, c_i32, c_void_p]),
    "hbj_spmv_tile_tables": (c_i32, [c_u32, c_u32, u32p, u32p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_u64, P(c_u64)]),
    "hbj_op_axpy": (c_i32, [c_void_p, c_u32, c_f32, f32p, f32p]),
    "hbj_op_scale": (c_i32, [c_void_p, c_u32, c_f32, f32p]),
    "hbj_op_dot": (c_i32, [c_void_p, c_u32, f32p, f32p, P(c_f32)]),
    "hbj_op_dot_noalloc": (c_i32, [c_void_p, c_u32, c_void_p, c_void_p, c_void_p, c_void_p]),
    "hbj_op_spmv": (c_i32, [c_void_p, c_u32, c_u32, u32p, u32p, f32p, f32p, f32p]),
    "hbj_op_block_jacobi": (c_i32, [c_void_p, c_u32, f32p, u32p, c_u32, f32p, f32p]),
    "hbj_op_update_scalars": (c_i32, [c_void_p, f32p, c_u32, c_u32, c_u32, c_u32, c_u32, c_u32, c_u32]),
    "hbj_system_create": (c_i32, [c_void_p, c_u32, c_u32, u32p, u32p, f32p, u32p, c_u32, P(c_void_p)]),
    "hbj_system_destroy": (None, [c_void_p]),
    "hbj_system_device_bytes": (c_u64, [c_void_p]),
    "hbj_system_get_info": (c_i32, [c_void_p, P(hbj_system_info)]),
    "hbj_system_get_lu": (c_i32, [c_void_p, c_void_p]),
    "hbj_system_set_spmv_variant": (c_i32, [c_void_p, c_u32]),
    "hbj_system_set_algorithm": (c_i32, [c_void_p, c_u32]),
    "hbj_system_autotune": (c_i32, [c_void_p, c_u32, P(c_u32), f64p]),
    "hbj_solve": (c_i32, [c_void_p, f32p, f32p, c_u32, c_f32, c_f32, P(hbj_result)]),
    "hbj_system_set_rhs": (c_i32, [c_void_p, f32p, c_void_p]),
    "hbj_system_solve_resident": (c_i32, [c_void_p, c_u32, c_f32, c_f32, P(hbj_result)]),
    "hbj_system_get_x": (c_i32, [c_void_p, f32p]),
    "hbj_system_spmv": (c_i32, [c_void_p, f32p, f32p]),
    "hbj_system_set_kernel_timing": (c_i32, [c_void_p, c_u32]),
    "hbj_system_kernel_timing": (c_i32, [c_void_p, f64p, np.ctypeslib.ndpointer(dtype=np.uint64)]),
    "hbj_solve_ir": (c_i32, [c_void_p, f64p, f64p, P(hbj_ir_params), P(hbj_result)]),
    "hbj_system_set_rhs64": (c_i32, [c_void_p, f64p, c_void_p]),
    "hbj_system_get_x64": (c_i32, [c_void_p, f64p]),
    "hbj_batch_create": (c_i32, [c_void_p, P(c_void_p), c_u32, P(c_void_p)]),
    "hbj_batch_destroy": (None, [c_void_p]),
    "hbj_batch_solve": (c_i32, [c_void_p, c_u32, c_f32, c_f32, P(hbj_result)]),
    "hbj_batch_solve_ir": (c_i32, [c_void_p, P(hbj_ir_params), P(hbj_result)]),
    "hbj_batch_autotune": (c_i32, [c_void_p, c_u32, ctypes.c_int, P(c_u32), f64p]),
    "hbj_batch_set_kernel_timing": (c_i32, [c_void_p, c_u32]),
    "hbj_batch_kernel_samples": (c_i32, [c_void_p, c_void_p, c_u64, P(c_u64), c_void_p]),
    "hbj_server_create": (c_i32, [c_void_p, P(hbj_server_config), P(c_void_p)]),
    "hbj_server_destroy": (None, [c_void_p]),
    "hbj_server_submit": (c_i32, [c_void_p, c_u32, c_u32, u32p, u32p, f32p, u32p, c_u32, f64p,
                                  P(hbj_ir_params), P(c_u64)]),
    "hbj_server_wait": (c_i32, [c_void_p, c_u64, c_void_p, P(hbj_result), P(c_f64)]),
    "hbj_server_stats_get": (c_i32, [c_void_p, P(hbj_server_stats)]),
    "hbj_server_request_bytes": (c_u64, [c_u32, c_u32, u32p, c_void_p, c_void_p, u32p, c_u32]),
    "hbj_case_load": (c_i32, [c_char_p, P(c_void_p)]),
    "hbj_case_free": (None, [c_void_p]),
    "hbj_case_view": (c_i32, [c_void_p, P(c_u32), P(c_u32), P(c_u32)] + [P(c_void_p)] * 6),
    "hbj_text_to_bin": (c_i32, [c_char_p, c_char_p, ctypes.c_int]),
    "hbj_vec_bin_write": (c_i32, [c_char_p, f32p, c_u32]),
    "hbj_vec_bin_read": (c_i32, [c_char_p, c_void_p, c_u32, P(c_u32)]),
    "hbj_compare_x": (c_i32, [f32p, f32p, c_u32, c_u32, P(hbj_compare_stats), c_void_p]),
    "hbj_stream_copy": (c_i32, [c_void_p, c_u64, c_u32, P(c_f64)]),
    "hbj_lattice_size": (c_i32, [c_u32, P(c_u32), P(c_u64)]),
    "hbj_lattice_generate": (c_i32, [c_u32, c_u64, c_f64, u32p, u32p, f32p, f32p, u32p]),
    "hbj_beam_element_stiffness": (c_i32, [c_f64, c_f64, c_f64, c_f64, c_f64, c_f64, c_f64, ctypes.c_int, f64p]),
}

_lib = None


def lib():
    """The loaded libhbj.so.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -m wgpu_solver_backend_amd._build` (no fallback path exists)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            if not hasattr(L, name):
                continue  # reported by tests/test_abi.py
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().hbj_last_error().decode()
