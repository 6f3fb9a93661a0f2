"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: imported only by tests, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(ORACLE_DIR)
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None

STATUS = {0: "OK", 1: "INVALID", 2: "ZERO_PIVOT", 3: "BREAKDOWN_PAP", 4: "BREAKDOWN_RZ",
          5: "NOT_CONVERGED"}


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        build()
    L = ctypes.CDLL(ORACLE_SO)
    u32, f32, f64 = ctypes.c_uint32, ctypes.c_float, ctypes.c_double
    L.or_set_threads.argtypes = [ctypes.c_int]
    L.or_get_threads.restype = ctypes.c_int
    L.or_last_error.restype = ctypes.c_char_p
    L.or_lu_factor_inplace_6.argtypes = [_f32p, u32]
    L.or_build_lu_blocks_6.argtypes = [u32, _u32p, u32, _u32p, _f32p, u32, _u32p, u32, _f32p]
    L.or_spmv.argtypes = [u32, _u32p, _u32p, _f32p, _f32p, _f32p]
    L.or_dot.argtypes = [u32, _f32p, _f32p]
    L.or_dot.restype = f32
    L.or_axpy.argtypes = [u32, f32, _f32p, _f32p]
    L.or_scale.argtypes = [u32, f32, _f32p]
    L.or_block_jacobi.argtypes = [u32, _f32p, _u32p, u32, _f32p, _f32p]
    L.or_update_scalars.argtypes = [_f32p, u32, u32, u32, u32, u32, u32]
    L.or_pcg.argtypes = [u32, _u32p, _u32p, _f32p, _f32p, _u32p, u32, _f32p, _f32p, u32, f32, f32,
                         ctypes.POINTER(u32), ctypes.c_void_p]
    L.or_ccg.argtypes = [u32, _u32p, _u32p, _f32p, _f32p, _u32p, u32, _f32p, _f32p, u32, f32, f32,
                         ctypes.POINTER(u32)]
    L.or_spmv64.argtypes = [u32, _u32p, _u32p, _f32p, _f64p, _f64p]
    L.or_true_rel_residual64.argtypes = [u32, _u32p, _u32p, _f32p, _f64p, _f64p]
    L.or_true_rel_residual64.restype = f64
    L.or_pcg64.argtypes = [u32, _u32p, _u32p, _f32p, _u32p, u32, _f64p, _f64p, u32, f64,
                           ctypes.POINTER(u32)]
    _lib = L
    return L


def set_threads(n):
    lib().or_set_threads(int(n))


def last_error():
    return lib().or_last_error().decode()


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def lu_factor_inplace_6(mat36, m):
    mat = _c(mat36, np.float32).copy()
    st = lib().or_lu_factor_inplace_6(mat, m)
    return st, mat


def build_lu_blocks_6(n, row_ptr, col_idx, values, block_starts):
    row_ptr, col_idx = _c(row_ptr, np.uint32), _c(col_idx, np.uint32)
    values, bs = _c(values, np.float32), _c(block_starts, np.uint32)
    nb = max(bs.size - 1, 0)
    out = np.zeros(max(36 * nb, 1), dtype=np.float32)
    st = lib().or_build_lu_blocks_6(n, row_ptr, row_ptr.size, col_idx, values, values.size, bs,
                                    bs.size, out)
    return st, out[: 36 * nb]


def spmv(row_ptr, col_idx, values, x):
    n = len(row_ptr) - 1
    y = np.zeros(max(n, 1), dtype=np.float32)
    lib().or_spmv(n, _c(row_ptr, np.uint32), _c(col_idx, np.uint32), _c(values, np.float32),
                  _c(x, np.float32), y)
    return y[:n]


def dot(a, b):
    a, b = _c(a, np.float32), _c(b, np.float32)
    if a.size == 0:
        a = np.zeros(1, np.float32)
        b = np.zeros(1, np.float32)
        return float(lib().or_dot(0, a, b))
    return float(lib().or_dot(a.size, a, b))


def axpy(alpha, x, y):
    y = _c(y, np.float32).copy()
    lib().or_axpy(y.size, alpha, _c(x, np.float32), y)
    return y


def scale(beta, x):
    x = _c(x, np.float32).copy()
    lib().or_scale(x.size, beta, x)
    return x


def block_jacobi(n, lu, block_starts, r, z0=None):
    z = np.zeros(n, np.float32) if z0 is None else _c(z0, np.float32).copy()
    lu = _c(lu, np.float32)
    if lu.size == 0:
        lu = np.zeros(1, np.float32)
    bs = _c(block_starts, np.uint32)
    lib().or_block_jacobi(n, lu, bs, bs.size, _c(r, np.float32), z)
    return z


def update_scalars(slots, i_pap, i_rz_new, i_rz_old, i_alpha, i_malpha, i_beta):
    s = _c(slots, np.float32).copy()
    lib().or_update_scalars(s, i_pap, i_rz_new, i_rz_old, i_alpha, i_malpha, i_beta)
    return s


def pcg(row_ptr, col_idx, values, lu, block_starts, b, x0, max_iter, rel_tol, abs_tol,
        history=False):
    """Returns (status, iterations, x, rr_history or None).  x is x0 unless status == 0."""
    n = len(row_ptr) - 1
    x = _c(x0, np.float32).copy()
    it = ctypes.c_uint32(0)
    hist = np.zeros(max(max_iter, 1), np.float32) if history else None
    lu = _c(lu, np.float32)
    if lu.size == 0:
        lu = np.zeros(1, np.float32)
    st = lib().or_pcg(n, _c(row_ptr, np.uint32), _c(col_idx, np.uint32), _c(values, np.float32),
                      lu, _c(block_starts, np.uint32), len(block_starts), _c(b, np.float32), x,
                      max_iter, rel_tol, abs_tol, ctypes.byref(it),
                      hist.ctypes.data if history else None)
    return st, it.value, x, (hist[: it.value] if history else None)


def ccg(row_ptr, col_idx, values, lu, block_starts, b, x0, max_iter, rel_tol, abs_tol):
    """One-reduction PCG: (status, iterations, x); x is x0 unless status == 0."""
    n = len(row_ptr) - 1
    x = _c(x0, np.float32).copy()
    it = ctypes.c_uint32(0)
    lu = _c(lu, np.float32)
    if lu.size == 0:
        lu = np.zeros(1, np.float32)
    st = lib().or_ccg(n, _c(row_ptr, np.uint32), _c(col_idx, np.uint32), _c(values, np.float32),
                      lu, _c(block_starts, np.uint32), len(block_starts), _c(b, np.float32), x,
                      max_iter, rel_tol, abs_tol, ctypes.byref(it))
    return st, it.value, x


def true_rel_residual64(row_ptr, col_idx, values, b, x):
    n = len(row_ptr) - 1
    return float(lib().or_true_rel_residual64(n, _c(row_ptr, np.uint32), _c(col_idx, np.uint32),
                                               _c(values, np.float32), _c(b, np.float64),
                                               _c(x, np.float64)))


def pcg64(row_ptr, col_idx, values, block_starts, b, x0, max_iter, rel_tol):
    n = len(row_ptr) - 1
    x = _c(x0, np.float64).copy()
    it = ctypes.c_uint32(0)
    st = lib().or_pcg64(n, _c(row_ptr, np.uint32), _c(col_idx, np.uint32), _c(values, np.float32),
                        _c(block_starts, np.uint32), len(block_starts), _c(b, np.float64), x,
                        max_iter, rel_tol, ctypes.byref(it))
    return st, it.value, x


GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_dataset(k):
    d = np.load(os.path.join(GOLDEN, f"ds{k}.npz"))
    return {key: d[key] for key in d.files}
