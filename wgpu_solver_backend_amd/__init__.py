"""MI355X-native block-Jacobi PCG behind the C ABI of include/hbj.h.

The names mirror the reference crate ``wgpu_solver_backend`` so that code
and tests written against it read the same:

    reference (crates/wgpu_solver_backend/src)        here
    gpu/context.rs      GpuContext::create             GpuContext.create
    compute.rs:37       lu_factor_inplace_6            lu_factor_inplace_6
    compute.rs:84       build_lu_blocks_from_csr_...   build_lu_blocks_from_csr_block_starts_6
    compute/*_exec.rs   SpmvExecutor, DotScalarExecutor, VecOpsExecutor,
                        BlockJacobiExecutor, PcgUpdateScalarsExecutor (same names)
    compute.rs:189      pcg_block_jacobi_csr_wgpu      pcg_block_jacobi_csr_wgpu

``Result<usize, String>`` maps to "returns the iteration count or raises
SolverError(status, message)" with the reference's message texts; ``x`` is
written only on success.  ``System``, ``Batch`` and ``Server`` expose the
device-resident, batched and served forms.  All compute runs in libhbj.so
(HIP, gfx950); there is no other path.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import hbj_ir_params, hbj_result, hbj_server_config, hbj_server_stats

__all__ = [
    "SolverError", "GpuContext", "lu_factor_inplace_6", "build_lu_blocks_from_csr_block_starts_6",
    "SpmvExecutor", "DotScalarExecutor", "VecOpsExecutor", "BlockJacobiExecutor",
    "PcgUpdateScalarsExecutor", "pcg_block_jacobi_csr_wgpu", "System", "Batch", "Server",
    "IrParams", "lattice", "lattice_size", "beam_element_stiffness", "text_to_bin",
    "load_case_dir", "write_x_bin", "read_x_bin", "compare_x", "CLI_PATH", "ITER_KERNELS",
    "ITER_KERNEL_NAMES", "SPMV_VARIANTS", "ALG_PCG", "ALG_PCG_ONE_REDUCTION",
]

import os as _os

CLI_PATH = _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "hbj_cli")
# launches per PCG iteration (HBJ_ITER_KERNELS in include/hbj.h)
ITER_KERNELS = 3
ITER_KERNEL_NAMES = ("spmv", "update", "finish")
# bit-identical SpMV variants (HBJ_SPMV_VARIANTS)
SPMV_VARIANTS = 6
# hbj_system_set_algorithm
ALG_PCG = 0
ALG_PCG_ONE_REDUCTION = 1

STATUS_NAMES = {0: "OK", 1: "INVALID", 2: "ZERO_PIVOT", 3: "BREAKDOWN_PAP", 4: "BREAKDOWN_RZ",
                5: "NOT_CONVERGED", 6: "DEVICE", 7: "OOM", 8: "NO_DEVICE", 9: "REJECTED"}


class SolverError(RuntimeError):
    """A non-OK hbj_status with the library's (reference-worded) message."""

    def __init__(self, status, message, result=None):
        super().__init__(message)
        self.status = int(status)
        self.status_name = STATUS_NAMES.get(self.status, str(status))
        self.result = result


def _check(st, result=None):
    if st != 0:
        raise SolverError(st, _lib.last_error(), result)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _nz(a, dt):
    """ndpointer arguments need a non-empty buffer even for n == 0."""
    return a if a.size else np.zeros(1, dtype=dt)


# ----------------------------------------------------------------- context

class GpuContext:
    """GpuContext (src/gpu/context.rs:39-177): one HIP device."""

    def __init__(self, handle, device):
        self._h = handle
        self.device = device

    @classmethod
    def create(cls, device=0):
        h = ctypes.c_void_p()
        _check(_lib.lib().hbj_ctx_create(int(device), ctypes.byref(h)))
        return cls(h, device)

    def describe(self):
        buf = ctypes.create_string_buffer(1024)
        _check(_lib.lib().hbj_ctx_describe(self._h, buf, len(buf)))
        return buf.value.decode()

    def stream_copy_us(self, nbytes, reps=20):
        """Mean duration (us) of a plain copy kernel moving nbytes in and nbytes out."""
        us = ctypes.c_double()
        _check(_lib.lib().hbj_stream_copy(self._h, int(nbytes), int(reps), ctypes.byref(us)))
        return us.value

    def close(self):
        if self._h:
            _lib.lib().hbj_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ------------------------------------------------------------ CPU setup

def lu_factor_inplace_6(mat, n):
    """compute.rs:37-71 on a 36-float row-major buffer; returns the factored copy."""
    m = _f32(mat).copy()
    if m.size < 36:
        raise SolverError(1, "lu_factor_inplace_6: mat must have len >= 36")
    _check(_lib.lib().hbj_lu_factor_inplace_6(m, int(n)))
    return m


def build_lu_blocks_from_csr_block_starts_6(n, row_ptr, col_idx, values, block_starts):
    """compute.rs:84-181: packed 6x6 LU per block (36 floats each)."""
    rp, ci, va, bs = _u32(row_ptr), _u32(col_idx), _f32(values), _u32(block_starts)
    nb = max(bs.size - 1, 0)
    out = np.zeros(max(36 * nb, 1), np.float32)
    _check(_lib.lib().hbj_build_lu_blocks_6(int(n), _nz(rp, np.uint32), rp.size, _nz(ci, np.uint32),
                                            ci.size, _nz(va, np.float32), va.size,
                                            _nz(bs, np.uint32), bs.size, out))
    return out[: 36 * nb]


def spmv_tile_tables(row_ptr, col_idx):
    """Tile-local SpMV tables of a CSR matrix (hbj_spmv_tile_tables): per 256-row
    tile (offset, count) into the column lists (count 0xFFFFFFFF: no list), each
    entry's index in its tile's list, and the concatenated sorted lists."""
    rp, ci = _u32(row_ptr), _u32(col_idx)
    n, nnz = rp.size - 1, ci.size
    nt = (n + 255) // 256
    off = np.zeros(max(nt, 1), np.uint32)
    cnt = np.zeros(max(nt, 1), np.uint32)
    lidx = np.zeros(max(nnz, 1), np.uint16)
    total = ctypes.c_uint64()
    L = _lib.lib()
    args = (n, nnz, _nz(rp, np.uint32), _nz(ci, np.uint32), off.ctypes.data, cnt.ctypes.data,
            lidx.ctypes.data)
    st = L.hbj_spmv_tile_tables(*args, None, 0, ctypes.byref(total))  # sizes the lists
    tcols = np.zeros(max(total.value, 1), np.uint32)
    if st != 0 and total.value > 0:
        st = L.hbj_spmv_tile_tables(*args, tcols.ctypes.data, total.value, ctypes.byref(total))
    _check(st)
    return off[:nt], cnt[:nt], lidx[:nnz], tcols[: total.value]


# ------------------------------------------------------------- executors

def sliced_layout_info(row_ptr, col_idx, values, allow_nz=True):
    """The sliced SpMV layout a system of this matrix gets (host only): dict of
    sliced (0/1), waves, nz_waves (waves storing only their nonzero values),
    value_words and run_words."""
    rp = _u32(row_ptr)
    ci = _u32(col_idx)
    va = _f32(values)
    inf = _lib.hbj_sliced_info()
    ci1, va1 = _nz(ci, np.uint32), _nz(va, np.float32)
    _check(_lib.lib().hbj_sliced_layout_info(rp.size - 1, ci.size, rp, ci1.ctypes.data,
                                             va1.ctypes.data, 1 if allow_nz else 0,
                                             ctypes.byref(inf)))
    return {f: getattr(inf, f) for f, _ in inf._fields_ if f != "reserved0"}


class VecOpsExecutor:
    """VecOpsExecutor (compute/vec_ops_exec.rs:23-239)."""

    def __init__(self, ctx):
        self.ctx = ctx

    @classmethod
    def create(cls, ctx):
        return cls(ctx)

    def axpy_inplace(self, x, y, alpha):
        """y = y + alpha*x (vec_ops_exec.rs:130-160); returns the new y."""
        x, y = _f32(x), _f32(y).copy()
        _check(_lib.lib().hbj_op_axpy(self.ctx._h, y.size, float(alpha), _nz(x, np.float32),
                                      _nz(y, np.float32)))
        return y

    def scale_inplace(self, x, beta):
        """x = x*beta (vec_ops_exec.rs:200-232); returns the new x."""
        x = _f32(x).copy()
        _check(_lib.lib().hbj_op_scale(self.ctx._h, x.size, float(beta), _nz(x, np.float32)))
        return x


class DotScalarExecutor:
    """DotScalarExecutor (compute/dot_scalar_exec.rs:16-271)."""

    def __init__(self, ctx, n_max, scalar_results_len):
        self.ctx, self.n_max, self.slots = ctx, n_max, np.zeros(scalar_results_len, np.float32)

    @classmethod
    def create(cls, ctx, n_max, scalar_results_len):
        return cls(ctx, n_max, scalar_results_len)

    def dot(self, a, b):
        a, b = _f32(a), _f32(b)
        out = ctypes.c_float()
        _check(_lib.lib().hbj_op_dot(self.ctx._h, a.size, _nz(a, np.float32), _nz(b, np.float32),
                                     ctypes.byref(out)))
        return out.value

    def dot_scalar_into(self, a, b, out_index):
        """encode_dot_scalar_into (:151-247) into the host copy of the slots."""
        if out_index >= self.slots.size:
            raise IndexError("DotScalarExecutor: out_index out of range")
        self.slots[out_index] = self.dot(a, b)
        return self.slots[out_index]


class SpmvExecutor:
    """SpmvExecutor (compute/spmv_exec.rs:17-148): owns a copy of the CSR."""

    def __init__(self, ctx, n_rows, row_ptr, col_idx, values):
        self.ctx, self.n_rows = ctx, int(n_rows)
        self.row_ptr, self.col_idx, self.values = _u32(row_ptr), _u32(col_idx), _f32(values)

    @classmethod
    def create(cls, ctx, n_rows, row_ptr, col_idx, values):
        return cls(ctx, n_rows, row_ptr, col_idx, values)

    def spmv(self, x):
        y = np.zeros(max(self.n_rows, 1), np.float32)
        _check(_lib.lib().hbj_op_spmv(self.ctx._h, self.n_rows, self.col_idx.size, self.row_ptr,
                                      _nz(self.col_idx, np.uint32), _nz(self.values, np.float32),
                                      _nz(_f32(x), np.float32), y))
        return y[: self.n_rows]


class BlockJacobiExecutor:
    """BlockJacobiExecutor (compute/block_jacobi_exec.rs:25-133)."""

    def __init__(self, ctx, n, lu_blocks, block_starts):
        self.ctx, self.n = ctx, int(n)
        self.lu_blocks, self.block_starts = _f32(lu_blocks), _u32(block_starts)

    @classmethod
    def create(cls, ctx, n, lu_blocks, block_starts):
        return cls(ctx, n, lu_blocks, block_starts)

    def apply(self, r, z=None):
        """z = M^-1 r (encode_apply :96-124); rows no block writes keep z's value."""
        z = np.zeros(self.n, np.float32) if z is None else _f32(z).copy()
        _check(_lib.lib().hbj_op_block_jacobi(self.ctx._h, self.n, _nz(self.lu_blocks, np.float32),
                                              self.block_starts, self.block_starts.size,
                                              _nz(_f32(r), np.float32), _nz(z, np.float32)))
        return z


class PcgUpdateScalarsExecutor:
    """PcgUpdateScalarsExecutor (compute/pcg_update_scalars_exec.rs:26-119)."""

    def __init__(self, ctx):
        self.ctx = ctx

    @classmethod
    def create(cls, ctx):
        return cls(ctx)

    def update_scalars(self, slots, p_ap, rz_new, rz_old, alpha, minus_alpha, beta):
        s = _f32(slots).copy()
        _check(_lib.lib().hbj_op_update_scalars(self.ctx._h, s, s.size, p_ap, rz_new, rz_old, alpha,
                                                minus_alpha, beta))
        return s


# ------------------------------------------------------------------ system

class IrParams:
    def __init__(self, max_outer=30, max_inner=5000, inner_rel_tol=1e-4, target_rel_res=1e-10):
        self.max_outer, self.max_inner = max_outer, max_inner
        self.inner_rel_tol, self.target_rel_res = inner_rel_tol, target_rel_res

    def c(self):
        return hbj_ir_params(self.max_outer, self.max_inner, self.inner_rel_tol, 0.0,
                             self.target_rel_res)


class System:
    """A CSR system resident in HBM (CSR, LU blocks, PCG vectors)."""

    def __init__(self, ctx, row_ptr, col_idx, values, block_starts):
        self.ctx = ctx
        self.row_ptr, self.col_idx = _u32(row_ptr), _u32(col_idx)
        self.values, self.block_starts = _f32(values), _u32(block_starts)
        self.n = self.row_ptr.size - 1
        self.nnz = self.col_idx.size
        h = ctypes.c_void_p()
        _check(_lib.lib().hbj_system_create(ctx._h, self.n, self.nnz, self.row_ptr,
                                            _nz(self.col_idx, np.uint32),
                                            _nz(self.values, np.float32), self.block_starts,
                                            self.block_starts.size, ctypes.byref(h)))
        self._h = h

    @property
    def device_bytes(self):
        return int(_lib.lib().hbj_system_device_bytes(self._h))

    def info(self):
        """Sizes, tile-local SpMV coverage, device bytes, the SpMV variant of
        the next solve and the bytes one SpMV streams (CSR accounting and the
        tile-local layout), as a dict."""
        inf = _lib.hbj_system_info()
        _check(_lib.lib().hbj_system_get_info(self._h, ctypes.byref(inf)))
        return {f: getattr(inf, f) for f, _ in inf._fields_}

    def lu(self):
        """The system's LU blocks (float32 [n_blocks, 6, 6]), as built on the
        device from its CSR when the system was created."""
        nb = self.info()["n_blocks"]
        out = np.zeros((nb, 6, 6), np.float32)
        _check(_lib.lib().hbj_system_get_lu(self._h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def set_spmv_variant(self, variant):
        """Select one of the bit-identical SpMV variants (0..SPMV_VARIANTS-1)."""
        _check(_lib.lib().hbj_system_set_spmv_variant(self._h, int(variant)))

    def set_algorithm(self, alg):
        """ALG_PCG (the reference's PCG, default) or ALG_PCG_ONE_REDUCTION
        (Chronopoulos-Gear: one reduction point, two launches per iteration;
        rounding differs from the reference's) for later plain solves."""
        _check(_lib.lib().hbj_system_set_algorithm(self._h, int(alg)))

    def autotune(self, iters=50):
        """Time `iters` PCG iterations with every SpMV variant, keep the fastest.
        Returns (best_variant, microseconds per iteration of each variant)."""
        best = ctypes.c_uint32(0)
        us = np.zeros(SPMV_VARIANTS, np.float64)
        _check(_lib.lib().hbj_system_autotune(self._h, int(iters), ctypes.byref(best), us))
        return int(best.value), us

    def solve(self, b, x0, max_iter, rel_tol, abs_tol, raise_on_error=True):
        """hbj_solve: returns (x, result).  x is x0 unless status is OK."""
        x = _f32(x0).copy()
        res = hbj_result()
        st = _lib.lib().hbj_solve(self._h, _nz(_f32(b), np.float32), _nz(x, np.float32),
                                  int(max_iter), float(rel_tol), float(abs_tol), ctypes.byref(res))
        if raise_on_error:
            _check(st, res)
        return x, res

    def set_rhs(self, b, x0=None):
        self._keep = None if x0 is None else _nz(_f32(x0), np.float32)
        _check(_lib.lib().hbj_system_set_rhs(self._h, _nz(_f32(b), np.float32),
                                             None if x0 is None else self._keep.ctypes.data))

    def solve_resident(self, max_iter, rel_tol, abs_tol, raise_on_error=True):
        res = hbj_result()
        st = _lib.lib().hbj_system_solve_resident(self._h, int(max_iter), float(rel_tol),
                                                  float(abs_tol), ctypes.byref(res))
        if raise_on_error:
            _check(st, res)
        return res

    def get_x(self):
        x = np.zeros(max(self.n, 1), np.float32)
        _check(_lib.lib().hbj_system_get_x(self._h, x))
        return x[: self.n]

    def spmv(self, x):
        """y = A x through the system's own SpMV (resident layout, chosen
        variant): SpmvExecutor::encode_spmv, spmv_exec.rs:122-147."""
        xx = np.zeros(max(self.n, 1), np.float32)
        xx[: self.n] = _f32(x)
        y = np.zeros(max(self.n, 1), np.float32)
        _check(_lib.lib().hbj_system_spmv(self._h, xx, y))
        return y[: self.n]

    def solve_ir(self, b64, x0_64=None, params=None, raise_on_error=True):
        p = (params or IrParams()).c()
        x = np.zeros(self.n, np.float64) if x0_64 is None else _f64(x0_64).copy()
        res = hbj_result()
        st = _lib.lib().hbj_solve_ir(self._h, _nz(_f64(b64), np.float64), _nz(x, np.float64),
                                     ctypes.byref(p), ctypes.byref(res))
        if raise_on_error:
            _check(st, res)
        return x, res

    def set_rhs64(self, b64, x0_64=None):
        self._keep64 = None if x0_64 is None else _f64(x0_64)
        _check(_lib.lib().hbj_system_set_rhs64(self._h, _nz(_f64(b64), np.float64),
                                               None if x0_64 is None else self._keep64.ctypes.data))

    def get_x64(self):
        x = np.zeros(max(self.n, 1), np.float64)
        _check(_lib.lib().hbj_system_get_x64(self._h, x))
        return x[: self.n]

    def set_kernel_timing(self, every):
        _check(_lib.lib().hbj_system_set_kernel_timing(self._h, int(every)))

    def kernel_timing(self):
        """Mean HIP-event duration (us) and sample count of each launch of an
        iteration: (spmv, update with alpha and the block solve, finish with
        beta, the stop test and the p-update)."""
        avg = np.zeros(ITER_KERNELS, np.float64)
        cnt = np.zeros(ITER_KERNELS, np.uint64)
        _check(_lib.lib().hbj_system_kernel_timing(self._h, avg, cnt))
        return avg, cnt

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().hbj_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent systems advanced together (hbj_batch_*)."""

    def __init__(self, ctx, systems):
        self.ctx, self.systems = ctx, list(systems)
        arr = (ctypes.c_void_p * len(self.systems))(*[s._h for s in self.systems])
        h = ctypes.c_void_p()
        _check(_lib.lib().hbj_batch_create(ctx._h, arr, len(self.systems), ctypes.byref(h)))
        self._h = h

    def solve(self, max_iter, rel_tol, abs_tol):
        res = (hbj_result * len(self.systems))()
        st = _lib.lib().hbj_batch_solve(self._h, int(max_iter), float(rel_tol), float(abs_tol), res)
        return st, list(res)

    def solve_ir(self, params=None):
        p = (params or IrParams()).c()
        res = (hbj_result * len(self.systems))()
        st = _lib.lib().hbj_batch_solve_ir(self._h, ctypes.byref(p), res)
        return st, list(res)

    def autotune(self, iters=20, ir=False):
        """hbj_batch_autotune: (best SpMV variant, us per iteration of each)."""
        best = ctypes.c_uint32()
        us = np.zeros(SPMV_VARIANTS, np.float64)
        _check(_lib.lib().hbj_batch_autotune(self._h, int(iters), int(bool(ir)), ctypes.byref(best), us))
        return int(best.value), us

    def set_kernel_timing(self, every):
        """hbj_batch_set_kernel_timing: sample every `every`-th launch set (0 = off)."""
        _check(_lib.lib().hbj_batch_set_kernel_timing(self._h, int(every)))

    def kernel_samples(self):
        """hbj_batch_kernel_samples: (records, sets) of the last timed solve --
        records[i] = (launch set, then each kernel's us, -1 where not
        launched); sets[j] = launch sets that found system j running."""
        n = ctypes.c_uint64()
        _check(_lib.lib().hbj_batch_kernel_samples(self._h, None, 0, ctypes.byref(n), None))
        rec = np.zeros((max(int(n.value), 1), 1 + ITER_KERNELS), np.float64)
        sets = np.zeros(len(self.systems), np.uint32)
        _check(_lib.lib().hbj_batch_kernel_samples(self._h, rec.ctypes.data, int(n.value),
                                                   ctypes.byref(n), sets.ctypes.data))
        return rec[: int(n.value)], sets

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().hbj_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop (hbj_server_*)."""

    def __init__(self, ctx, hbm_budget_bytes, max_active=256, chunk_iters=64):
        cfg = hbj_server_config(int(hbm_budget_bytes), int(max_active), int(chunk_iters))
        h = ctypes.c_void_p()
        _check(_lib.lib().hbj_server_create(ctx._h, ctypes.byref(cfg), ctypes.byref(h)))
        self._h, self.ctx = h, ctx

    def submit(self, row_ptr, col_idx, values, block_starts, b64, params=None):
        p = (params or IrParams()).c()
        rp, ci, va, bs = _u32(row_ptr), _u32(col_idx), _f32(values), _u32(block_starts)
        t = ctypes.c_uint64()
        _check(_lib.lib().hbj_server_submit(self._h, rp.size - 1, ci.size, rp, _nz(ci, np.uint32),
                                            _nz(va, np.float32), bs, bs.size,
                                            _nz(_f64(b64), np.float64), ctypes.byref(p),
                                            ctypes.byref(t)))
        return t.value

    def wait(self, ticket, n=None, raise_on_error=False):
        res = hbj_result()
        lat = ctypes.c_double()
        x = np.zeros(max(n or 0, 1), np.float64) if n is not None else None
        st = _lib.lib().hbj_server_wait(self._h, ctypes.c_uint64(ticket),
                                        x.ctypes.data if x is not None else None, ctypes.byref(res),
                                        ctypes.byref(lat))
        if raise_on_error:
            _check(st, res)
        return st, (x[:n] if x is not None else None), res, lat.value

    @staticmethod
    def request_bytes(row_ptr, col_idx, values, block_starts):
        """Device bytes the request's system reserves from the budget (its exact
        block in the server's arena)."""
        rp, ci, va, bs = _u32(row_ptr), _u32(col_idx), _f32(values), _u32(block_starts)
        ci1, va1 = _nz(ci, np.uint32), _nz(va, np.float32)  # kept alive across the call
        nb = _lib.lib().hbj_server_request_bytes(rp.size - 1, ci.size, rp, ci1.ctypes.data,
                                                 va1.ctypes.data, bs, bs.size)
        if nb == 0:
            raise SolverError(1, _lib.last_error())
        return int(nb)

    def stats(self):
        s = hbj_server_stats()
        _check(_lib.lib().hbj_server_stats_get(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().hbj_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pcg_block_jacobi_csr_wgpu(n, b, x, max_iter, rel_tol, abs_tol, ctx, spmv_exec, vec_ops_exec,
                              dot_scalar_exec, block_jacobi_exec, pcg_update_scalars_exec):
    """compute.rs:189-495: returns the iteration count; x (float32 ndarray) is updated in
    place on success only; raises SolverError with the reference's texts otherwise."""
    b = _f32(b)
    if b.size != n or x.size != n:
        raise SolverError(1, f"PCG(BlockJacobiGpu): dimension mismatch: n={n}, b len {b.size}, "
                             f"x len {x.size}")
    key = (id(spmv_exec), id(block_jacobi_exec))
    sys_ = getattr(spmv_exec, "_system", None)
    if sys_ is None or getattr(spmv_exec, "_system_key", None) != key:
        sys_ = System(ctx, spmv_exec.row_ptr, spmv_exec.col_idx, spmv_exec.values,
                      block_jacobi_exec.block_starts)
        spmv_exec._system, spmv_exec._system_key = sys_, key
    xo, res = sys_.solve(b, x, max_iter, rel_tol, abs_tol)
    x[:] = xo
    return int(res.iterations)


# ------------------------------------------------------- synthetic lattice

def lattice_size(N):
    n = ctypes.c_uint32()
    nnz = ctypes.c_uint64()
    _check(_lib.lib().hbj_lattice_size(int(N), ctypes.byref(n), ctypes.byref(nnz)))
    return n.value, nnz.value


def lattice(N, seed=0, e_log10_half_range=0.0):
    """Beam-lattice system of N^3 nodes: dict of row_ptr, col_idx, values, rhs, block_starts."""
    n, nnz = lattice_size(N)
    rp = np.zeros(n + 1, np.uint32)
    ci = np.zeros(nnz, np.uint32)
    va = np.zeros(nnz, np.float32)
    b = np.zeros(n, np.float32)
    bs = np.zeros(n // 6 + 1, np.uint32)
    _check(_lib.lib().hbj_lattice_generate(int(N), int(seed), float(e_log10_half_range), rp, ci,
                                           va, b, bs))
    return {"n": n, "nnz": nnz, "row_ptr": rp, "col_idx": ci, "values": va, "rhs": b,
            "block_starts": bs, "N": N}


def beam_element_stiffness(E, G, A, Iy, Iz, J, L, direction):
    k = np.zeros(144, np.float64)
    _check(_lib.lib().hbj_beam_element_stiffness(E, G, A, Iy, Iz, J, L, int(direction), k))
    return k.reshape(12, 12)


# ------------------------------------------------------- case-directory IO

def text_to_bin(input_txt, out_dir, force=False):
    """convert main.rs:104-494 (text-to-bin)."""
    _check(_lib.lib().hbj_text_to_bin(str(input_txt).encode(), str(out_dir).encode(), int(force)))


def load_case_dir(case_dir):
    """loaders.rs:8-42: dict of row_ptr, col_idx, values, b, x0, block_starts (copies)."""
    L = _lib.lib()
    h = ctypes.c_void_p()
    _check(L.hbj_case_load(str(case_dir).encode(), ctypes.byref(h)))
    try:
        n, nnz, ns = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        ptrs = [ctypes.c_void_p() for _ in range(6)]
        _check(L.hbj_case_view(h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(ns),
                               *[ctypes.byref(p) for p in ptrs]))

        def arr(p, count, dt):
            if count == 0:
                return np.zeros(0, dt)
            ct = ctypes.c_uint32 if dt == np.uint32 else ctypes.c_float
            return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=(count,)).astype(dt).copy()

        N, NNZ, NS = n.value, nnz.value, ns.value
        return {"n": N, "nnz": NNZ, "row_ptr": arr(ptrs[0], N + 1, np.uint32),
                "col_idx": arr(ptrs[1], NNZ, np.uint32), "values": arr(ptrs[2], NNZ, np.float32),
                "b": arr(ptrs[3], N, np.float32), "x0": arr(ptrs[4], N, np.float32),
                "block_starts": arr(ptrs[5], NS, np.uint32)}
    finally:
        L.hbj_case_free(h)


def write_x_bin(path, x):
    x = _f32(x)
    _check(_lib.lib().hbj_vec_bin_write(str(path).encode(), _nz(x, np.float32), x.size))


def read_x_bin(path):
    L = _lib.lib()
    n = ctypes.c_uint32()
    _check(L.hbj_vec_bin_read(str(path).encode(), None, 0, ctypes.byref(n)))
    x = np.zeros(max(n.value, 1), np.float32)
    _check(L.hbj_vec_bin_read(str(path).encode(), x.ctypes.data, x.size, ctypes.byref(n)))
    return x[: n.value]


def compare_x(x_ref, x, top_k=10):
    """cli main.rs:603-662: dict with n, l2_ref, l2_err, rel_l2, rmse, max_abs_err, worst."""
    xr, xx = _f32(x_ref), _f32(x)
    if xr.size != xx.size:
        raise SolverError(1, f"Length mismatch: x_ref len {xr.size} vs x len {xx.size}")
    st = _lib.hbj_compare_stats()
    worst = np.zeros(max(top_k, 1), np.uint32)
    _check(_lib.lib().hbj_compare_x(_nz(xr, np.float32), _nz(xx, np.float32), xr.size, int(top_k),
                                    ctypes.byref(st), worst.ctypes.data))
    d = {k: getattr(st, k) for k, _ in st._fields_ if not k.startswith("reserved")}
    d["worst"] = worst[: st.n_worst].tolist()
    return d
