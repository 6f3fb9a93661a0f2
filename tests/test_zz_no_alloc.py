"""hbj_op_dot_noalloc: the dot product that allocates nothing (device inputs,
partials in the library's fixed device array) against hbj_op_dot, bit for
bit, on calls in a row with shrinking inputs, the largest at the limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """Device memory through the HIP runtime libhbj.so itself links (a second
    runtime in the process, such as the one bundled with torch, cannot open
    the device once this one has)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [ctypes.c_void_p]

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(nbytes, 4)) == 0
        return p

    def h2d(self, dst, arr):
        if arr.size:
            assert self.L.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1) == 0  # hipMemcpyHostToDevice

    def d2h(self, arr, src):
        assert self.L.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2) == 0  # hipMemcpyDeviceToHost


def _noalloc(hbj, ctx, hip, a, b, stream):
    L = hbj._lib.lib()
    da, db, dout = hip.malloc(a.nbytes), hip.malloc(b.nbytes), hip.malloc(4)
    try:
        hip.h2d(da, a)
        hip.h2d(db, b)
        hip.h2d(dout, np.full(1, np.nan, np.float32))
        st = L.hbj_op_dot_noalloc(ctx._h, a.size, da if a.size else None, db if b.size else None,
                                  dout, stream)
        assert st == 0, hbj._lib.last_error()
        # (the context's stream does not synchronise with the null stream)
        assert hip.L.hipDeviceSynchronize() == 0
        out = np.zeros(1, np.float32)
        hip.d2h(out, dout)
        return out[0]
    finally:
        for p in (da, db, dout):
            hip.L.hipFree(p)


def _reference(hbj, ctx, a, b):
    L = hbj._lib.lib()
    ref = ctypes.c_float(float("nan"))
    a1 = a if a.size else np.zeros(1, np.float32)
    b1 = b if b.size else np.zeros(1, np.float32)
    assert L.hbj_op_dot(ctx._h, a.size, a1, b1, ctypes.byref(ref)) == 0
    return np.float32(ref.value)


def test_dot_noalloc_matches_op_dot(ctx, hbj, oracle):
    hip = _Hip()
    own = ctypes.c_void_p()
    assert hip.L.hipStreamCreate(ctypes.byref(own)) == 0
    limit = 1 << 26
    rng = np.random.default_rng(17)
    # the limit (two reduction levels through the scratch), then smaller inputs;
    # on a caller's stream and on the context's own (stream = NULL)
    for n, use_own in ((limit, True), (5_000_003, False), (70_001, True), (1, False), (0, True)):
        print(f"dot noalloc n={n}", flush=True)
        a = rng.uniform(-1, 1, n).astype(np.float32)
        b = rng.uniform(-1, 1, n).astype(np.float32)
        got = np.float32(_noalloc(hbj, ctx, hip, a, b, own if use_own else None))
        ref = _reference(hbj, ctx, a, b)
        assert got.view(np.uint32) == ref.view(np.uint32), (n, got, ref)
        if 0 < n <= 5_000_003:
            assert got.view(np.uint32) == np.float32(oracle.dot(a, b)).view(np.uint32)
    hip.L.hipStreamDestroy(own)


def test_dot_noalloc_rejects_above_limit(ctx, hbj):
    L = hbj._lib.lib()
    dummy = ctypes.c_void_p(1)  # never dereferenced: the size check comes first
    st = L.hbj_op_dot_noalloc(ctx._h, (1 << 26) + 1, dummy, dummy, dummy, None)
    assert st == 1
    assert "exceeds the limit" in hbj._lib.last_error()
    assert L.hbj_op_dot_noalloc(ctx._h, 4, None, None, dummy, None) == 1
