"""chan_mlp_forward_noalloc (include/channel_closure.h): inference that
allocates nothing, staging its BF16 inputs in the library's static device
array.  Checked through the C-ABI against chan_mlp_forward, which runs the
same kernels on buffers it allocates: the outputs must be bitwise equal, on
two calls in a row, the second with fewer rows (the array keeps the first
call's contents past the second call's rows)."""
import ctypes

import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _model(n, seed):
    from channel_amd.closure import Mlp
    rng = np.random.default_rng(seed)
    p = (mo.init_params(seed) + 0.05 * rng.standard_normal(mo.NPARAM)).astype(np.float32)
    m = Mlp(n)
    m.set_params(p)
    return m, rng


class _Hip:
    """Device buffers for the test through the HIP runtime that the library
    itself links (torch's wheel carries a runtime of its own)."""

    def __init__(self):
        from channel_amd import lib
        lib.load()
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def put(self, dst, a):  # host -> device
        assert self.rt.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0

    def get(self, a, src):  # device -> host
        assert self.rt.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0

    def free(self, p):
        assert self.rt.hipFree(p) == 0


def _noalloc(m, X):
    from channel_amd import lib
    hip = _Hip()
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.full((X.shape[0], mo.OUT), np.nan, dtype=np.float32)
    dX, dY = hip.alloc(X.nbytes), hip.alloc(Y.nbytes)
    try:
        hip.put(dX, X)
        hip.put(dY, Y)
        lib.check(lib.load().chan_mlp_forward_noalloc(m.h, dX, X.shape[0], dY))
        hip.get(Y, dY)
    finally:
        hip.free(dX)
        hip.free(dY)
    return Y


def test_noalloc_forward_equals_forward_on_two_calls():
    from channel_amd import lib
    n1, n2 = 50000, 1237
    print("[case] noalloc forward", n1, n2, flush=True)
    m, rng = _model(n1, 11)
    for n in (n1, n2):
        X = rng.standard_normal((n, mo.IN)).astype(np.float32)
        want = m.forward(X)
        got = _noalloc(m, X)
        assert np.all(np.isfinite(got))
        np.testing.assert_array_equal(got, want)
    assert lib.CHAN_MLP_NOALLOC_MAX == 65536
    m.close()


def test_noalloc_forward_rejects_oversize_calls():
    from channel_amd import lib
    print("[case] noalloc oversize", flush=True)
    m, _ = _model(16, 12)
    n = lib.CHAN_MLP_NOALLOC_MAX + 1
    hip = _Hip()
    buf = hip.alloc(64)  # never read: the size check comes first
    rc = lib.load().chan_mlp_forward_noalloc(m.h, buf, n, buf)
    hip.free(buf)
    assert rc == lib.CHAN_E_ARG
    assert b"exceed" in lib.load().chan_last_error()
    m.close()
