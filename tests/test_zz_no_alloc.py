"""The allocation-free solve (qp_solve2d_device_noalloc) against the allocating
one (qp_solve2d_device) on the same device buffers, through the C-ABI: two
calls in a row on one context, the second on a smaller grid of the other
kernel family (so it runs on what the first call left in the shared scratch
array), then a third back at the first size; and its argument errors."""
import ctypes

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DevBuf:
    """n doubles of device memory owned by the test (the entry point under test allocates nothing)."""

    def __init__(self, n):
        self.n = n
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(8 * n)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        assert a.size == self.n
        assert hip().hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(8 * self.n), 1) == 0

    def get(self, shape):
        out = np.empty(self.n)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(8 * self.n), 2) == 0
        return out.reshape(shape)

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = ctypes.c_void_p()


def alloc_solve(qp_ctx, src, d1, d2, D):
    from quetzalpic_amd import capi
    n1, n2 = src.shape
    f = capi.Fourier2d(qp_ctx, n1, n2)
    a, b = DevBuf(n1 * n2), DevBuf(n1 * n2)
    a.put(src)
    capi._check(capi.load().qp_solve2d_device(f.h, d1, d2, D, a.p, b.p), "qp_solve2d_device")
    qp_ctx.synchronize()
    out = b.get((n1, n2))
    a.free()
    b.free()
    f.close()
    return out


def noalloc_solve(qp_ctx, src, d1, d2, D, inplace=False):
    from quetzalpic_amd import capi
    n1, n2 = src.shape
    a = DevBuf(n1 * n2)
    b = a if inplace else DevBuf(n1 * n2)
    a.put(src)
    capi._check(capi.load().qp_solve2d_device_noalloc(qp_ctx.h, n1, n2, d1, d2, D, a.p, b.p),
                "qp_solve2d_device_noalloc")
    qp_ctx.synchronize()
    out = b.get((n1, n2))
    a.free()
    if not inplace:
        b.free()
    return out


# (n1, n2, D): register-resident FFT kernels (both sides 2^p or 3 * 2^p) and
# the generic ones; each case shrinks after the one before it
CASES = [
    [(256, 256, 0.0), (64, 96, 10.0), (256, 256, 0.0)],
    [(1024, 1024, 10.0), (100, 60, 0.0), (7, 9, 10.0)],
]


@pytest.mark.parametrize("seq", CASES, ids=["wave-then-small", "largest-then-generic"])
def test_noalloc_equals_allocating_solve(qp_ctx, seq):
    for i, (n1, n2, D) in enumerate(seq):
        rng = np.random.default_rng(11 * n1 + n2 + i)
        src = rng.standard_normal((n1, n2))
        d1, d2 = 20.0 / n1, 24.0 / n2
        want = alloc_solve(qp_ctx, src, d1, d2, D)
        got = noalloc_solve(qp_ctx, src, d1, d2, D, inplace=(i == 1))
        err = rel_err(got, want)
        print(f"[noalloc] call {i}: {n1}x{n2} D={D}: max {err[0]:.2e} rms {err[1]:.2e}", flush=True)
        # same kernels; only the twiddles and eigenvalue terms are computed on the
        # device (sincospi) instead of the host (cosl/sinl): rounding-level differences
        assert err[0] < 1e-12 and err[1] < 1e-12, (n1, n2, D, err)


def test_noalloc_rejects_bad_arguments(qp_ctx):
    from quetzalpic_amd import capi
    lib = capi.load()
    a = DevBuf(16)
    for n1, n2 in ((0, 4), (4, 0), (1025, 4), (4, 2048)):
        assert lib.qp_solve2d_device_noalloc(qp_ctx.h, n1, n2, 1.0, 1.0, 0.0, a.p, a.p) == capi.QP_ERR_ARG
        assert b"outside [1, 1024]" in lib.qp_last_error()
    assert lib.qp_solve2d_device_noalloc(qp_ctx.h, 4, 4, 0.0, 1.0, 0.0, a.p, a.p) == capi.QP_ERR_ARG
    assert lib.qp_solve2d_device_noalloc(qp_ctx.h, 4, 4, 1.0, 1.0, 0.0, None, a.p) == capi.QP_ERR_ARG
    a.free()
