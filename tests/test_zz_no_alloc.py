"""usbs_sketch_update_noalloc (device arrays in, device arrays out, no
allocation) against usbs_sketch_update (host arrays, same kernels) through
the C ABI: two calls in a row on one context, the second with smaller sizes,
in FP64 and in FP32 storage.  The two entry points run the same kernels over
the same row tiles, so their results must agree bit for bit."""
import ctypes

import numpy as np
import pytest

from usbs_amd import _lib

pytestmark = pytest.mark.gpu

_lib.EXTRA_SIGNATURES.update({
    "usbs_sketch_update": [_lib._vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                           _lib._dp, _lib._dp, _lib._dp, ctypes.c_double, ctypes.c_int32, _lib._dp,
                           _lib._dp],
    "usbs_sketch_update_noalloc": [_lib._vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                   ctypes.c_int32, ctypes.c_int64, _lib._vp, _lib._vp, _lib._vp,
                                   ctypes.c_double, ctypes.c_int32, _lib._vp, _lib._vp],
})


def _host_call(ctx, F, Om, Pb, eta, kc, prec):
    n, k = F.shape
    r = 0 if Om is None else Om.shape[1]
    P = np.zeros_like(Pb)
    Pb2 = np.zeros_like(Pb)
    ctx.check(_lib.lib().usbs_sketch_update(ctx.h, n, k, kc, r, _lib.dptr(F),
                                            None if Om is None else _lib.dptr(Om), _lib.dptr(Pb),
                                            eta, prec, _lib.dptr(P), _lib.dptr(Pb2)))
    return P, Pb2


class _Hip:
    """The HIP runtime the library is linked against (the caller's side of the
    no-allocation call: the caller owns the device arrays).  torch carries a
    HIP runtime of its own, whose device memory this process's library
    runtime does not share, so the test allocates through this one."""

    def __init__(self):
        import ctypes.util  # noqa: F401
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p.value

    def put(self, dst, arr):
        assert self.h.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1) == 0   # host to device

    def get(self, src, arr):
        assert self.h.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2) == 0   # device to host

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _noalloc_call(ctx, F, Om, Pb, eta, kc, prec, ld):
    dt = np.float64 if prec == 0 else np.float32
    n, k = F.shape
    r = 0 if Om is None else Om.shape[1]
    hip = _Hip()

    def colmajor(a):  # n x p row-major host -> p x ld host block (column-major with ld)
        a = a.reshape(n, -1)
        out = np.zeros((a.shape[1], ld), dtype=dt)
        out[:, :n] = a.T.astype(dt)
        return out

    def dev(a):
        h = colmajor(a)
        d = hip.alloc(h.nbytes)
        hip.put(d, h)
        return d, h

    try:
        dF, _ = dev(F)
        dPb, pb_host = dev(Pb)
        dO = dev(Om)[0] if r > 0 else None
        nanblock = np.full_like(pb_host, np.nan)
        dP, dPb2 = dev_nan = (hip.alloc(nanblock.nbytes), hip.alloc(nanblock.nbytes))
        for d in dev_nan:
            hip.put(d, nanblock)
        assert hip.h.hipDeviceSynchronize() == 0
        ctx.check(_lib.lib().usbs_sketch_update_noalloc(ctx.h, n, k, kc, r, ld, dF, dO, dPb, eta,
                                                        prec, dP, dPb2))
        outs = []
        for d in (dP, dPb2, dPb):
            h = np.empty_like(pb_host)
            hip.get(d, h)
            outs.append(h)
    finally:
        hip.free()
    assert np.array_equal(outs[2], pb_host)  # the input sketch is left unchanged
    return tuple(o[:, :n].astype(np.float64).T.reshape(Pb.shape) for o in outs[:2])


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_sketch_update_equals_host_entry_point(gpu_ctx, prec):
    rng = np.random.default_rng(31 + prec)
    cases = [(70001, 11, 10, 10, 70016 + 64), (3001, 5, 3, 4, 3072)]
    for n, k, kc, r, ld in cases:
        print("case: noalloc sketch update n=%d k=%d kc=%d r=%d ld=%d prec=%d" % (n, k, kc, r, ld, prec),
              flush=True)
        F, Om, Pb = (rng.standard_normal((n, k)), rng.standard_normal((n, r)),
                     rng.standard_normal((n, r)))
        eta = float(rng.random())
        P_h, Pb_h = _host_call(gpu_ctx, F, Om, Pb, eta, kc, prec)
        P_d, Pb_d = _noalloc_call(gpu_ctx, F, Om, Pb, eta, kc, prec, ld)
        assert np.array_equal(P_d, P_h) and np.array_equal(Pb_d, Pb_h), (n, prec)


def test_noalloc_diag_only_and_size_limit(gpu_ctx):
    rng = np.random.default_rng(5)
    n, k, kc = 4097, 6, 4
    F, Pb = rng.standard_normal((n, k)), rng.standard_normal(n)
    print("case: noalloc diag-only n=%d" % n, flush=True)
    P_h, Pb_h = _host_call(gpu_ctx, F, None, Pb, 0.25, kc, 0)
    P_d, Pb_d = _noalloc_call(gpu_ctx, F, None, Pb, 0.25, kc, 0, 4160)
    assert np.array_equal(P_d, P_h) and np.array_equal(Pb_d, Pb_h)
    # tile partials beyond the fixed array (2^18 doubles) are refused before any
    # launch: n = 2^24 has 512 row tiles of 32768 rows, k = 16, r = 48
    big_n, tile = 1 << 24, 512
    while (big_n + tile - 1) // tile > 512:
        tile *= 2
    assert 1024 + ((big_n + tile - 1) // tile) * 16 * 48 > (1 << 18)
    with pytest.raises(_lib.USBSError) as e:
        gpu_ctx.check(_lib.lib().usbs_sketch_update_noalloc(
            gpu_ctx.h, 1 << 24, 16, 16, 48, 1 << 24, 1, 1, 1, 0.5, 0, 1, 1))
    assert e.value.code == -1
