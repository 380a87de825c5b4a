"""Nystrom reconstruction and max-cut rounding on the GPU.

Mirror of utils/common.py:99-116 (reconstruct_from_sketch) and
utils/maxcut_helpers.py:292-302 (compute_max_cut) over the C ABI.
"""
import ctypes

import numpy as np

from . import _lib
from .sparse import coo_parts

_lib.EXTRA_SIGNATURES.update({
    "usbs_reconstruct_from_sketch": [_lib._vp, ctypes.c_int64, ctypes.c_int32, _lib._dp, _lib._dp,
                                     _lib._dp, _lib._dp],
    "usbs_maxcut_round": [_lib._vp, ctypes.c_int32, _lib._dp, _lib._dp, _lib._dp,
                          ctypes.POINTER(ctypes.c_int8)],
})


def reconstruct_from_sketch(Omega, P, approx_eps=1e-6, context=None, device=0):
    """(E, Lambda) of the fixed-rank Nystrom approximation of X from its
    sketch P = X Omega (Tropp, Yurtsever, Udell, Cevher 2017)."""
    if approx_eps != 1e-6:
        raise ValueError("only the reference's approx_eps=1e-6 is implemented")
    Omega, P = _lib.as_f64(Omega), _lib.as_f64(P)
    n, r = Omega.shape
    ctx = context if context is not None else _lib.Context(device)
    E = np.zeros((n, r))
    lam = np.zeros(r)
    ctx.check(_lib.lib().usbs_reconstruct_from_sketch(ctx.h, n, r, _lib.dptr(Omega), _lib.dptr(P),
                                                      _lib.dptr(E), _lib.dptr(lam)))
    return E, lam


def maxcut_round(C, P, Omega, context=None, device=0):
    """Rounding of the sketch against cost C: returns (best value, the r
    candidate values, sign vectors n x r int8)."""
    P, Omega = _lib.as_f64(P), _lib.as_f64(Omega)
    n, r = P.shape
    ctx = context if context is not None else _lib.Context(device)
    if context is None or ctx.n != n:
        rows, cols, vals, shape = coo_parts(C)
        ctx.set_problem(n, n, rows, cols, vals, np.stack([np.arange(n)] * 3, 1), np.ones(n),
                        np.ones(n), np.zeros(n), np.ones(n))
    values = np.zeros(r)
    signs = np.zeros((n, r), dtype=np.int8)
    ctx.check(_lib.lib().usbs_maxcut_round(ctx.h, r, _lib.dptr(P), _lib.dptr(Omega), _lib.dptr(values),
                                           signs.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))))
    return float(values.max()), values, signs
