"""The constraint operator A(.) and its adjoint with the reference's
signatures (solver/utils.py:11-37), computed by the HIP library.

A constraint set is a list of triples (c, i, j) with values A_data:
constraint c reads entry (i, j) of X.  ``apply_A_operator_slim(m, ...,
u)`` is the constraint map of u u^T, ``apply_A_operator_batched`` sums it
over the columns of ``vecs``; ``apply_A_adjoint_slim(n, ..., z, u)`` is
A*(z) u and ``apply_A_adjoint_batched`` applies A*(z) to every column.
The trailing keyword arguments pick the device and an optional context.
"""
import ctypes

import numpy as np

from . import _lib

_lib.EXTRA_SIGNATURES.update({
    "usbs_apply_A_operator": [_lib._vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _lib._ip,
                              _lib._dp, ctypes.c_int32, _lib._dp, _lib._dp],
    "usbs_apply_A_adjoint": [_lib._vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _lib._ip,
                             _lib._dp, _lib._dp, ctypes.c_int32, _lib._dp, _lib._dp],
})


def _triples(A_data, A_indices):
    A_indices = _lib.as_i32(np.asarray(A_indices).reshape(-1, 3))
    A_data = _lib.as_f64(A_data).ravel()
    if A_data.size != A_indices.shape[0]:
        raise ValueError("A_data has %d entries, A_indices %d rows" % (A_data.size, A_indices.shape[0]))
    return A_data, A_indices


def _operator(m, A_data, A_indices, U, context, device):
    A_data, A_indices = _triples(A_data, A_indices)
    U = _lib.as_f64(U)
    n, p = U.shape
    ctx = context if context is not None else _lib.Context(device)
    out = np.zeros(int(m))
    ctx.check(_lib.lib().usbs_apply_A_operator(ctx.h, n, int(m), A_data.size, _lib.iptr(A_indices),
                                               _lib.dptr(A_data), p, _lib.dptr(U), _lib.dptr(out)))
    return out


def _adjoint(n, A_data, A_indices, z, U, context, device):
    A_data, A_indices = _triples(A_data, A_indices)
    U = _lib.as_f64(U)
    if U.shape[0] != n:
        raise ValueError("vectors have %d rows, expected n = %d" % (U.shape[0], n))
    z = _lib.as_f64(z).ravel()
    p = U.shape[1]
    ctx = context if context is not None else _lib.Context(device)
    out = np.zeros((int(n), p))
    ctx.check(_lib.lib().usbs_apply_A_adjoint(ctx.h, int(n), z.size, A_data.size, _lib.iptr(A_indices),
                                              _lib.dptr(A_data), _lib.dptr(z), p, _lib.dptr(U),
                                              _lib.dptr(out)))
    return out


def apply_A_operator_slim(m, A_data, A_indices, u, context=None, device=0):
    """z[c] = sum of A_data * u[i] * u[j] over constraint c's triples (solver/utils.py:11-16)."""
    return _operator(m, A_data, A_indices, np.asarray(u, dtype=np.float64).reshape(-1, 1), context,
                     device)


def apply_A_adjoint_slim(n, A_data, A_indices, z, u, context=None, device=0):
    """out[i] = sum of A_data * z[c] * u[j] over the triples of row i (solver/utils.py:19-24)."""
    return _adjoint(n, A_data, A_indices, z, np.asarray(u, dtype=np.float64).reshape(-1, 1),
                    context, device)[:, 0]


def apply_A_operator_batched(m, A_data, A_indices, vecs, context=None, device=0):
    """apply_A_operator_slim summed over the columns of vecs (solver/utils.py:27-31)."""
    return _operator(m, A_data, A_indices, vecs, context, device)


def apply_A_adjoint_batched(n, A_data, A_indices, z, vecs, context=None, device=0):
    """apply_A_adjoint_slim on every column of vecs (solver/utils.py:34-37)."""
    return _adjoint(n, A_data, A_indices, z, vecs, context, device)
