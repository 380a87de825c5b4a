"""eigsh_smallest with the reference's signature, computed by the HIP library.

Mirrors solver/lanczos.py:239-284: the num_desired smallest eigenpairs of
the operator C + A*(adjoint_left_vec) by thick-restart Lanczos with full
CGS2 reorthogonalisation.  The trailing keyword arguments select the backend
(only "hip" exists here) and the execution form.
"""
import ctypes

import numpy as np

from . import _lib
from .sparse import coo_parts


def _check_backend(backend):
    if backend != "hip":
        raise ValueError("backend=%r is not available in usbs_amd (only 'hip'); the JAX path "
                         "lives in the reference" % (backend,))


def eigsh_smallest(n, C, A_data, A_indices, adjoint_left_vec, q0, num_desired, inner_iterations,
                   max_restarts, tolerance, backend="hip", precision="fp64", exec_mode="perop",
                   device=0, return_info=False, context=None):
    """Find the `num_desired` smallest eigenvalues of C + A*(adjoint_left_vec).

    Returns (eigenvalues (num_desired,), eigenvectors (n, num_desired)); with
    return_info also a dict {restarts, num_converged, steps, breakdowns}.
    precision: "fp64" or "mixed" (FP32 basis and operator values, FP64
    accumulation and FP64 Ritz solve).  exec_mode: "perop" or "persistent".
    """
    _check_backend(backend)
    rows, cols, vals, shape = coo_parts(C)
    if shape != (n, n):
        raise ValueError("C has shape %s, expected (%d, %d)" % (shape, n, n))
    A_indices = np.asarray(A_indices).reshape(-1, 3)
    z = _lib.as_f64(adjoint_left_vec)
    m = z.size
    ctx = context if context is not None else _lib.Context(device)
    ctx.set_problem(n, m, rows, cols, vals, A_indices, A_data, np.zeros(m), np.zeros(m), np.ones(m))
    q0 = _lib.as_f64(q0)
    if q0.size != n:
        raise ValueError("q0 must have n entries")
    vals_out = np.zeros(num_desired)
    vecs_out = np.zeros((n, num_desired))
    info = np.zeros(4, dtype=np.int32)
    ctx.check(_lib.lib().usbs_lanczos_smallest(
        ctx.h, 1.0, _lib.dptr(z), _lib.dptr(q0), int(num_desired), int(inner_iterations),
        int(max_restarts), float(tolerance), _lib.PRECISION[precision], _lib.EXEC_MODE[exec_mode],
        _lib.dptr(vals_out), _lib.dptr(vecs_out), _lib.iptr(info)))
    if return_info:
        return vals_out, vecs_out, dict(restarts=int(info[0]), num_converged=int(info[1]),
                                        steps=int(info[2]), breakdowns=int(info[3]))
    return vals_out, vecs_out


_PP = ctypes.POINTER(_lib._dp)
_lib.EXTRA_SIGNATURES["usbs_lanczos_smallest_batch"] = [
    ctypes.POINTER(_lib._vp), ctypes.c_int, ctypes.c_double, _PP, _PP, ctypes.c_int, ctypes.c_int,
    ctypes.c_int, ctypes.c_double, ctypes.c_int, _PP, _PP, _lib._ip]


def eigsh_smallest_batch(problems, num_desired, inner_iterations, max_restarts, tolerance,
                         precision="fp64", contexts=None, device=0, return_info=False):
    """Batched eigsh_smallest: ``problems`` is a list of (n, C, A_data,
    A_indices, adjoint_left_vec, q0) tuples of possibly different sizes; all
    instances advance together in one persistent launch per restart.  Each
    result is bitwise the one eigsh_smallest(..., exec_mode="persistent")
    returns for that instance alone.  Returns a list of (eigenvalues,
    eigenvectors) pairs (and a list of info dicts with return_info)."""
    k = len(problems)
    ctxs = contexts if contexts is not None else [_lib.Context(device) for _ in range(k)]
    zs, qs, outs_v, outs_e = [], [], [], []
    for (n, C, A_data, A_indices, z, q0), ctx in zip(problems, ctxs):
        rows, cols, vals, shape = coo_parts(C)
        if shape != (n, n):
            raise ValueError("C has shape %s, expected (%d, %d)" % (shape, n, n))
        z = _lib.as_f64(z)
        m = z.size
        ctx.set_problem(n, m, rows, cols, vals, np.asarray(A_indices).reshape(-1, 3), A_data,
                        np.zeros(m), np.zeros(m), np.ones(m))
        q0 = _lib.as_f64(q0)
        if q0.size != n:
            raise ValueError("q0 must have n entries")
        zs.append(z)
        qs.append(q0)
        outs_v.append(np.zeros(num_desired))
        outs_e.append(np.zeros((n, num_desired)))
    info = np.zeros(4 * k, dtype=np.int32)
    handles = (_lib._vp * k)(*[c.h for c in ctxs])
    arr = lambda xs: (_lib._dp * k)(*[_lib.dptr(x) for x in xs])  # noqa: E731
    ctxs[0].check(_lib.lib().usbs_lanczos_smallest_batch(
        handles, k, 1.0, arr(zs), arr(qs), int(num_desired), int(inner_iterations),
        int(max_restarts), float(tolerance), _lib.PRECISION[precision], arr(outs_v), arr(outs_e),
        _lib.iptr(info)))
    res = list(zip(outs_v, outs_e))
    if return_info:
        infos = [dict(restarts=int(info[4 * i]), num_converged=int(info[4 * i + 1]),
                      steps=int(info[4 * i + 2]), breakdowns=int(info[4 * i + 3])) for i in range(k)]
        return res, infos
    return res
