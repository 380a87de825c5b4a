"""SDPState, scaling and the Nystrom reconstruction in FP64 NumPy (oracle code).

Restates utils/common.py:11-121.  ``C`` is a scipy.sparse matrix here where
the reference holds a BCOO; only ``C @ x``, ``C.data`` and the pattern are used.
"""
from collections import namedtuple

import numpy as np
import scipy.sparse

SDPState = namedtuple("SDPState", ["C", "A_indices", "A_data", "b", "b_ineq_mask", "X", "P",
                                   "Omega", "y", "z", "tr_X", "primal_obj", "SCALE_C", "SCALE_X",
                                   "SCALE_A"])


def scale_sdp_state(s):
    """utils/common.py:29-61."""
    A_indices = np.asarray(s.A_indices)
    return s._replace(
        C=(s.C * s.SCALE_C).tocsr() if scipy.sparse.issparse(s.C) else s.C * s.SCALE_C,
        A_data=s.A_data * np.asarray(s.SCALE_A)[A_indices[:, 0]],
        b=s.b * s.SCALE_X * s.SCALE_A,
        X=None if s.X is None else s.X * s.SCALE_X,
        P=None if s.P is None else s.P * s.SCALE_X,
        z=s.z * s.SCALE_A * s.SCALE_X,
        tr_X=s.tr_X * s.SCALE_X,
        primal_obj=s.primal_obj * s.SCALE_X * s.SCALE_C)


def unscale_sdp_state(s):
    """utils/common.py:64-96."""
    A_indices = np.asarray(s.A_indices)
    return s._replace(
        C=(s.C / s.SCALE_C).tocsr() if scipy.sparse.issparse(s.C) else s.C / s.SCALE_C,
        A_data=s.A_data / np.asarray(s.SCALE_A)[A_indices[:, 0]],
        b=(s.b / s.SCALE_A) / s.SCALE_X,
        X=None if s.X is None else s.X / s.SCALE_X,
        P=None if s.P is None else s.P / s.SCALE_X,
        z=(s.z / s.SCALE_A) / s.SCALE_X,
        tr_X=s.tr_X / s.SCALE_X,
        primal_obj=s.primal_obj / (s.SCALE_X * s.SCALE_C))


def reconstruct_from_sketch(Omega, P, approx_eps=1e-6):
    """Fixed-rank Nystrom reconstruction (utils/common.py:99-116; Tropp,
    Yurtsever, Udell, Cevher 2017).  A non-positive-definite core B gives NaN
    output, as XLA's Cholesky does."""
    n = Omega.shape[0]
    rho = np.sqrt(n) * approx_eps * np.linalg.norm(P, ord=2)
    P_rho = P + rho * Omega
    B = Omega.T @ P_rho
    B = 0.5 * (B + B.T)
    try:
        L = np.linalg.cholesky(B)
    except np.linalg.LinAlgError:
        r = Omega.shape[1]
        return np.full((n, r), np.nan), np.full((r,), np.nan)
    M = np.linalg.lstsq(L, P_rho.T, rcond=None)[0].T
    E, Rho, _ = np.linalg.svd(M, full_matrices=False)
    Lambda = np.clip(Rho ** 2 - rho, 0, np.inf)
    return E, Lambda


def apply_A_operator_mx(n, m, A_data, A_indices, X):
    """<A_c, X> for a dense X (utils/common.py:119-121)."""
    A_indices = np.asarray(A_indices)
    return np.bincount(A_indices[:, 0], weights=A_data * X[A_indices[:, 1], A_indices[:, 2]],
                       minlength=m)
