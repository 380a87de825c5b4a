"""Max-cut SDP construction and rounding in FP64 NumPy (oracle code).

Restates utils/maxcut_helpers.py:26-92 and 292-302.  Omega is injected by
the caller (the reference draws it from jax.random with PRNGKey(0)).
"""
import numpy as np
import scipy.sparse

from .common import SDPState, scale_sdp_state, reconstruct_from_sketch


def get_all_problem_data(n):
    """diag(X) = 1 as triples (i, i, i) (utils/maxcut_helpers.py:26-33)."""
    r = np.arange(n)[:, None]
    A_indices = np.hstack(3 * [r]).astype(np.int64)
    return np.ones(n), A_indices, np.ones(n), np.zeros(n)


def laplacian_cost(W):
    """C = 0.25 * sym(diag(W 1) - W) (utils/maxcut_helpers.py:38-42)."""
    n = W.shape[0]
    W = scipy.sparse.csr_matrix(W, dtype=np.float64)
    C = scipy.sparse.diags(np.asarray(W @ np.ones(n)).ravel(), 0, shape=(n, n)) - W
    C = 0.5 * (C + C.T)
    return (0.25 * C).tocsr()


def initialize_state(W, sketch_dim, Omega=None):
    """Cold-start state (utils/maxcut_helpers.py:36-92); SCALE_* = 1."""
    C = laplacian_cost(W)
    n = C.shape[0]
    A_data, A_indices, b, mask = get_all_problem_data(n)
    m = n
    if sketch_dim == -1:
        X, P = np.zeros((n, n)), None
        Omega = None
    elif sketch_dim > 0:
        X = None
        assert Omega is not None and Omega.shape == (n, sketch_dim)
        P = np.zeros_like(Omega)
    else:
        raise ValueError("Invalid value for sketch_dim")
    s = SDPState(C=C, A_indices=A_indices, A_data=A_data, b=b, b_ineq_mask=mask, X=X, P=P,
                 Omega=Omega, y=np.zeros(m), z=np.zeros(m), tr_X=0.0, primal_obj=0.0,
                 SCALE_C=1.0, SCALE_X=1.0, SCALE_A=np.ones(m))
    return scale_sdp_state(s)


def sign_vectors(P, Omega):
    """sgn = 2 (E > 0) - 1 for the Nystrom factor E (utils/maxcut_helpers.py:299-300)."""
    E, _ = reconstruct_from_sketch(Omega, P)
    return 2 * (E > 0).astype(np.float64) - 1


def compute_max_cut(P, Omega, callback_static_args, callback_nonstatic_args):
    """max_r diag(sgn^T C sgn) (utils/maxcut_helpers.py:292-302)."""
    C = callback_nonstatic_args
    sgn = sign_vectors(P, Omega)
    return float(np.max(np.sum(sgn * (C @ sgn), axis=0)))


def cut_values_bruteforce(W):
    """Every cut value of a small weighted graph (test pin for rounding)."""
    W = np.asarray(scipy.sparse.csr_matrix(W).todense())
    n = W.shape[0]
    out = {}
    for mask in range(1 << n):
        s = np.array([1.0 if (mask >> i) & 1 else -1.0 for i in range(n)])
        iu, ju = np.triu_indices(n, 1)
        out[mask] = float(np.sum(W[iu, ju] * (s[iu] != s[ju])))
    return out
