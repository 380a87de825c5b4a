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


def _laplacian_cost_and_data(W):
    C = laplacian_cost(W)
    n = C.shape[0]
    A_data, A_indices, b, mask = get_all_problem_data(n)
    return C, n, A_data, A_indices, b, mask


def _diag(C):
    return np.asarray(C.diagonal(), dtype=np.float64)


def _warm_scales(C, n):
    """SCALE_X = 1/n, SCALE_C = 1/|C.data| (utils/maxcut_helpers.py:134-136)."""
    return 1.0 / float(n), 1.0 / np.linalg.norm(C.data)


def implicit_warm_start_state(old, W, sketch_dim, Omega_pad):
    """utils/maxcut_helpers.py:95-161.  The new graph W has n >= n_old
    vertices; P and Omega are padded (new rows of Omega: Omega_pad, drawn by
    the reference from jax.random with PRNGKey(n)); y, z padded with zeros;
    the primal objective corrected for the change of C's diagonal; the state
    rescaled by SCALE_X = 1/n, SCALE_C = 1/|C|_F."""
    from .common import unscale_sdp_state
    old = unscale_sdp_state(old)
    C, n, A_data, A_indices, b, mask = _laplacian_cost_and_data(W)
    n_old = old.C.shape[0]
    X, P, Omega = old.X, old.P, old.Omega
    if old.X is not None:
        X = np.zeros((n, n))
        X[:n_old, :n_old] = old.X
    if old.P is not None:
        P = np.zeros((n, sketch_dim))
        P[:n_old] = old.P
        Omega = np.concatenate([old.Omega, Omega_pad], axis=0)
    y = np.zeros(n)
    y[:old.b.shape[0]] = old.y
    z = np.zeros(n)
    z[:old.b.shape[0]] = old.z
    primal_obj = old.primal_obj - _diag(old.C) @ old.z + _diag(C) @ z
    SX, SC = _warm_scales(C, n)
    s = SDPState(C=C, A_indices=A_indices, A_data=A_data, b=b, b_ineq_mask=mask, X=X, P=P,
                 Omega=Omega, y=y, z=z, tr_X=old.tr_X, primal_obj=primal_obj, SCALE_C=SC,
                 SCALE_X=SX, SCALE_A=np.ones(n))
    return scale_sdp_state(s)


def explicit_warm_start_state(old, W, sketch_dim, Omega_new):
    """utils/maxcut_helpers.py:164-229.  The old sketch is reconstructed
    (Nystrom, utils/common.py:99-116), its eigenvalues shifted so that their
    sum is the old trace, the factor padded to the new size, and P, z, tr_X
    and the primal objective recomputed from it with a fresh test matrix
    (Omega_new, PRNGKey(n) in the reference)."""
    from .common import unscale_sdp_state
    from .operators import apply_A_operator_batched
    old = unscale_sdp_state(old)
    C, n, A_data, A_indices, b, mask = _laplacian_cost_and_data(W)
    n_old = old.C.shape[0]
    X, P, Omega = old.X, old.P, old.Omega
    if old.X is not None:
        X = np.zeros((n, n))
        X[:n_old, :n_old] = old.X
        z = np.bincount(A_indices[:, 0], weights=A_data * X[A_indices[:, 1], A_indices[:, 2]],
                        minlength=n)
        tr_X = np.trace(X)
        primal_obj = float(np.sum(C.multiply(X)))
    if old.P is not None:
        Omega = Omega_new
        E, Lam = reconstruct_from_sketch(old.Omega, old.P)
        tr_offset = (old.tr_X - np.sum(Lam)) / Lam.shape[0]
        Lam = Lam + tr_offset
        Ep = np.zeros((n, E.shape[1]))
        Ep[:n_old] = E
        S = Ep * np.sqrt(Lam)[None, :]
        P = S @ (S.T @ Omega)
        z = apply_A_operator_batched(n, A_data, A_indices, S)
        tr_X = np.sum(Lam)
        primal_obj = float(np.trace(S.T @ (C @ S)))
    y = np.zeros(n)
    y[:old.b.shape[0]] = old.y
    SX, SC = _warm_scales(C, n)
    s = SDPState(C=C, A_indices=A_indices, A_data=A_data, b=b, b_ineq_mask=mask, X=X, P=P,
                 Omega=Omega, y=y, z=z, tr_X=tr_X, primal_obj=primal_obj, SCALE_C=SC, SCALE_X=SX,
                 SCALE_A=np.ones(n))
    return scale_sdp_state(s)


def dual_only_warm_start_state(old, W, sketch_dim, Omega_new):
    """utils/maxcut_helpers.py:232-289: only y carries over (padded); the
    primal starts at zero with a fresh test matrix (PRNGKey(0) in the
    reference), rescaled as the other warm starts."""
    from .common import unscale_sdp_state
    old = unscale_sdp_state(old)
    C, n, A_data, A_indices, b, mask = _laplacian_cost_and_data(W)
    if sketch_dim == -1:
        X, Omega, P = np.zeros((n, n)), None, None
    elif sketch_dim > 0:
        X, Omega, P = None, Omega_new, np.zeros_like(Omega_new)
    else:
        raise ValueError("Invalid value for sketch_dim")
    y = np.zeros(n)
    y[:old.b.shape[0]] = old.y
    SX, SC = _warm_scales(C, n)
    s = SDPState(C=C, A_indices=A_indices, A_data=A_data, b=b, b_ineq_mask=mask, X=X, P=P,
                 Omega=Omega, y=y, z=np.zeros(n), tr_X=0.0, primal_obj=0.0, SCALE_C=SC,
                 SCALE_X=SX, SCALE_A=np.ones(n))
    return scale_sdp_state(s)
