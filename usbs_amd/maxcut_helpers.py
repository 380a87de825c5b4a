"""Max-cut SDP construction and the rounding callback, mirroring
utils/maxcut_helpers.py:26-92 and 292-302.

The problem build is host-side setup (like the reference's scipy code); the
rounding is computed by the HIP library (Nystrom reconstruction from the
device-resident sketch, then integer cut sums).  ``compute_max_cut`` passed
as ``callback_fn`` to ``usbs_amd.usbs.usbs`` selects that native rounding.
"""
import numpy as np
import scipy.sparse

from . import problems
from .common import SDPState, scale_sdp_state, unscale_sdp_state
from .sparse import BCOO


def get_all_problem_data(C):
    """diag(X) = 1 as triples (i, i, i) (utils/maxcut_helpers.py:26-33)."""
    n = C.shape[0]
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


def initialize_state(C, sketch_dim, omega_seed=0, Omega=None):
    """Cold start (utils/maxcut_helpers.py:36-92).  ``C`` is the graph's
    weighted adjacency (what the reference loads from a .mat file).  Omega is
    a seeded standard normal n x sketch_dim matrix (the reference draws it
    from jax.random with PRNGKey(0)); pass Omega to inject one."""
    L = laplacian_cost(C)
    n = L.shape[0]
    Cb = BCOO.from_scipy_sparse(L)
    A_data, A_indices, b, b_ineq_mask = get_all_problem_data(L)
    m = b.shape[0]
    if sketch_dim == -1:
        X, P, Omega = np.zeros((n, n)), None, None
    elif sketch_dim > 0:
        X = None
        if Omega is None:
            Omega = problems.gaussian_matrix(n, sketch_dim, omega_seed)
        P = np.zeros_like(Omega)
    else:
        raise ValueError("Invalid value for sketch_dim")
    s = SDPState(C=Cb, A_indices=A_indices, A_data=A_data, b=b, b_ineq_mask=b_ineq_mask, X=X, P=P,
                 Omega=Omega, y=np.zeros(m), z=np.zeros(m), tr_X=0.0, primal_obj=0.0, SCALE_C=1.0,
                 SCALE_X=1.0, SCALE_A=np.ones(m))
    return scale_sdp_state(s)


def compute_max_cut(P, Omega, callback_static_args, callback_nonstatic_args):
    """Rounding callback (utils/maxcut_helpers.py:292-302).  Inside
    ``usbs(..., backend="hip")`` this function is recognised and the rounding
    runs natively on the device-resident sketch; called directly it runs the
    same HIP kernels on the given host arrays."""
    from .rounding import maxcut_round
    value, _, _ = maxcut_round(callback_nonstatic_args, P, Omega)
    return value


# -- warm starts (utils/maxcut_helpers.py:95-289) ---------------------------
# The reference draws its new test-matrix rows from jax.random (PRNGKey(n) or
# PRNGKey(0)), a stream this package cannot reproduce; each transform takes
# them as an argument and defaults to a seeded normal matrix with the same
# seed number.

def _new_problem(C):
    L = laplacian_cost(C)
    n = L.shape[0]
    A_data, A_indices, b, mask = get_all_problem_data(L)
    return L, n, A_data, A_indices, b, mask


def _warm_scales(L, n):
    """SCALE_X = 1/n, SCALE_C = 1/|C|_F over the stored entries
    (utils/maxcut_helpers.py:134-136)."""
    return 1.0 / float(n), 1.0 / float(np.linalg.norm(L.data))


def _padded(v, n):
    out = np.zeros(n)
    v = np.asarray(v, dtype=np.float64)
    out[:v.shape[0]] = v
    return out


def get_implicit_warm_start_state(old_sdp_state, C, sketch_dim, Omega_pad=None):
    """utils/maxcut_helpers.py:95-161: the new graph's adjacency C has
    n >= n_old vertices; P and Omega are padded (new rows of Omega:
    Omega_pad, default seeded normal with seed n), y and z zero-padded, the
    primal objective corrected for the change of C's diagonal, and the state
    rescaled by SCALE_X = 1/n, SCALE_C = 1/|C|_F."""
    if old_sdp_state.X is not None:
        raise NotImplementedError("dense-X storage is not provided by the HIP backend")
    if sketch_dim != np.asarray(old_sdp_state.Omega).shape[1]:
        raise ValueError("sketch_dim must equal the old sketch's")
    old = unscale_sdp_state(old_sdp_state)
    L, n, A_data, A_indices, b, mask = _new_problem(C)
    n_old = old.C.shape[0]
    if Omega_pad is None:
        Omega_pad = problems.gaussian_matrix(n - n_old, sketch_dim, n)
    P = np.zeros((n, sketch_dim))
    P[:n_old] = old.P
    Omega = np.concatenate([np.asarray(old.Omega, dtype=np.float64), Omega_pad], axis=0)
    y, z = _padded(old.y, n), _padded(old.z, n)
    old_diag = np.asarray(_as_scipy(old.C).diagonal(), dtype=np.float64)
    primal_obj = (float(old.primal_obj) - float(old_diag @ np.asarray(old.z))
                  + float(np.asarray(L.diagonal()) @ z))
    SX, SC = _warm_scales(L, n)
    s = SDPState(C=BCOO.from_scipy_sparse(L), A_indices=A_indices, A_data=A_data, b=b,
                 b_ineq_mask=mask, X=None, P=P, Omega=Omega, y=y, z=z, tr_X=float(old.tr_X),
                 primal_obj=primal_obj, SCALE_C=SC, SCALE_X=SX, SCALE_A=np.ones(n))
    return scale_sdp_state(s)


def get_explicit_warm_start_state(old_sdp_state, C, sketch_dim, Omega=None, context=None):
    """utils/maxcut_helpers.py:164-229: the old sketch is reconstructed on the
    device (Nystrom, utils/common.py:99-116), its eigenvalues shifted so that
    they sum to the old trace, the factor padded to the new size, and P, z
    (the constraint map, on the device), tr_X and the primal objective
    recomputed from it with a fresh test matrix Omega (default seeded normal
    with seed n); rescaled as the implicit warm start."""
    from .operators import apply_A_operator_batched
    from .rounding import reconstruct_from_sketch
    if old_sdp_state.X is not None:
        raise NotImplementedError("dense-X storage is not provided by the HIP backend")
    old = unscale_sdp_state(old_sdp_state)
    L, n, A_data, A_indices, b, mask = _new_problem(C)
    n_old = old.C.shape[0]
    if Omega is None:
        Omega = problems.gaussian_matrix(n, sketch_dim, n)
    E, Lam = reconstruct_from_sketch(old.Omega, old.P, context=context)
    Lam = Lam + (float(old.tr_X) - float(np.sum(Lam))) / Lam.shape[0]
    Ep = np.zeros((n, E.shape[1]))
    Ep[:n_old] = E
    S = Ep * np.sqrt(Lam)[None, :]
    P = S @ (S.T @ Omega)
    z = apply_A_operator_batched(n, A_data, A_indices, S, context=context)
    primal_obj = float(np.sum(S * (L @ S)))
    SX, SC = _warm_scales(L, n)
    s = SDPState(C=BCOO.from_scipy_sparse(L), A_indices=A_indices, A_data=A_data, b=b,
                 b_ineq_mask=mask, X=None, P=P, Omega=Omega, y=_padded(old.y, n), z=z,
                 tr_X=float(np.sum(Lam)), primal_obj=primal_obj, SCALE_C=SC, SCALE_X=SX,
                 SCALE_A=np.ones(n))
    return scale_sdp_state(s)


def get_dual_only_warm_start_state(old_sdp_state, C, sketch_dim, Omega=None):
    """utils/maxcut_helpers.py:232-289: only y carries over (zero-padded); P,
    z, tr_X and the primal objective start at zero with a fresh test matrix
    (default seeded normal with seed 0); rescaled as the other warm starts."""
    if sketch_dim <= 0:
        raise NotImplementedError("the HIP backend keeps a Nystrom sketch (sketch_dim > 0)")
    old = unscale_sdp_state(old_sdp_state)
    L, n, A_data, A_indices, b, mask = _new_problem(C)
    if Omega is None:
        Omega = problems.gaussian_matrix(n, sketch_dim, 0)
    SX, SC = _warm_scales(L, n)
    s = SDPState(C=BCOO.from_scipy_sparse(L), A_indices=A_indices, A_data=A_data, b=b,
                 b_ineq_mask=mask, X=None, P=np.zeros_like(Omega), Omega=Omega,
                 y=_padded(old.y, n), z=np.zeros(n), tr_X=0.0, primal_obj=0.0, SCALE_C=SC,
                 SCALE_X=SX, SCALE_A=np.ones(n))
    return scale_sdp_state(s)


def _as_scipy(C):
    return C.to_scipy() if isinstance(C, BCOO) else scipy.sparse.csr_matrix(C)
