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
from .common import SDPState, scale_sdp_state
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
        raise NotImplementedError("dense-X storage (sketch_dim=-1) is not provided by the HIP backend")
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
