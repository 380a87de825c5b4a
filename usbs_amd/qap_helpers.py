"""The lifted QAP SDP, mirroring utils/qap_helpers.py:157-381.

A QAP with flow matrix W and distance matrix D (both l x l) is lifted to an
SDP over X of order N = l^2 + 1: row/column 0 is the homogenising border,
entry 1 + l*a + i stands for "facility a sits at location i".  The problem
build is host-side setup, as in the reference (which builds it with JAX and
scales it with scipy); the solve runs on the device through the same
usbs()/cgal() entry points as max-cut.

Layout facts the device operator relies on (and that the reference fixes):

* ``C[1 + l*a + i, 1 + l*b + j] = W[a, b] * D[i, j]``, entries in the order
  W's nonzeros (row-major) x D's nonzeros (row-major)
  (utils/qap_helpers.py:157-172), negated by initialize_state.
* constraint rows, in this order (utils/qap_helpers.py:175-318):
  X00 = 1 | diag(Y) = vec(B) (N-1 rows) | B1 = 1 (l) | 1'B = 1' (l) |
  tr_1(Y) = I (l^2) | tr_2(Y) = I (l^2) | -Y <= 0 on C's upper triangle.
"""
import ctypes

import numpy as np
import scipy.sparse
import scipy.sparse.linalg

from . import _lib, problems
from .common import SDPState, scale_sdp_state
from .sparse import BCOO


def build_objective_matrix(D, W):
    """C = W (x) D padded by one leading row and column
    (utils/qap_helpers.py:157-172).  D, W: dense l x l arrays; zeros are
    not stored."""
    D = np.asarray(D, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    l = D.shape[0]
    wi, wj = np.nonzero(W)
    di, dj = np.nonzero(D)
    rows = 1 + (l * wi)[:, None] + di[None, :]
    cols = 1 + (l * wj)[:, None] + dj[None, :]
    data = W[wi, wj][:, None] * D[di, dj][None, :]
    idx = np.stack([rows.ravel(), cols.ravel()], axis=1)
    return BCOO((data.ravel(), idx), shape=(l * l + 1, l * l + 1))


def _pairs_with_border(cons, v):
    """Triples (c, v, 0), (c, 0, v) interleaved per entry, data 1 each."""
    t = np.zeros((2 * v.size, 3), dtype=np.int64)
    t[0::2, 0] = cons
    t[1::2, 0] = cons
    t[0::2, 1] = v
    t[1::2, 2] = v
    return t


def get_all_problem_data(C):
    """(A_data, A_indices, b, b_ineq_mask) of the lifted QAP for a cost
    matrix C of order l^2 + 1 (utils/qap_helpers.py:175-318)."""
    N = C.shape[0]
    l = int(round(np.sqrt(N - 1)))
    if l * l + 1 != N:
        raise ValueError("C must have order l^2 + 1")
    trip, data, b, mask = [], [], [], []

    def add(t, d, bb, mm):
        trip.append(np.asarray(t, dtype=np.int64).reshape(-1, 3))
        data.append(np.asarray(d, dtype=np.float64))
        b.append(np.asarray(bb, dtype=np.float64))
        mask.append(np.asarray(mm, dtype=np.float64))

    base = 0
    # X(0, 0) = 1
    add([[0, 0, 0]], [1.0], [1.0], [0.0])
    base += 1
    # diag(Y) = vec(B): -X[j,0] - X[0,j] + 2 X[j,j] = 0 for j = 1..N-1
    j = np.arange(1, N)
    t = np.zeros((3 * (N - 1), 3), dtype=np.int64)
    t[:, 0] = base + np.repeat(np.arange(N - 1), 3)
    t[0::3, 1] = j
    t[1::3, 2] = j
    t[2::3, 1] = j
    t[2::3, 2] = j
    add(t, np.tile([-1.0, -1.0, 2.0], N - 1), np.zeros(N - 1), np.zeros(N - 1))
    base += N - 1
    # B 1 = 1 (row a sums entries 1 + l*a + i over i) and 1'B = 1' (column i)
    grid = np.arange(1, N).reshape(l, l)
    for order in (grid, grid.T):
        v = order.ravel()
        add(_pairs_with_border(base + np.repeat(np.arange(l), l), v), np.ones(2 * (N - 1)),
            2.0 * np.ones(l), np.zeros(l))
        base += l
    # tr_1(Y) = I: row (j1, j2) sums Y[(d, j1), (d, j2)] over blocks d;
    # tr_2(Y) = I: row (j1, j2) sums Y[(j1, d), (j2, d)] over inner index d.
    j1, j2, d = np.meshgrid(np.arange(l), np.arange(l), np.arange(l), indexing="ij")
    for ca, cb in ((1 + j1 + l * d, 1 + j2 + l * d), (1 + l * j1 + d, 1 + l * j2 + d)):
        cons = base + (j1 * l + j2).ravel()
        t1 = np.stack([cons, ca.ravel(), cb.ravel()], axis=1)
        t2 = np.stack([cons, cb.ravel(), ca.ravel()], axis=1)
        add(np.concatenate([t1, t2]), np.ones(2 * t1.shape[0]),
            2.0 * (np.arange(l)[:, None] == np.arange(l)[None, :]).ravel(), np.zeros(l * l))
        base += l * l
    # -Y <= 0 on the entries of C's upper triangle (diagonal included)
    ci = np.asarray(C.indices)
    up = ci[ci[:, 0] <= ci[:, 1]]
    cons = base + np.arange(up.shape[0])
    t1 = np.stack([cons, up[:, 0], up[:, 1]], axis=1)
    t2 = np.stack([cons, up[:, 1], up[:, 0]], axis=1)
    add(np.concatenate([t1, t2]), np.full(2 * up.shape[0], -0.5), np.zeros(up.shape[0]),
        np.ones(up.shape[0]))
    return (np.concatenate(data), np.concatenate(trip), np.concatenate(b), np.concatenate(mask))


def constraint_norm(A_data, A_indices, m, N):
    """Spectral norm of the m x N^2 constraint matrix (duplicate triples
    summed), as sqrt of the largest eigenvalue of A A^T by Lanczos
    (scipy's eigsh, as utils/qap_helpers.py:333-336)."""
    A = scipy.sparse.coo_matrix((A_data, (A_indices[:, 0], A_indices[:, 1] * N + A_indices[:, 2])),
                                shape=(m, N * N)).tocsr()
    G = (A @ A.T).tocsr()
    if m <= 3:
        return float(np.sqrt(np.linalg.eigvalsh(G.toarray())[-1]))
    lam = scipy.sparse.linalg.eigsh(G, k=1, which="LM", return_eigenvectors=False,
                                    maxiter=np.iinfo(np.int32).max)[0]
    return float(np.sqrt(lam))


def initialize_state(C, sketch_dim, omega_seed=0, Omega=None):
    """Cold start of the lifted QAP (utils/qap_helpers.py:321-381): negate C,
    SCALE_X = 1/(l+1), SCALE_C = 1/||C||_F, SCALE_A = row norms of A divided
    by ||A||_2, then scale.  Omega is a seeded standard normal N x sketch_dim
    matrix (the reference draws it from jax.random with PRNGKey(0))."""
    C = -C
    A_data, A_indices, b, b_ineq_mask = get_all_problem_data(C)
    N = C.shape[0]
    m = b.shape[0]
    l = int(round(np.sqrt(N - 1)))
    SCALE_X = 1.0 / float(l + 1)
    SCALE_C = 1.0 / float(np.linalg.norm(C.data))
    SCALE_A = 1.0 / np.sqrt(np.bincount(A_indices[:, 0], weights=A_data ** 2, minlength=m))
    SCALE_A = SCALE_A / constraint_norm(SCALE_A[A_indices[:, 0]] * A_data, A_indices, m, N)
    if sketch_dim == -1:
        raise NotImplementedError("dense-X storage (sketch_dim=-1) is not provided by the HIP backend")
    elif sketch_dim > 0:
        if Omega is None:
            Omega = problems.gaussian_matrix(N, sketch_dim, omega_seed)
        P = np.zeros_like(Omega)
    else:
        raise ValueError("Invalid value for sketch_dim")
    s = SDPState(C=C, A_indices=A_indices, A_data=A_data, b=b, b_ineq_mask=b_ineq_mask, X=None,
                 P=P, Omega=Omega, y=np.zeros(m), z=np.zeros(m), tr_X=0.0, primal_obj=0.0,
                 SCALE_C=SCALE_C, SCALE_X=SCALE_X, SCALE_A=SCALE_A)
    return scale_sdp_state(s)


def qap_objective(D, W, perm):
    """sum_{a,b} W[a,b] D[perm[a], perm[b]]: the QAP cost of placing
    facility a at location perm[a]."""
    perm = np.asarray(perm)
    return float(np.sum(np.asarray(W) * np.asarray(D)[np.ix_(perm, perm)]))


_lib.EXTRA_SIGNATURES["usbs_qap_kron_matvec"] = [
    ctypes.c_void_p, ctypes.c_int32, _lib._dp, _lib._dp, ctypes.c_int64, _lib._dp, ctypes.c_int32,
    ctypes.c_int32, _lib._dp, _lib._dp]

KRON_PRECISION = {"fp64": 0, "bf16": 2}


def kron_matvec(D, W, Y, precision="bf16", reps=1, context=None, device=0, return_time=False):
    """C v for a batch of vectors, C = W (x) D padded (build_objective_matrix),
    on the device: Y is (nvec, l, l) (or (nvec, l*l)) with Y[b, c, j] =
    v_b[1 + l c + j]; returns U (nvec, l, l) with (C v_b)[1 + l a + i] =
    U[b, a, i] ((C v_b)[0] = 0).  precision "bf16": BF16 operands on the
    matrix cores, FP32 accumulation; "fp64": FP64 FMA.  With return_time
    also the mean kernel time (ms) of ``reps`` launches on device-resident
    inputs."""
    D = np.ascontiguousarray(D, dtype=np.float64)
    W = np.ascontiguousarray(W, dtype=np.float64)
    l = D.shape[0]
    if D.shape != (l, l) or W.shape != (l, l):
        raise ValueError("D and W must be l x l")
    Y = np.ascontiguousarray(Y, dtype=np.float64).reshape(-1, l, l)
    U = np.zeros_like(Y)
    ms = ctypes.c_double(0.0)
    ctx = context if context is not None else _lib.Context(device)
    ctx.check(_lib.lib().usbs_qap_kron_matvec(ctx.h, l, _lib.dptr(W), _lib.dptr(D), Y.shape[0],
                                              _lib.dptr(Y), KRON_PRECISION[precision], int(reps),
                                              _lib.dptr(U), ctypes.byref(ms)))
    return (U, ms.value) if return_time else U


_lib.EXTRA_SIGNATURES.update({
    "usbs_munkres": [ctypes.c_int32, _lib._dp, _lib._ip],
    "usbs_qap_round": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _lib._dp, _lib._dp, _lib._dp,
                       _lib._dp, _lib._dp, _lib._ip, _lib._dp],
    "usbs_set_qap_rounding": [ctypes.c_void_p, ctypes.c_int32, _lib._dp, _lib._dp],
})


def munkres(n, cost_mx):
    """Minimum-cost assignment of an n x n cost matrix as the reference's
    mask (1 at the assigned entries; utils/munkres.py:14-120), by the
    library's host-side Hungarian algorithm with the reference's step order,
    tie-breaking and n^2-step bound."""
    C = np.ascontiguousarray(cost_mx, dtype=np.float64)
    if C.shape != (n, n):
        raise ValueError("cost_mx must be n x n")
    out = np.zeros((n, n), dtype=np.int32)
    rc = _lib.lib().usbs_munkres(int(n), _lib.dptr(C), _lib.iptr(out))
    if rc != 0:
        raise _lib.USBSError(rc, "usbs_munkres failed")
    return out.astype(np.float64)


def _W_D(callback_nonstatic_args):
    W = np.ascontiguousarray(callback_nonstatic_args["W"], dtype=np.float64)
    D = np.ascontiguousarray(callback_nonstatic_args["D"], dtype=np.float64)
    l = W.shape[0]
    if W.shape != (l, l) or D.shape != (l, l):
        raise ValueError("W and D must be l x l")
    return l, W, D


def qap_round(P, Omega, callback_static_args, callback_nonstatic_args, context=None, device=0):
    """QAP rounding callback (utils/qap_helpers.py:625-653): the smallest
    QAP objective trace(W Pi D Pi^T) over the assignments rounded from the
    Nystrom columns of the sketch and their negatives.  The reference passes
    l pickled in callback_static_args; here l is W's order (nothing is
    unpickled).  callback_nonstatic_args: {"W": flow, "D": distance}.
    Passed as callback_fn to usbs(), it runs natively at every condition
    check; called directly, it runs the same code on the given arrays."""
    l, W, D = _W_D(callback_nonstatic_args)
    P, Omega = _lib.as_f64(P), _lib.as_f64(Omega)
    r = P.shape[1]
    ctx = context if context is not None else _lib.Context(device)
    best = ctypes.c_double(0.0)
    ctx.check(_lib.lib().usbs_qap_round(ctx.h, l, r, _lib.dptr(P), _lib.dptr(Omega), _lib.dptr(W),
                                        _lib.dptr(D), None, None, ctypes.byref(best)))
    return best.value


def set_rounding(ctx, callback_nonstatic_args):
    """Hand W and D to a context for usbs()'s native QAP rounding."""
    l, W, D = _W_D(callback_nonstatic_args)
    ctx.check(_lib.lib().usbs_set_qap_rounding(ctx.h, l, _lib.dptr(W), _lib.dptr(D)))
