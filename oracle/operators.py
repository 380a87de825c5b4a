"""Constraint operators A(.), A*(.), svec and Q_base in FP64 NumPy.

Restates solver/utils.py:11-65.  A constraint tensor is a list of triples
(c, i, j) with values A_data: A_c[i, j] += A_data.  Oracle code: see
oracle/__init__.py.
"""
import numpy as np
import scipy.sparse


def apply_A_operator_slim(m, A_data, A_indices, u):
    """z[c] += A_data * u[i] * u[j]  (solver/utils.py:11-16)."""
    A_indices = np.asarray(A_indices)
    w = A_data * u[A_indices[:, 1]] * u[A_indices[:, 2]]
    return np.bincount(A_indices[:, 0], weights=w, minlength=m).astype(np.float64)


def apply_A_adjoint_slim(n, A_data, A_indices, z, u):
    """out[i] += A_data * z[c] * u[j]  (solver/utils.py:19-24)."""
    A_indices = np.asarray(A_indices)
    w = A_data * z[A_indices[:, 0]] * u[A_indices[:, 2]]
    return np.bincount(A_indices[:, 1], weights=w, minlength=n).astype(np.float64)


def apply_A_operator_batched(m, A_data, A_indices, vecs):
    """sum over columns of apply_A_operator_slim  (solver/utils.py:27-31)."""
    A_indices = np.asarray(A_indices)
    w = A_data[:, None] * vecs[A_indices[:, 1], :] * vecs[A_indices[:, 2], :]
    return np.bincount(A_indices[:, 0], weights=w.sum(axis=1), minlength=m).astype(np.float64)


def apply_A_adjoint_batched(n, A_data, A_indices, z, vecs):
    """column-wise apply_A_adjoint_slim  (solver/utils.py:34-37)."""
    A_indices = np.asarray(A_indices)
    coef = A_data * z[A_indices[:, 0]]
    out = np.zeros((n, vecs.shape[1]))
    for col in range(vecs.shape[1]):
        out[:, col] = np.bincount(A_indices[:, 1], weights=coef * vecs[A_indices[:, 2], col],
                                  minlength=n)
    return out


def create_svec_matrix(k):
    """Dense svec basis U (k(k+1)/2 x k^2), rows in np.tril_indices order
    (solver/utils.py:40-50).  Off-diagonal pairs carry 1/sqrt(2) so that
    svec is an isometry on symmetric matrices."""
    U = np.zeros((k * (k + 1) // 2, k * k))
    for a, (b, c) in enumerate(zip(*np.tril_indices(k))):
        if b == c:
            U[a, b * k + c] = 1.0
        else:
            U[a, b * k + c] = 1.0 / np.sqrt(2.0)
            U[a, c * k + b] = U[a, b * k + c]
    return U


def create_Q_base(m, k, U, A_data, A_indices, V):
    """sum_c svec(V^T A_c V) svec(V^T A_c V)^T  (solver/utils.py:53-65)."""
    A_indices = np.asarray(A_indices)
    outer = A_data[:, None, None] * (V[A_indices[:, 1], :][:, :, None] * V[A_indices[:, 2], :][:, None, :])
    nA = A_indices.shape[0]
    gather = scipy.sparse.csr_matrix((np.ones(nA), (A_indices[:, 0], np.arange(nA))), shape=(m, nA))
    base = gather @ outer.reshape(nA, k * k)
    svec_proj = U @ base.T            # (d, m)
    return svec_proj @ svec_proj.T
