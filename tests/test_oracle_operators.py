"""Pins of oracle/operators.py (solver/utils.py:11-65) by brute force on
dense constraint tensors and by the svec isometry."""
import numpy as np

from oracle import operators as ops


def _random_triples(rng, n, m, nA):
    A_indices = np.stack([rng.integers(0, m, nA), rng.integers(0, n, nA), rng.integers(0, n, nA)], 1)
    return rng.standard_normal(nA), A_indices


def _dense_A(n, m, A_data, A_indices):
    A = np.zeros((m, n, n))
    for d, (c, i, j) in zip(A_data, A_indices):
        A[c, i, j] += d
    return A


def test_operator_and_adjoint_brute_force():
    rng = np.random.default_rng(0)
    n, m, nA = 7, 5, 40
    A_data, A_indices = _random_triples(rng, n, m, nA)
    A = _dense_A(n, m, A_data, A_indices)
    u, z = rng.standard_normal(n), rng.standard_normal(m)
    # A(u u^T)_c = <A_c, u u^T>, A*(z) u = (sum_c z_c A_c) u
    np.testing.assert_allclose(ops.apply_A_operator_slim(m, A_data, A_indices, u),
                               np.einsum("cij,i,j->c", A, u, u), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ops.apply_A_adjoint_slim(n, A_data, A_indices, z, u),
                               np.einsum("c,cij->ij", z, A) @ u, rtol=1e-12, atol=1e-12)
    V = rng.standard_normal((n, 3))
    np.testing.assert_allclose(ops.apply_A_operator_batched(m, A_data, A_indices, V),
                               np.einsum("cij,ia,ja->c", A, V, V), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ops.apply_A_adjoint_batched(n, A_data, A_indices, z, V),
                               np.einsum("c,cij->ij", z, A) @ V, rtol=1e-12, atol=1e-12)
    # adjointness: <A(uu^T), z> = <u, A*(z) u>
    assert np.isclose(ops.apply_A_operator_slim(m, A_data, A_indices, u) @ z,
                      u @ ops.apply_A_adjoint_slim(n, A_data, A_indices, z, u))


def test_svec_is_isometry():
    rng = np.random.default_rng(1)
    for k in (1, 2, 5, 11):
        U = ops.create_svec_matrix(k)
        A, B = rng.standard_normal((k, k)), rng.standard_normal((k, k))
        A, B = A + A.T, B + B.T
        assert np.isclose((U @ A.ravel()) @ (U @ B.ravel()), np.sum(A * B))
        # U^T U is the orthogonal projector onto symmetric matrices
        np.testing.assert_allclose((U.T @ (U @ A.ravel())).reshape(k, k), A, atol=1e-12)


def test_Q_base_brute_force():
    rng = np.random.default_rng(2)
    n, m, nA, k = 6, 4, 25, 3
    A_data, A_indices = _random_triples(rng, n, m, nA)
    A = _dense_A(n, m, A_data, A_indices)
    V = rng.standard_normal((n, k))
    U = ops.create_svec_matrix(k)
    Q = ops.create_Q_base(m, k, U, A_data, A_indices, V)
    ref = np.zeros_like(Q)
    for c in range(m):
        M = V.T @ A[c] @ V
        s = np.array([M[b, a] if a == b else (M[b, a] + M[a, b]) / np.sqrt(2)
                      for b in range(k) for a in range(b + 1)])
        ref += np.outer(s, s)
    np.testing.assert_allclose(Q, ref, rtol=1e-12, atol=1e-12)
