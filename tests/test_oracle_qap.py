"""The lifted QAP builder (utils/qap_helpers.py:157-381): the oracle's
entry-by-entry restatement pinned by mathematics, and the library's
vectorised host builder checked against it exactly."""
import numpy as np
import pytest

from oracle import qap as oq
from oracle.common import unscale_sdp_state
from usbs_amd import problems, qap_helpers


def _instance(l, seed):
    D, F = problems.qap_instance(l, seed)
    return D, F


def test_objective_matrix_is_padded_kronecker():
    """C[1:, 1:] = W (x) D (numpy.kron, W outer), border empty."""
    D, W = _instance(5, 1)
    rows, cols, vals = oq.build_objective_matrix(D.tolist(), W.tolist())
    N = 26
    C = np.zeros((N, N))
    C[rows, cols] = vals
    np.testing.assert_array_equal(C[1:, 1:], np.kron(W, D))
    assert not C[0].any() and not C[:, 0].any()


def test_lifted_assignment_is_feasible_and_prices_the_qap():
    """For a permutation (not an involution, so W and D are not
    interchangeable), X = x x^T with x = [1; vec(P)] satisfies every equality
    row exactly and every inequality row, and <W (x) D, X> is the QAP cost
    sum W[a,b] D[perm a, perm b].  A non-assignment (two facilities on one
    location) breaks the row/column-sum rows."""
    l = 5
    D, W = _instance(l, 2)
    N = l * l + 1
    rows, cols, vals = oq.build_objective_matrix(D.tolist(), W.tolist())
    A_data, A_ind, b, mask = oq.get_all_problem_data(N, rows, cols)
    A = oq.constraint_matrix(A_data, A_ind, b.size, N)
    perm = [2, 0, 4, 1, 3]
    x = oq.lifted_assignment(perm)
    AX = A @ np.outer(x, x).ravel()
    eq = mask == 0
    np.testing.assert_allclose(AX[eq], b[eq], atol=1e-12)
    assert np.all(AX[~eq] <= b[~eq] + 1e-12)
    C = np.zeros((N, N))
    C[rows, cols] = vals
    cost = sum(W[a, c] * D[perm[a], perm[c]] for a in range(l) for c in range(l))
    assert x @ C @ x == pytest.approx(cost, rel=1e-14)
    assert qap_helpers.qap_objective(D, W, perm) == pytest.approx(cost, rel=1e-14)
    bad = oq.lifted_assignment([2, 2, 4, 1, 3])
    AXb = A @ np.outer(bad, bad).ravel()
    assert np.max(np.abs(AXb[eq] - b[eq])) >= 1.0


def test_constraint_count_and_families():
    """m = 1 + (N-1) + 2l + 2 l^2 + #upper-triangle entries of C; the
    partial-trace rows of I itself: tr_1(I) = tr_2(I) = l I restricted to the
    Y block gives 2l on diagonal rows and 0 elsewhere before scaling."""
    l = 4
    D, W = _instance(l, 3)
    N = l * l + 1
    rows, cols, vals = oq.build_objective_matrix(D.tolist(), W.tolist())
    A_data, A_ind, b, mask = oq.get_all_problem_data(N, rows, cols)
    ntriu = sum(1 for r, c in zip(rows, cols) if r <= c)
    assert b.size == 1 + (N - 1) + 2 * l + 2 * l * l + ntriu
    assert mask.sum() == ntriu
    A = oq.constraint_matrix(A_data, A_ind, b.size, N)
    Y = np.zeros((N, N))
    Y[1:, 1:] = np.eye(N - 1)
    AY = A @ Y.ravel()
    base = 1 + (N - 1) + 2 * l
    tr = AY[base:base + 2 * l * l].reshape(2, l, l)
    for t in range(2):
        np.testing.assert_array_equal(tr[t], 2 * l * np.eye(l))


def test_scaled_constraint_operator_has_unit_norm():
    """After initialize_state the scaled constraint matrix has spectral norm
    1 (the defining property of SCALE_A, utils/qap_helpers.py:329-337),
    checked with a dense SVD against the library's eigsh-based value."""
    l = 4
    D, W = _instance(l, 4)
    N = l * l + 1
    Om = problems.gaussian_matrix(N, l, 1)
    C = qap_helpers.build_objective_matrix(D, W)
    s = qap_helpers.initialize_state(C, l, Omega=Om)
    A = oq.constraint_matrix(s.A_data, s.A_indices, s.b.size, N)
    assert np.linalg.norm(A, 2) == pytest.approx(1.0, rel=1e-10)
    so, _ = oq.initialize_state(D, W, l, Om)
    np.testing.assert_allclose(s.SCALE_A, so.SCALE_A, rtol=1e-12)
    assert s.SCALE_C == pytest.approx(so.SCALE_C, rel=1e-15)
    assert s.SCALE_X == so.SCALE_X == 1.0 / (l + 1)
    u = unscale_sdp_state(so)
    np.testing.assert_allclose(u.b[:1 + N - 1], np.r_[1.0, np.zeros(N - 1)], atol=1e-15)


@pytest.mark.parametrize("l,seed", [(4, 5), (5, 6), (7, 7)])
def test_library_builder_equals_oracle(l, seed):
    """The vectorised builder reproduces the entry-by-entry one exactly, in
    the same row and triple order."""
    D, W = _instance(l, seed)
    C = qap_helpers.build_objective_matrix(D, W)
    rows, cols, vals = oq.build_objective_matrix(D.tolist(), W.tolist())
    np.testing.assert_array_equal(C.indices[:, 0], rows)
    np.testing.assert_array_equal(C.indices[:, 1], cols)
    np.testing.assert_array_equal(C.data, vals)
    got = qap_helpers.get_all_problem_data(-C)
    want = oq.get_all_problem_data(l * l + 1, rows, cols)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)

