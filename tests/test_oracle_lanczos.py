"""Pins of oracle/lanczos.py (solver/lanczos.py) against dense eigensolvers,
the torus closed form and hand-worked arrowhead / restart cases."""
import numpy as np
import scipy.sparse

from oracle import lanczos as olz, maxcut as omc
from usbs_amd import problems


def _op(W, seed):
    C = omc.laplacian_cost(W)
    n = C.shape[0]
    A_data, A_indices, _, _ = omc.get_all_problem_data(n)
    y = np.random.default_rng(seed).standard_normal(n)
    return (-C).tocsr(), A_data, A_indices, y, problems.unit_start_vector(n, seed)


def test_matches_dense_eigh():
    negC, A_data, A_indices, y, q0 = _op(problems.erdos_renyi(150, 0.1, seed=0), 1)
    n = negC.shape[0]
    vals, vecs = olz.eigsh_smallest(n, negC, A_data, A_indices, y, q0, 3, 32, 50, 1e-10)
    M = (negC + scipy.sparse.diags(y)).toarray()
    dense = np.linalg.eigvalsh(M)
    np.testing.assert_allclose(vals, dense[:3], rtol=1e-9)
    for j in range(3):
        assert np.linalg.norm(M @ vecs[:, j] - vals[j] * vecs[:, j]) < 1e-8
    np.testing.assert_allclose(vecs.T @ vecs, np.eye(3), atol=1e-12)


def test_torus_closed_form():
    """Unit torus R x S: Laplacian eigenvalues 4 - 2cos(2 pi a/R) - 2cos(2 pi b/S);
    the smallest eigenvalue of -C = -L/4 is -max/4."""
    R, S = 10, 20
    W = problems.torus(R, S, seed=0, signed=False)
    negC, A_data, A_indices, _, q0 = _op(W, 2)
    n = R * S
    lam = sorted(4 - 2 * np.cos(2 * np.pi * a / R) - 2 * np.cos(2 * np.pi * b / S)
                 for a in range(R) for b in range(S))
    vals, _ = olz.eigsh_smallest(n, negC, A_data, A_indices, np.zeros(n), q0, 1, 32, 100, 1e-10)
    np.testing.assert_allclose(vals[0], -lam[-1] / 4, rtol=1e-10)


def test_arrowhead_hand_cases():
    a = np.array([1.0, 2.0, 3.0, 4.0])
    b = np.array([0.5, 0.25, 0.125, 9.0])
    # k = 0: plain tridiagonal with beta on the first super-diagonal
    T0 = olz._build_arrowhead_matrix(a, b, 0)
    np.testing.assert_array_equal(T0, [[1, .5, 0, 0], [.5, 2, .25, 0], [0, .25, 3, .125], [0, 0, .125, 4]])
    # k = 2: arrow column at index 2 holding beta[0:2], band from index 2
    T2 = olz._build_arrowhead_matrix(a, b, 2)
    np.testing.assert_array_equal(T2, [[1, 0, .5, 0], [0, 2, .25, 0], [.5, .25, 3, .125], [0, 0, .125, 4]])


def test_pick_saved_hand_case():
    lam = np.arange(10, dtype=float)  # evenly spaced Ritz values
    saved, gamma = olz._pick_saved_ritz_values(lam, 0, 3)
    # delta = min(10-3, 2*10//5) = 4; gamma_j = j/(j+4) for j+4 <= 10 (clamped index 9 at j=6)
    assert np.isclose(gamma[5], 5 / 9) and np.isclose(gamma[6], 6 / 9) and gamma[7] == 0
    # argmax is j=6 -> keep 0..6 and nothing past 6+4=10
    assert saved.tolist() == [True] * 7 + [False] * 3


def test_breakdown_small_n():
    negC, A_data, A_indices, y, q0 = _op(problems.erdos_renyi(12, 0.5, seed=3), 4)
    n = negC.shape[0]
    vals, vecs = olz.eigsh_smallest(n, negC, A_data, A_indices, y, q0, 2, n, 3, 1e-9)
    dense = np.linalg.eigvalsh((negC + scipy.sparse.diags(y)).toarray())
    np.testing.assert_allclose(vals, dense[:2], rtol=1e-9)


def test_breakdown_normal_stream():
    x = olz.breakdown_normal(5, 200000)
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1) < 0.01
    assert np.array_equal(x[:10], olz.breakdown_normal(5, 10))
