"""Pins of oracle/ipm.py: the prox QP against scipy's SLSQP on the same
convex problem, and the spectral lower bound against its closed form."""
import numpy as np
import scipy.optimize
import scipy.sparse

from oracle import ipm, maxcut as omc, operators as ops
from usbs_amd import problems


def _setup(k, seed):
    rng = np.random.default_rng(seed)
    W = problems.erdos_renyi(40, 0.2, seed=seed)
    C = omc.laplacian_cost(W)
    n = C.shape[0]
    A_data, A_indices, b, mask = omc.get_all_problem_data(n)
    V, _ = np.linalg.qr(rng.standard_normal((n, k)))
    U = ops.create_svec_matrix(k)
    y = rng.standard_normal(n)
    z_bar = rng.random(n) * 0.5
    return C, A_data, A_indices, b, U, V, y, z_bar, n


def test_quad_subproblem_matches_slsqp():
    k = 2
    C, A_data, A_indices, b, U, V, y, z_bar, n = _setup(k, 0)
    tu, rho, pb, trXb = 2.0 * n, 1.0, 3.0, 50.0
    ups = np.zeros(n)
    Q11, q12, q22, h1, h2, trXb2 = ipm.quad_subprob_data(C, A_data, A_indices, U, b, ups, tu, rho, pb,
                                                          z_bar, trXb, y, V, n, n, k)
    S_init = tu * 0.99 * np.eye(k) / (k + 1)
    eta_init = (tu / trXb2) * 0.99 / (k + 1)
    (eta, S), st = ipm.solve_quad_subprob_ipm(C, A_data, A_indices, U, b, ups, tu, rho, pb, z_bar,
                                              trXb, y, V, n, n, k, eta_init, S_init, 1e-10, 100,
                                              return_info=True)
    Sn, en = S / tu, eta * trXb2 / tu   # back to the QP's scaled variables

    def f(x):
        L = np.array([[x[0], 0], [x[1], x[2]]])
        s = U @ (L @ L.T).ravel()
        return 0.5 * (s @ Q11 @ s + 2 * x[3] * (q12 @ s) + q22 * x[3] ** 2) + h1 @ s + h2 * x[3]

    cons = [{"type": "ineq", "fun": lambda x: 1 - (x[0] ** 2 + x[1] ** 2 + x[2] ** 2) - x[3]},
            {"type": "ineq", "fun": lambda x: x[3]}]
    best = None
    for start in ([0.5, 0.0, 0.5, 0.1], [0.1, 0.3, 0.2, 0.5], [0.7, -0.1, 0.1, 0.0]):
        res = scipy.optimize.minimize(f, start, constraints=cons, method="SLSQP",
                                      options={"ftol": 1e-14, "maxiter": 1000})
        best = res if best is None or res.fun < best.fun else best
    L = np.linalg.cholesky(Sn + 1e-300 * np.eye(2))
    val_ipm = f([L[0, 0], L[1, 0], L[1, 1], en])
    assert val_ipm <= best.fun + 1e-6 * max(1.0, abs(best.fun))
    assert np.linalg.eigvalsh(Sn)[0] > 0 and en > 0 and np.trace(Sn) + en <= 1 + 1e-9


def test_lower_bound_closed_form():
    """min of g1.svec(S) + g2 eta over {S >= 0, eta >= 0, tr S + eta <= 1} is
    min(0, lambda_min(svec^{-1} g1), g2): the lower bound is b.y minus it."""
    k = 3
    C, A_data, A_indices, b, U, V, y, z_bar, n = _setup(k, 1)
    # V spanned by eigenvectors of A*(y) - C, as the bundle is in USBS (the
    # reference picks omega from diag(T) only, so it needs a near-diagonal T)
    M = ops.apply_A_adjoint_batched(n, A_data, A_indices, y, np.eye(n)) - C.toarray()
    V = np.linalg.eigh(M)[1][:, [0, 3, 7]]
    tu, pb, trXb = 2.0 * n, 4.0, 60.0
    lb = ipm.compute_lb_spec_est_ipm(C, A_data, A_indices, U, b, tu, pb, z_bar, trXb, y, V, n, k,
                                     1e-12, 200)
    G = V.T @ ops.apply_A_adjoint_batched(n, A_data, A_indices, y, V) - V.T @ (C @ V)
    g2 = (tu / trXb) * (z_bar @ y - pb)
    closed = b @ y - min(0.0, tu * np.linalg.eigvalsh(0.5 * (G + G.T))[0], g2)
    assert abs(lb - closed) <= 1e-6 * max(1.0, abs(closed))
