"""Pins of the oracle's loop schedule, Nystrom reconstruction, rounding and
the end-to-end USBS loop (closed forms of the max-cut SDP)."""
import numpy as np
import pytest

from oracle import common as ocommon, loop as oloop, maxcut as omc, usbs as ousbs
from usbs_amd import problems


def test_geometric_schedule_hand_case():
    # base 2, maxiter 10: num = ceil(log2 9) + 1 = 5 -> [1,2,4,8,16]; keep
    # prefixes with cumsum < 10 -> [1,2,4]; remainder 3.
    assert oloop.geometric_schedule(10, 2.0).tolist() == [1, 2, 4, 3]


def test_while_loop_counts():
    calls = []
    out = oloop.while_loop(lambda v: (calls.append(v) or v < 5), lambda v: v + 1, 0, 100)
    assert out == 5 and calls == [0, 1, 2, 3, 4, 5]
    assert oloop.while_loop(lambda v: True, lambda v: v + 1, 0, 7) == 7


def test_nystrom_exact_recovery_small_shift():
    """Tropp et al. 2017: rank(X) <= r is recovered exactly (shift -> 0)."""
    rng = np.random.default_rng(0)
    n, r, rank = 300, 8, 5
    G = rng.standard_normal((n, rank))
    Om = rng.standard_normal((n, r))
    E, lam = ocommon.reconstruct_from_sketch(Om, G @ (G.T @ Om), approx_eps=1e-15)
    np.testing.assert_allclose((E * lam) @ E.T, G @ G.T, atol=1e-8 * np.abs(G @ G.T).max())
    np.testing.assert_allclose(np.sort(lam)[::-1][:rank], np.linalg.eigvalsh(G.T @ G)[::-1], rtol=1e-8)


def test_nystrom_zero_sketch_nan():
    E, lam = ocommon.reconstruct_from_sketch(np.ones((5, 2)), np.zeros((5, 2)))
    assert np.isnan(E).all()


def test_rounding_bruteforce():
    W = problems.erdos_renyi(12, 0.5, seed=2, signed=True)
    cuts = omc.cut_values_bruteforce(W)
    C = omc.laplacian_cost(W)
    for mask in (0, 5, 1234, 4095):
        s = np.array([1.0 if (mask >> i) & 1 else -1.0 for i in range(12)])
        assert s @ (C @ s) == cuts[mask]   # exact: the cut value as an edge sum
    best = max(cuts, key=cuts.get)
    s = np.array([1.0 if (best >> i) & 1 else -1.0 for i in range(12)])
    Om = np.random.default_rng(3).standard_normal((12, 4))
    P = (np.outer(s, s) + 0.01 * np.eye(12)) @ Om
    assert omc.compute_max_cut(P, Om, None, C) == cuts[best]


def _solve(W, iters, rho, eps):
    n = W.shape[0]
    Om = problems.gaussian_matrix(n, 4, 0)
    s = omc.initialize_state(W, 4, Om)
    q0 = problems.unit_start_vector(n, 1)
    log = []
    _, fin = ousbs.usbs(s, n, n, 2.0 * n, 2.0, rho, 0.25, 4, 1, iters, 1e9, eps, eps, eps, min(n, 32),
                        50, 1e-10, 100, 1.0, None, None, None, q0, log=log)
    return fin, log


@pytest.mark.parametrize("name,W,value", [
    ("bipartite torus", problems.torus(6, 8, seed=0, signed=False), 96.0),
    ("K_8", problems.erdos_renyi(8, 1.0, seed=0), 16.0),
    ("C_5", problems.torus(1, 5, seed=0, signed=False), 2.5 * (1 + np.cos(np.pi / 5))),
])
def test_usbs_closed_forms(name, W, value):
    fin, log = _solve(W, 400, 1.0, 1e-7)
    assert abs(fin.primal_obj - value) <= 1e-4 * value, (name, fin.primal_obj, value, len(log))


@pytest.mark.parametrize("name,W,value,tol", [
    ("bipartite torus", problems.torus(6, 8, seed=0, signed=False), 96.0, 1e-9),
    ("K_8", problems.erdos_renyi(8, 1.0, seed=0), 16.0, 1e-3),
])
def test_cgal_closed_forms(name, W, value, tol):
    """The CGAL restatement reaches the max-cut SDP value of graphs whose
    value is known in closed form (bipartite: |E|; K_n: n^2/4)."""
    from oracle import cgal as ocg
    n = W.shape[0]
    st = omc.initialize_state(W, 4, problems.gaussian_matrix(n, 4, 0))
    out, _ = ocg.cgal(st, n, n, float(n), 1.0, 2000, 1e9, 1e-6, 1e-6, 1e-6, min(n, 32), 50, 1e-10,
                      1.0, None, None, None, problems.unit_start_vector(n, 1))
    assert abs(out.primal_obj - value) <= tol * value, (name, out.primal_obj)
