"""The warm-start transforms of utils/maxcut_helpers.py:95-289.

Oracle pins (closed forms, independent of the oracle's own formulas):
* implicit, same graph: unscaling the new state gives back the old state's
  y, z, P, trace and objective (the diagonal correction vanishes);
* implicit, grown graph: the objective correction is the change of
  <diag C, z> over the old vertices, and the scales are 1/n and 1/|C|_F;
* explicit: for a sketch of an exactly rank-r PSD X with a well-separated
  spectrum, the new state is the sketch of X + delta E E^T (E: X's
  eigenvectors, delta spreading the trace deficit over the r columns):
  its trace equals the old trace, z = diag(X) + delta rowsum(E^2), and the
  objective <C, X + delta E E^T>;
* dual-only: y padded, everything primal zero.
The library's host-side transforms (implicit, dual-only) must equal the
oracle's array for array; the explicit one runs on the device (GPU tests)."""
import numpy as np
import scipy.sparse

from oracle import common as ocommon, maxcut as omc
from usbs_amd import maxcut_helpers, problems
from usbs_amd.sparse import BCOO


def _old_state(W, r, seed):
    n = W.shape[0]
    rng = np.random.default_rng(seed)
    so = omc.initialize_state(W, r, problems.gaussian_matrix(n, r, 1))
    return so._replace(y=rng.standard_normal(n), z=1.0 + 0.1 * rng.standard_normal(n),
                       P=rng.standard_normal((n, r)), tr_X=float(n) * 0.97,
                       primal_obj=123.456)


def _to_pkg(s):
    return s._replace(C=BCOO.from_scipy_sparse(s.C))


def test_implicit_same_graph_round_trip():
    W = problems.erdos_renyi(200, 0.05, seed=1, signed=True)
    old = _old_state(W, 5, 2)
    new = omc.implicit_warm_start_state(old, W, 5, np.zeros((0, 5)))
    assert new.SCALE_X == 1.0 / 200
    assert np.isclose(new.SCALE_C, 1.0 / np.linalg.norm(omc.laplacian_cost(W).data))
    u = ocommon.unscale_sdp_state(new)
    np.testing.assert_allclose(u.y, old.y, rtol=0, atol=0)
    np.testing.assert_allclose(u.z, old.z, rtol=1e-15)
    np.testing.assert_allclose(u.P, old.P, rtol=1e-15)
    assert np.isclose(u.tr_X, old.tr_X, rtol=1e-15) and np.isclose(u.primal_obj, old.primal_obj,
                                                                    rtol=1e-13)


def test_implicit_grown_graph_and_library_equal_oracle():
    W = problems.erdos_renyi(240, 0.05, seed=3, signed=True)
    W_old = scipy.sparse.csr_matrix(W)[:200, :200]
    old = _old_state(W_old, 6, 4)
    pad = np.random.default_rng(5).standard_normal((40, 6))
    new = omc.implicit_warm_start_state(old, W, 6, pad)
    u = ocommon.unscale_sdp_state(new)
    d_old = omc.laplacian_cost(W_old).diagonal()
    d_new = omc.laplacian_cost(W).diagonal()
    expect = old.primal_obj - d_old @ old.z + d_new[:200] @ old.z
    assert np.isclose(u.primal_obj, expect, rtol=1e-13)
    assert np.array_equal(u.y[200:], np.zeros(40)) and np.array_equal(new.Omega[200:], pad)
    lib = maxcut_helpers.get_implicit_warm_start_state(_to_pkg(old), W, 6, Omega_pad=pad)
    for f in ("y", "z", "P", "Omega", "b", "A_data"):
        assert np.array_equal(np.asarray(getattr(lib, f)), np.asarray(getattr(new, f))), f
    for f in ("tr_X", "primal_obj", "SCALE_C", "SCALE_X"):
        assert float(getattr(lib, f)) == float(getattr(new, f)), f
    assert abs(lib.C.to_scipy() - new.C).max() == 0.0


def test_explicit_reconstructs_an_exact_sketch():
    n, r = 150, 5
    W = problems.erdos_renyi(n, 0.06, seed=6)
    rng = np.random.default_rng(7)
    E0, _ = np.linalg.qr(rng.standard_normal((n, r)))
    lam0 = np.array([9.0, 7.0, 5.0, 3.0, 1.0])
    X = (E0 * lam0) @ E0.T
    Om_old = rng.standard_normal((n, r))
    old = omc.initialize_state(W, r, Om_old)._replace(P=X @ Om_old, tr_X=30.0)
    Om_new = rng.standard_normal((n, r))
    new = omc.explicit_warm_start_state(old, W, r, Om_new)
    u = ocommon.unscale_sdp_state(new)
    delta = (30.0 - lam0.sum()) / r
    Xh = X + delta * (E0 @ E0.T)
    C = omc.laplacian_cost(W)
    # the Nystrom reconstruction shifts the sketch by rho = sqrt(n) 1e-6 |P|_2
    # (utils/common.py:103-113), which moves each reconstructed eigenvalue by
    # a multiple of rho (12 rho here, measured); the trace stays exact
    rho = np.sqrt(n) * 1e-6 * np.linalg.norm(old.P, 2)
    tol = 30 * rho
    assert np.isclose(u.tr_X, 30.0, rtol=1e-12)
    np.testing.assert_allclose(u.z, np.diag(Xh), rtol=0, atol=tol)
    np.testing.assert_allclose(u.P, Xh @ Om_new, rtol=0, atol=tol * np.abs(Om_new).sum(0).max())
    assert abs(u.primal_obj - np.sum(C.multiply(Xh))) <= tol * np.abs(C).sum()
    # the tolerance still catches a dropped term: without the trace deficit
    # spread over the columns, z misses by delta rowsum(E^2), over 3x the bound
    assert np.abs(u.z - np.diag(X)).max() > 3 * tol


def test_dual_only_keeps_only_y():
    W = problems.erdos_renyi(120, 0.05, seed=8)
    W_old = scipy.sparse.csr_matrix(W)[:100, :100]
    old = _old_state(W_old, 4, 9)
    Om = np.random.default_rng(10).standard_normal((120, 4))
    new = omc.dual_only_warm_start_state(old, W, 4, Om)
    assert np.array_equal(new.y[:100], old.y) and not new.y[100:].any()
    assert not new.z.any() and not new.P.any() and new.tr_X == 0.0 and new.primal_obj == 0.0
    lib = maxcut_helpers.get_dual_only_warm_start_state(_to_pkg(old), W, 4, Omega=Om)
    for f in ("y", "z", "P", "Omega"):
        assert np.array_equal(np.asarray(getattr(lib, f)), np.asarray(getattr(new, f))), f
    assert lib.SCALE_C == new.SCALE_C and lib.SCALE_X == new.SCALE_X
