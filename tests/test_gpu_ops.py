"""Parity of the individual hot-path operations (HIP C-ABI entry points) with
the FP64 oracle: sketch update, objective/constraint map/dual candidate,
Nystrom reconstruction and max-cut rounding."""
import ctypes

import numpy as np
import pytest

from oracle import common as ocommon, maxcut as omc, operators as oops
from usbs_amd import _lib, problems
from usbs_amd.rounding import maxcut_round, reconstruct_from_sketch

pytestmark = pytest.mark.gpu

_lib.EXTRA_SIGNATURES.update({
    "usbs_sketch_update": [_lib._vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                           _lib._dp, _lib._dp, _lib._dp, ctypes.c_double, ctypes.c_int32, _lib._dp,
                           _lib._dp],
    "usbs_objective_infeas": [_lib._vp, ctypes.c_int32, ctypes.c_int32, _lib._dp, ctypes.c_double,
                              _lib._dp, _lib._dp, _lib._dp, ctypes.c_double, _lib._dp, _lib._dp,
                              _lib._dp, _lib._dp],
})


def _maxcut_ctx(ctx, W):
    C = omc.laplacian_cost(W).tocoo()
    n = C.shape[0]
    A_data, A_indices, b, mask = omc.get_all_problem_data(n)
    ctx.set_problem(n, n, C.row, C.col, C.data, A_indices, A_data, b, mask, np.ones(n))
    return C.tocsr(), A_data, A_indices, b, mask


@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 2e-6)])
def test_sketch_update(gpu_ctx, prec, tol):
    rng = np.random.default_rng(0)
    n, k, kc, r = 3001, 11, 10, 10
    F, Om, Pb = rng.standard_normal((n, k)), rng.standard_normal((n, r)), rng.standard_normal((n, r))
    eta = 0.37
    P = np.zeros((n, r)); Pb2 = np.zeros((n, r))
    gpu_ctx.check(_lib.lib().usbs_sketch_update(gpu_ctx.h, n, k, kc, r, _lib.dptr(F), _lib.dptr(Om),
                                                _lib.dptr(Pb), eta, prec, _lib.dptr(P), _lib.dptr(Pb2)))
    # solver/usbs.py:580 and :636 as restated by the oracle loop
    P_ref = eta * Pb + F @ (F.T @ Om)
    Pb_ref = eta * Pb + F[:, :kc] @ (F[:, :kc].T @ Om)
    scale = np.abs(P_ref).max()
    assert np.abs(P - P_ref).max() <= tol * scale
    assert np.abs(Pb2 - Pb_ref).max() <= tol * scale


def test_sketch_update_diag_only(gpu_ctx):
    rng = np.random.default_rng(1)
    n, k, kc = 777, 5, 3
    F, Pb = rng.standard_normal((n, k)), rng.standard_normal(n)
    P = np.zeros(n); Pb2 = np.zeros(n)
    gpu_ctx.check(_lib.lib().usbs_sketch_update(gpu_ctx.h, n, k, kc, 0, _lib.dptr(F), None,
                                                _lib.dptr(Pb), 0.5, 0, _lib.dptr(P), _lib.dptr(Pb2)))
    np.testing.assert_allclose(P, 0.5 * Pb + np.sum(F * F, 1), rtol=1e-13)
    np.testing.assert_allclose(Pb2, 0.5 * Pb + np.sum(F[:, :kc] ** 2, 1), rtol=1e-13)


def test_objective_infeas(gpu_ctx):
    W = problems.erdos_renyi(500, 0.05, seed=2, signed=True)
    C, A_data, A_indices, b, mask = _maxcut_ctx(gpu_ctx, W)
    n = m = C.shape[0]
    rng = np.random.default_rng(3)
    k, kc, eta, rho = 11, 10, 0.8, 0.01
    F = rng.standard_normal((n, k)) * 0.1
    zbar, y = rng.random(m), rng.standard_normal(m)
    out = np.zeros(6); z = np.zeros(m); zbn = np.zeros(m); yc = np.zeros(m)
    gpu_ctx.check(_lib.lib().usbs_objective_infeas(gpu_ctx.h, k, kc, _lib.dptr(F), eta, _lib.dptr(zbar),
                                                   _lib.dptr(y), None, rho, _lib.dptr(z), _lib.dptr(zbn),
                                                   _lib.dptr(yc), _lib.dptr(out)))
    z_ref = eta * zbar + oops.apply_A_operator_batched(m, A_data, A_indices, F)
    zbn_ref = eta * zbar + oops.apply_A_operator_batched(m, A_data, A_indices, F[:, :kc])
    yc_ref = y - (1.0 / rho) * (b - z_ref)
    np.testing.assert_allclose(z, z_ref, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(zbn, zbn_ref, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(yc, yc_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(out[0], np.trace(F.T @ (C @ F)), rtol=1e-12)
    np.testing.assert_allclose(out[1], np.trace(F[:, :kc].T @ (C @ F[:, :kc])), rtol=1e-12)
    np.testing.assert_allclose(out[2], b @ yc_ref, rtol=1e-12)
    np.testing.assert_allclose(out[3], zbar @ yc_ref, rtol=1e-12)
    infeas = np.clip(z_ref - b, 0, None) + (1 - mask) * np.clip(b - z_ref, 0, None)
    np.testing.assert_allclose(out[4], np.sum(infeas ** 2), rtol=1e-12)
    np.testing.assert_allclose(out[5], np.max(np.abs(infeas)), rtol=1e-13)


def test_reconstruct_matches_oracle(gpu_ctx):
    rng = np.random.default_rng(4)
    n, r, rank = 2000, 10, 6
    G = rng.standard_normal((n, rank))
    Om = rng.standard_normal((n, r))
    P = G @ (G.T @ Om)
    E, lam = reconstruct_from_sketch(Om, P, context=gpu_ctx)
    E_o, lam_o = ocommon.reconstruct_from_sketch(Om, P)
    np.testing.assert_allclose(lam, lam_o, rtol=1e-7, atol=1e-6 * lam_o.max())
    for j in range(rank):  # well-separated part of the spectrum, up to sign
        s = np.sign(E[:, j] @ E_o[:, j])
        np.testing.assert_allclose(s * E[:, j], E_o[:, j], atol=1e-6)


def test_reconstruct_zero_sketch_is_nan(gpu_ctx):
    n, r = 100, 5
    Om = np.random.default_rng(5).standard_normal((n, r))
    E, lam = reconstruct_from_sketch(Om, np.zeros((n, r)), context=gpu_ctx)
    assert np.isnan(lam).all()


def test_rounding_bruteforce_small_graph(gpu_ctx):
    """A sketch of X = s s^T for the best cut s must round to that cut's value."""
    W = problems.erdos_renyi(14, 0.4, seed=6)
    cuts = omc.cut_values_bruteforce(W)
    best_mask = max(cuts, key=cuts.get)
    s = np.array([1.0 if (best_mask >> i) & 1 else -1.0 for i in range(14)])
    Om = np.random.default_rng(7).standard_normal((14, 3))
    P = (np.outer(s, s) + 0.01 * np.eye(14)) @ Om
    C = omc.laplacian_cost(W)
    val, vals, signs = maxcut_round(C, P, Om)
    assert val == cuts[best_mask]
    # every candidate value is the exact integer cut of its sign vector
    for j in range(3):
        mask = sum(1 << i for i in range(14) if signs[i, j] > 0)
        assert vals[j] == cuts[mask]


def test_rounding_matches_oracle(gpu_ctx):
    """A sketch of a rank-r PSD matrix with a well-separated spectrum: every
    Nystrom column is determined up to sign, so each GPU sign vector must
    equal the oracle's (up to a global flip) and every candidate cut value,
    and the best one, must equal the oracle's exactly."""
    W = problems.torus(30, 40, seed=8, signed=True)
    C = omc.laplacian_cost(W)
    n, r = C.shape[0], 10
    rng = np.random.default_rng(9)
    G, _ = np.linalg.qr(rng.standard_normal((n, r)))
    G = G * np.arange(r, 0, -1.0)[None, :]            # singular values 10, 9, ..., 1
    Om = rng.standard_normal((n, r))
    P = G @ (G.T @ Om)
    val, vals, signs = maxcut_round(C, P, Om)
    sg_o = omc.sign_vectors(P, Om)
    agree = np.abs((signs.astype(float) * sg_o).sum(0)) == n
    assert agree.all(), agree
    oracle_vals = np.sum(sg_o * (C @ sg_o), 0)
    assert np.array_equal(np.sort(vals), np.sort(oracle_vals))
    assert val == omc.compute_max_cut(P, Om, None, C)


def test_rounding_rank_deficient_sketch_exact_cuts(gpu_ctx):
    """P of rank 4 < r = 10: six Nystrom columns are noise and their signs are
    not determined by the sketch, so no sign agreement is asserted.  Every
    candidate value must still be the exact integer cut of its own sign
    vector, recomputed on the CPU, and the four determined columns must agree
    with the oracle's."""
    W = problems.torus(30, 40, seed=8, signed=True)
    C = omc.laplacian_cost(W)
    n = C.shape[0]
    rng = np.random.default_rng(9)
    G = rng.standard_normal((n, 4))
    Om = rng.standard_normal((n, 10))
    P = G @ (G.T @ Om)
    val, vals, signs = maxcut_round(C, P, Om)
    cpu_vals = np.sum(signs * (C @ signs.astype(float)), 0)
    assert np.array_equal(vals, cpu_vals)
    assert val == cpu_vals.max()
    sg_o = omc.sign_vectors(P, Om)
    agree = np.abs((signs.astype(float) * sg_o).sum(0)) == n
    assert agree[:4].all(), agree
