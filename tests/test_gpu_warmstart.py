"""Warm starts (utils/maxcut_helpers.py:95-289) on the GPU: the explicit
transform, whose Nystrom reconstruction and constraint map run on the
device, against the oracle's; and usbs() from each kind of warm-started
state against the oracle's loop from the oracle's state, with the trajectory
rule of tests/test_gpu_configs.py (1e-6 relative, or 20x the oracle's own
response to a 1e-15 perturbation of q0) and exact integer cuts."""
import numpy as np
import pytest
import scipy.sparse

from oracle import maxcut as omc, usbs as ousbs
from usbs_amd import maxcut_helpers, problems
from usbs_amd.sparse import BCOO
from usbs_amd.usbs import usbs

pytestmark = pytest.mark.gpu


def _exact_sketch_state(W, r, seed, n_old=None):
    """An old state on the first n_old vertices whose sketch is that of an
    exactly rank-r PSD matrix with a well-separated spectrum (so that the
    Nystrom factor is determined up to column signs)."""
    W_old = scipy.sparse.csr_matrix(W)[:n_old, :n_old] if n_old else W
    n = W_old.shape[0]
    rng = np.random.default_rng(seed)
    E0, _ = np.linalg.qr(rng.standard_normal((n, r)))
    X = (E0 * np.linspace(2.0 * n / r, 0.2 * n / r, r)) @ E0.T
    Om = rng.standard_normal((n, r))
    so = omc.initialize_state(W_old, r, Om)
    return so._replace(P=X @ Om, tr_X=float(n), y=rng.standard_normal(n) * 0.1,
                       z=np.diag(X).copy(), primal_obj=float(np.sum(so.C.multiply(X))))


def _pkg(s):
    return s._replace(C=BCOO.from_scipy_sparse(s.C))


def test_explicit_transform_matches_oracle(gpu_ctx):
    W = problems.erdos_renyi(600, 0.03, seed=21, signed=True)
    old = _exact_sketch_state(W, 6, 22, n_old=560)
    Om_new = np.random.default_rng(23).standard_normal((600, 6))
    print("case: explicit warm start n 560 -> 600, r 6", flush=True)
    o = omc.explicit_warm_start_state(old, W, 6, Om_new)
    g = maxcut_helpers.get_explicit_warm_start_state(_pkg(old), W, 6, Omega=Om_new,
                                                     context=gpu_ctx)
    for f in ("P", "z"):
        a, b = np.asarray(getattr(o, f)), np.asarray(getattr(g, f))
        np.testing.assert_allclose(b, a, rtol=0, atol=1e-9 * np.abs(a).max(), err_msg=f)
    for f in ("tr_X", "primal_obj", "SCALE_C", "SCALE_X"):
        assert np.isclose(float(getattr(g, f)), float(getattr(o, f)), rtol=1e-10), f
    assert np.array_equal(g.y, o.y) and np.array_equal(g.Omega, o.Omega)


def _solve_pair(so, sg, n, iters, ctx, q0):
    q1 = q0 + 1e-15 * np.random.default_rng(5).standard_normal(n)
    q1 /= np.linalg.norm(q1)
    tu = 2.0 * n * so.SCALE_X
    logs = []
    for st, q in ((so, q0), (so, q1)):
        log = []
        ousbs.usbs(st, n, n, tu, 2.0, 1.0, 0.25, 4, 1, iters, 1e9, -1.0, -1.0, -1.0, 32, 10, 1e-7,
                   100, 1.0, omc.compute_max_cut, None, st.C / st.SCALE_C, q, log=log)
        logs.append(log)
    log_g = []
    usbs(sg, n, n, 2.0 * n * sg.SCALE_X, 2.0, 1.0, 0.25, 4, 1, iters, 1e9, -1.0, -1.0, -1.0, 32, 10,
         1e-7, 100, 1.0, maxcut_helpers.compute_max_cut, None, sg.C / sg.SCALE_C, q0=q0, log=log_g,
         context=ctx)
    log_o, log_p = logs
    assert len(log_o) == len(log_g) == iters + 1
    for a, b, c in zip(log_o, log_g, log_p):
        for key in ("primal_obj", "obj_ub", "obj_gap", "infeas_gap", "max_infeas", "dual_obj"):
            if a["t"] == 0 and key in ("obj_ub", "obj_gap", "infeas_gap", "max_infeas"):
                continue  # the t = 0 record carries the initial placeholders
            tol = max(1e-6 * abs(a[key]), 20.0 * abs(a[key] - c[key]), 1e-12)
            assert abs(b[key] - a[key]) <= tol, (key, int(a["t"]), b[key], a[key], tol)
        # the oracle prices cuts in FP64 with C / SCALE_C, whose entries are the
        # graph's quarter-integer weights up to rounding; the library's cuts are
        # the exact integer sums, so the oracle's value is taken to the nearest 1/4
        a_cut, c_cut = round(4 * a["callback_val"]) / 4, round(4 * c["callback_val"]) / 4
        assert abs(a["callback_val"] - a_cut) <= 1e-9 * (1.0 + abs(a_cut))
        if a_cut == c_cut:
            assert b["callback_val"] == a_cut, (int(a["t"]), b["callback_val"], a["callback_val"])


@pytest.mark.parametrize("kind", ["implicit", "explicit", "dual_only"])
def test_usbs_from_warm_state_matches_oracle(gpu_ctx, kind):
    W = problems.erdos_renyi(300, 0.05, seed=31, signed=True)
    n, r = 300, 5
    old = _exact_sketch_state(W, r, 32, n_old=270)
    rng = np.random.default_rng(33)
    print("case: usbs from a %s warm start, n 270 -> 300, 3 iterations" % kind, flush=True)
    if kind == "implicit":
        pad = rng.standard_normal((30, r))
        so = omc.implicit_warm_start_state(old, W, r, pad)
        sg = maxcut_helpers.get_implicit_warm_start_state(_pkg(old), W, r, Omega_pad=pad)
    elif kind == "explicit":
        Om = rng.standard_normal((n, r))
        so = omc.explicit_warm_start_state(old, W, r, Om)
        sg = maxcut_helpers.get_explicit_warm_start_state(_pkg(old), W, r, Omega=Om, context=gpu_ctx)
    else:
        Om = rng.standard_normal((n, r))
        so = omc.dual_only_warm_start_state(old, W, r, Om)
        sg = maxcut_helpers.get_dual_only_warm_start_state(_pkg(old), W, r, Omega=Om)
    _solve_pair(so, sg, n, 3, gpu_ctx, problems.unit_start_vector(n, 34))
