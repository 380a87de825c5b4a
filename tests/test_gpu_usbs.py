"""End-to-end usbs() on the GPU against the FP64 oracle loop
(solver/usbs.py:436-789 restated in oracle/usbs.py)."""
import numpy as np
import pytest

from oracle import maxcut as omc, usbs as ousbs
from usbs_amd import maxcut_helpers, problems
from usbs_amd.usbs import usbs

pytestmark = pytest.mark.gpu

HP = dict(trace_factor=2.0, beta=0.25, max_time=1e9, obj_gap_eps=-1.0, infeas_gap_eps=-1.0,
          max_infeas_eps=-1.0, lanczos_max_restarts=10, subprob_eps=1e-7, subprob_max_iters=100,
          cond_exp_base=1.0)


def _run_both(W, iters, k_curr, k_past, r, rho, precision="fp64", exec_mode="perop", ctx=None):
    n = W.shape[0]
    Om = problems.gaussian_matrix(n, r, 1)
    q0 = problems.unit_start_vector(n, 2)
    so = omc.initialize_state(W, r, Om)
    log_o = []
    ousbs.usbs(so, n, n, 2.0 * n, HP["trace_factor"], rho, HP["beta"], k_curr, k_past, iters,
               HP["max_time"], -1.0, -1.0, -1.0, min(n, 32), 10, 1e-7, 100, 1.0,
               omc.compute_max_cut, None, so.C, q0, log=log_o)
    sg = maxcut_helpers.initialize_state(W, r, Omega=Om)
    log_g = []
    out = usbs(sg, n, n, 2.0 * n, HP["trace_factor"], rho, HP["beta"], k_curr, k_past, iters,
               HP["max_time"], -1.0, -1.0, -1.0, min(n, 32), 10, 1e-7, 100, 1.0,
               maxcut_helpers.compute_max_cut, None, sg.C, precision=precision, exec_mode=exec_mode,
               q0=q0, log=log_g, context=ctx)
    return log_o, log_g, out


@pytest.mark.parametrize("exec_mode", ["perop", "persistent"])
def test_usbs_fp64_trajectory_matches_oracle(gpu_ctx, exec_mode):
    """The first iterations agree with the oracle to 1e-5 relative.  Later
    iterates drift apart by design: the IPM's Newton systems have condition
    ~1/mu and stop at mu <= subprob_eps, so rounding-level differences grow to
    the 1e-7 tolerance level in one iteration and are then amplified by the
    prox dynamics (see the converged-value tests below)."""
    W = problems.erdos_renyi(120, 0.08, seed=11)
    print("case: ER(120,0.08) k=4+1 r=5 rho=1 exec=%s" % exec_mode, flush=True)
    log_o, log_g, _ = _run_both(W, 4, 4, 1, 5, 1.0, exec_mode=exec_mode, ctx=gpu_ctx)
    assert len(log_o) == len(log_g) == 5
    for a, b in zip(log_o[1:], log_g[1:]):
        for key in ("primal_obj", "obj_ub", "pen_dual_obj", "lb_spec_est", "obj_gap", "infeas_gap"):
            np.testing.assert_allclose(b[key], a[key], rtol=1e-5, atol=1e-9, err_msg=key)
        assert b["callback_val"] == a["callback_val"]


def test_usbs_reference_config_er200(gpu_ctx):
    """k_curr=10, k_past=1, sketch 10, rho=0.01 (the reference's max-cut
    settings): the first two iterations agree with the oracle to 1e-6
    relative, or to 20x the oracle's own response to a rounding-level
    perturbation of q0 when that is larger.  At rho = 0.01 the dual step
    (b - z)/rho amplifies rounding: a 1e-15 relative change of q0 moves the
    oracle's infeasibility gap by ~1e-5 at the second iteration."""
    W = problems.erdos_renyi(200, 0.06, seed=0)
    n = W.shape[0]
    log_o, log_g, _ = _run_both(W, 2, 10, 1, 10, 0.01, ctx=gpu_ctx)
    q0 = problems.unit_start_vector(n, 2)
    q1 = q0 + 1e-15 * np.random.default_rng(5).standard_normal(n)
    so = omc.initialize_state(W, 10, problems.gaussian_matrix(n, 10, 1))
    log_p = []
    ousbs.usbs(so, n, n, 2.0 * n, HP["trace_factor"], 0.01, HP["beta"], 10, 1, 2, HP["max_time"],
               -1.0, -1.0, -1.0, 32, 10, 1e-7, 100, 1.0, omc.compute_max_cut, None, so.C,
               q1 / np.linalg.norm(q1), log=log_p)
    for a, b, c in zip(log_o[1:], log_g[1:], log_p[1:]):
        for key in ("primal_obj", "obj_ub", "obj_gap", "infeas_gap", "max_infeas"):
            tol = max(1e-6 * abs(a[key]), 20.0 * abs(a[key] - c[key]), 1e-12)
            assert abs(b[key] - a[key]) <= tol, (key, int(a["t"]), b[key], a[key], tol)
        assert b["callback_val"] == a["callback_val"]


@pytest.mark.parametrize("name,W,value", [
    ("bipartite torus", problems.torus(6, 8, seed=0, signed=False), 96.0),
    ("K_8", problems.erdos_renyi(8, 1.0, seed=0), 16.0),
    ("C_5", problems.torus(1, 5, seed=0, signed=False), 2.5 * (1 + np.cos(np.pi / 5))),
])
@pytest.mark.parametrize("precision,exec_mode", [("fp64", "perop"), ("mixed", "persistent")])
def test_usbs_closed_forms(gpu_ctx, name, W, value, precision, exec_mode):
    """SDP values known in closed form, reached to 1e-4 relative."""
    n = W.shape[0]
    sg = maxcut_helpers.initialize_state(W, 4, omega_seed=0)
    out = usbs(sg, n, n, 2.0 * n, 2.0, 1.0, 0.25, 4, 1, 400, 1e9, 1e-7, 1e-7, 1e-7, min(n, 32), 50,
               1e-10, 100, 1.0, maxcut_helpers.compute_max_cut, None, sg.C, precision=precision,
               exec_mode=exec_mode, q0=problems.unit_start_vector(n, 1), context=gpu_ctx)
    assert abs(out.primal_obj - value) <= 1e-4 * value, (name, out.primal_obj, value)


def test_usbs_mixed_precision_close_to_fp64(gpu_ctx):
    W = problems.erdos_renyi(150, 0.08, seed=12, signed=True)
    log_o, log_g, _ = _run_both(W, 3, 4, 1, 5, 1.0, precision="mixed", exec_mode="persistent",
                                ctx=gpu_ctx)
    for a, b in zip(log_o[1:], log_g[1:]):
        np.testing.assert_allclose(b["primal_obj"], a["primal_obj"], rtol=1e-5)


def test_usbs_tune_form_same_results(gpu_ctx):
    """exec_mode "tune" times the first Lanczos call per-op and persistent and
    keeps the faster; both forms give bitwise the same Lanczos results, so the
    whole solve equals the per-op and the persistent solves exactly."""
    W = problems.erdos_renyi(150, 0.08, seed=12)
    outs = {}
    for mode in ("perop", "persistent", "tune"):
        print("case: ER(150,0.08) 3 its mixed exec=%s" % mode, flush=True)
        outs[mode] = _gpu_only(W, 3, 4, 1, 5, 1.0, "mixed", mode, gpu_ctx)
    base_log, base = outs["perop"]
    for mode in ("persistent", "tune"):
        log_g, out = outs[mode]
        assert np.array_equal(out.y, base.y) and np.array_equal(out.z, base.z), mode
        assert out.primal_obj == base.primal_obj, mode
        assert [r["obj_gap"] for r in log_g] == [r["obj_gap"] for r in base_log], mode


def _gpu_only(W, iters, k_curr, k_past, r, rho, precision, exec_mode, ctx):
    n = W.shape[0]
    Om = problems.gaussian_matrix(n, r, 1)
    q0 = problems.unit_start_vector(n, 2)
    sg = maxcut_helpers.initialize_state(W, r, Omega=Om)
    log_g = []
    out = usbs(sg, n, n, 2.0 * n, HP["trace_factor"], rho, HP["beta"], k_curr, k_past, iters,
               HP["max_time"], -1.0, -1.0, -1.0, min(n, 32), 10, 1e-7, 100, 1.0,
               maxcut_helpers.compute_max_cut, None, sg.C, precision=precision, exec_mode=exec_mode,
               q0=q0, log=log_g, context=ctx)
    return log_g, out


def test_usbs_reordered_nodes_same_sdp(gpu_ctx):
    """A solve of the RCM-renumbered problem is a solve of the same SDP: the
    first iterations' objective and dual values agree with the solve in the
    input numbering to rounding, and the sketch comes back in input order."""
    from usbs_amd.reorder import permute_state, unpermute_rows
    n = 300
    W = problems.random_geometric(n, 0.12, 3)
    Om = problems.gaussian_matrix(n, 5, 1)
    q0 = problems.unit_start_vector(n, 2)
    st = maxcut_helpers.initialize_state(W, 5, Omega=Om)
    st2, q02, perm = permute_state(st, q0)
    assert not np.array_equal(perm, np.arange(n))
    logs, outs = [], []
    for s, q in ((st, q0), (st2, q02)):
        print("case: RGG(300) 3 its fp64 perop, numbering %d" % len(outs), flush=True)
        log = []
        outs.append(usbs(s, n, n, 2.0 * n, HP["trace_factor"], 1.0, HP["beta"], 4, 1, 3,
                         HP["max_time"], -1.0, -1.0, -1.0, 32, 10, 1e-7, 100, 1.0,
                         maxcut_helpers.compute_max_cut, None, s.C, precision="fp64",
                         exec_mode="perop", q0=q, log=log, context=gpu_ctx))
        logs.append(log)
    for a, b in zip(logs[0][1:], logs[1][1:]):
        np.testing.assert_allclose(b["primal_obj"], a["primal_obj"], rtol=1e-8)
        np.testing.assert_allclose(b["dual_obj"], a["dual_obj"], rtol=1e-8)
    np.testing.assert_allclose(outs[1].y, outs[0].y, rtol=1e-7, atol=1e-9)
    P0, P1 = outs[0].P, unpermute_rows(outs[1].P, perm)
    # P accumulates Lanczos vectors converged to an absolute residual of 1e-7
    # (subprob_eps): a different node order changes their rounding, and with
    # it P, at that scale relative to |P|
    np.testing.assert_allclose(P1, P0, rtol=1e-6, atol=1e-7 * np.abs(P0).max())


def test_usbs_scaled_state_and_cut_scale(gpu_ctx):
    """A state scaled as the reference's warm starts scale it (SCALE_X = 1/n,
    SCALE_C = 1/|C|_F, utils/maxcut_helpers.py:134-135), with the rounding
    callback given C / SCALE_C as the reference's scripts pass it
    (scripts/warm_start_maxcut_usbs.py:121): the records (unscaled primal
    objective) follow the oracle, and every cut value is the oracle's exact
    integer cut of the unscaled graph."""
    from oracle import common as ocommon
    from usbs_amd import common
    W = problems.erdos_renyi(120, 0.08, seed=11)
    n = W.shape[0]
    Om = problems.gaussian_matrix(n, 5, 1)
    q0 = problems.unit_start_vector(n, 2)
    C0 = omc.laplacian_cost(W)
    sc, sx = 1.0 / np.linalg.norm(C0.data), 1.0 / n
    so = ocommon.scale_sdp_state(omc.initialize_state(W, 5, Om)._replace(SCALE_C=sc, SCALE_X=sx))
    sg = common.scale_sdp_state(maxcut_helpers.initialize_state(W, 5, Omega=Om)._replace(
        SCALE_C=sc, SCALE_X=sx))
    tu = 2.0 * n * sx
    log_o, log_g = [], []
    ousbs.usbs(so, n, n, tu, 2.0, 1.0, 0.25, 4, 1, 4, 1e9, -1.0, -1.0, -1.0, 32, 10, 1e-7, 100,
               1.0, omc.compute_max_cut, None, so.C / sc, q0, log=log_o)
    usbs(sg, n, n, tu, 2.0, 1.0, 0.25, 4, 1, 4, 1e9, -1.0, -1.0, -1.0, 32, 10, 1e-7, 100, 1.0,
         maxcut_helpers.compute_max_cut, None, sg.C / sc, q0=q0, log=log_g, context=gpu_ctx)
    assert len(log_o) == len(log_g) == 5
    for a, b in zip(log_o[1:], log_g[1:]):
        for key in ("primal_obj", "obj_ub", "obj_gap", "infeas_gap"):
            np.testing.assert_allclose(b[key], a[key], rtol=1e-5, atol=1e-9, err_msg=key)
        assert b["callback_val"] == a["callback_val"]
        assert b["callback_val"] == int(b["callback_val"]) and b["callback_val"] > 0


def test_usbs_geometric_condition_schedule(gpu_ctx):
    """cond_exp_base = 2 (the QAP scripts' default): condition checks, and so
    records and roundings, at t = 0, 1, 3, 7, 15 for max_iters = 15, as
    utils/loop.py:85-98 schedules them; the early records follow the oracle."""
    W = problems.erdos_renyi(120, 0.08, seed=11)
    n = W.shape[0]
    Om = problems.gaussian_matrix(n, 5, 1)
    q0 = problems.unit_start_vector(n, 2)
    so = omc.initialize_state(W, 5, Om)
    sg = maxcut_helpers.initialize_state(W, 5, Omega=Om)
    log_o, log_g = [], []
    ousbs.usbs(so, n, n, 2.0 * n, 2.0, 1.0, 0.25, 4, 1, 15, 1e9, -1.0, -1.0, -1.0, 32, 10, 1e-7,
               100, 2.0, omc.compute_max_cut, None, so.C, q0, log=log_o)
    out = usbs(sg, n, n, 2.0 * n, 2.0, 1.0, 0.25, 4, 1, 15, 1e9, -1.0, -1.0, -1.0, 32, 10, 1e-7,
               100, 2.0, maxcut_helpers.compute_max_cut, None, sg.C, q0=q0, log=log_g,
               context=gpu_ctx)
    assert [int(r["t"]) for r in log_o] == [0, 1, 3, 7, 15]
    assert [int(r["t"]) for r in log_g] == [0, 1, 3, 7, 15]
    for a, b in zip(log_o[1:3], log_g[1:3]):
        np.testing.assert_allclose(b["primal_obj"], a["primal_obj"], rtol=1e-5)
        assert b["callback_val"] == a["callback_val"]
    assert np.isfinite(out.primal_obj)


def test_usbs_prints_the_reference_log_line(gpu_ctx, capsys):
    """print_log=True prints one line per condition check in the format of
    solver/usbs.py:514-530, carrying the record's values."""
    import re
    W = problems.erdos_renyi(100, 0.08, seed=3)
    n = W.shape[0]
    sg = maxcut_helpers.initialize_state(W, 4, omega_seed=0)
    log = []
    usbs(sg, n, n, 2.0 * n, 2.0, 1.0, 0.25, 4, 1, 3, 1e9, -1.0, -1.0, -1.0, 32, 10, 1e-7, 100, 1.0,
         maxcut_helpers.compute_max_cut, None, sg.C, q0=problems.unit_start_vector(n, 1), log=log,
         context=gpu_ctx, print_log=True)
    lines = [l for l in capsys.readouterr().out.splitlines() if l.startswith("t: ")]
    assert len(lines) == len(log) == 4
    for line, rec in zip(lines, log):
        m = re.search(r"t:\s(\d+) - curr_time:\s([\.0-9]+).*primal_obj:\s([-e\.0-9]+).*"
                      r"callback_val:\s(.+)$", line)
        assert int(m.group(1)) == rec["t"] and float(m.group(3)) == rec["primal_obj"]
        assert float(m.group(4)) == rec["callback_val"]
        assert abs(float(m.group(2)) - rec["curr_time"]) < 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_usbs_dense_X_matches_oracle(gpu_ctx, precision):
    """Dense-X storage (sketch_dim = -1, solver/usbs.py:575-576, 631-632): the
    records and the final X (n x n) follow the oracle's, with the trajectory
    rule of the other tests (FP64; mixed within 1e-5 of it), and diag(X) is
    the constraint value z of the same iterate (max-cut: A(X) = diag X, up to
    the clipping of S's rounding-level negative eigenvalues in z)."""
    from oracle import common as ocommon
    W = problems.erdos_renyi(150, 0.06, seed=14, signed=True)
    n = W.shape[0]
    q0 = problems.unit_start_vector(n, 2)
    q1 = q0 + 1e-15 * np.random.default_rng(5).standard_normal(n)
    so = omc.initialize_state(W, -1)
    sg = maxcut_helpers.initialize_state(W, -1)
    assert sg.X.shape == (n, n) and sg.Omega is None

    def run_o(q):
        log = []
        out, _ = ousbs.usbs(so, n, n, 2.0 * n, 2.0, 1.0, 0.25, 4, 1, 4, 1e9, -1.0, -1.0, -1.0, 32,
                            10, 1e-7, 100, 1.0, None, None, None, q, log=log)
        return out, log

    (out_o, log_o), (out_p, _) = run_o(q0), run_o(q1 / np.linalg.norm(q1))
    log_g = []
    print("case: dense X, ER(150), 4 iterations, %s" % precision, flush=True)
    out_g = usbs(sg, n, n, 2.0 * n, 2.0, 1.0, 0.25, 4, 1, 4, 1e9, -1.0, -1.0, -1.0, 32, 10, 1e-7,
                 100, 1.0, None, None, None, precision=precision, q0=q0, log=log_g, context=gpu_ctx)
    assert out_g.P is None and out_g.X.shape == (n, n)
    rel = 1e-5 if precision == "mixed" else 1e-6
    for a, b in zip(log_o[1:], log_g[1:]):
        np.testing.assert_allclose(b["primal_obj"], a["primal_obj"], rtol=rel)
    scale = np.abs(out_o.X).max()
    # FP32 storage perturbs the bundle at 1e-7, which four prox steps amplify:
    # the mixed X is held to 1e-4 of the scale, as the other mixed-vs-FP64 checks
    tol = max((1e-4 if precision == "mixed" else 1e-6) * scale,
              20.0 * np.abs(out_o.X - out_p.X).max())
    assert np.abs(out_g.X - out_o.X).max() <= tol, (np.abs(out_g.X - out_o.X).max(), tol)
    assert abs(np.trace(out_g.X) - out_g.tr_X) <= 1e-9 * abs(out_g.tr_X)
    np.testing.assert_allclose(np.diag(out_g.X), out_g.z, rtol=1e-6, atol=1e-9)
