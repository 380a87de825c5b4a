"""cgal() on the GPU against the FP64 oracle loop (solver/cgal.py:21-243
restated in oracle/cgal.py), and the max-cut SDP closed forms."""
import numpy as np
import pytest

from oracle import cgal as ocg, maxcut as omc
from usbs_amd import maxcut_helpers, problems
from usbs_amd.cgal import cgal

pytestmark = pytest.mark.gpu


def _run_both(W, iters, r, beta0, precision="fp64", exec_mode="perop", ctx=None):
    n = W.shape[0]
    Om = problems.gaussian_matrix(n, r, 1)
    q0 = problems.unit_start_vector(n, 2)
    so = omc.initialize_state(W, r, Om)
    log_o = []
    ocg.cgal(so, n, n, float(n), beta0, iters, 1e9, -1.0, -1.0, -1.0, min(n, 32), 10, 1e-7, 1.0,
             omc.compute_max_cut, None, so.C, q0, log=log_o)
    sg = maxcut_helpers.initialize_state(W, r, Omega=Om)
    log_g = []
    out = cgal(sg, n, n, float(n), beta0, iters, 1e9, -1.0, -1.0, -1.0, min(n, 32), 10, 1e-7, 1.0,
               maxcut_helpers.compute_max_cut, None, sg.C, precision=precision,
               exec_mode=exec_mode, q0=q0, log=log_g, context=ctx)
    return log_o, log_g, out


@pytest.mark.parametrize("exec_mode", ["perop", "persistent"])
def test_cgal_fp64_trajectory_matches_oracle(gpu_ctx, exec_mode):
    """Objective, gaps and cut values of the first iterations agree with the
    oracle (FP64, rel 1e-6; the cut values exactly)."""
    W = problems.erdos_renyi(120, 0.08, seed=11)
    print("case: cgal ER(120,0.08) r=5 exec=%s" % exec_mode, flush=True)
    log_o, log_g, _ = _run_both(W, 8, 5, 1.0, exec_mode=exec_mode, ctx=gpu_ctx)
    assert len(log_o) == len(log_g) == 9
    for a, b in zip(log_o[1:], log_g[1:]):
        for key in ("primal_obj", "obj_gap", "infeas_gap", "max_infeas"):
            np.testing.assert_allclose(b[key], a[key], rtol=1e-6, atol=1e-12, err_msg=key)
        assert b["callback_val"] == a["callback_val"]


@pytest.mark.parametrize("name,W,value", [
    ("bipartite torus", problems.torus(6, 8, seed=0, signed=False), 96.0),
    ("K_8", problems.erdos_renyi(8, 1.0, seed=0), 16.0),
])
@pytest.mark.parametrize("precision,exec_mode", [("fp64", "perop"), ("mixed", "persistent")])
def test_cgal_closed_forms(gpu_ctx, name, W, value, precision, exec_mode):
    """SDP values known in closed form (bipartite |E|, K_n n^2/4) to 1e-3
    relative; CGAL is a first-order method, 3000 iterations."""
    n = W.shape[0]
    sg = maxcut_helpers.initialize_state(W, 4, omega_seed=0)
    out = cgal(sg, n, n, float(n), 1.0, 3000, 1e9, 1e-4, 1e-4, 1e-4, min(n, 32), 50, 1e-10, 1.0,
               None, None, None, precision=precision, exec_mode=exec_mode,
               q0=problems.unit_start_vector(n, 1), context=gpu_ctx)
    assert abs(out.primal_obj - value) <= 1e-3 * value, (name, out.primal_obj, value)


def test_cgal_mixed_precision_close_to_fp64(gpu_ctx):
    W = problems.erdos_renyi(150, 0.08, seed=12, signed=True)
    log_o, log_g, _ = _run_both(W, 5, 4, 1.0, precision="mixed", exec_mode="persistent",
                                ctx=gpu_ctx)
    for a, b in zip(log_o[1:], log_g[1:]):
        np.testing.assert_allclose(b["primal_obj"], a["primal_obj"], rtol=1e-5)
