"""The lifted QAP SDP (utils/qap_helpers.py:157-381) solved on the GPU:
Lanczos on its operator against a dense eigensolver at the full l = 48 size,
and usbs() against the FP64 oracle loop."""
import numpy as np
import pytest
import scipy.sparse

from oracle import lanczos as olz, qap as oq, usbs as ousbs
from usbs_amd import problems, qap_helpers
from usbs_amd.lanczos import eigsh_smallest
from usbs_amd.usbs import usbs

pytestmark = pytest.mark.gpu

# warm_start_qap_usbs.py defaults, with trace_factor 2 (trace_ub = 2 after
# SCALE_X = 1/(l+1)) and a condition check every iteration
QAP = dict(trace_factor=2.0, rho=0.1, beta=0.25, k_curr=2, k_past=0, restarts=100,
           eps=1e-7, sub_iters=15)


def _state(l, seed, lib=True):
    D, W = problems.qap_instance(l, seed)
    N = l * l + 1
    Om = problems.gaussian_matrix(N, l, 1)
    if lib:
        s = qap_helpers.initialize_state(qap_helpers.build_objective_matrix(D, W), l, Omega=Om)
        return s._replace(C=s.C.to_scipy()), N
    return oq.initialize_state(D, W, l, Om)[0], N


def _operator(s, N, y):
    """Dense C + A*(y) of a scaled state."""
    A = s.A_indices
    M = scipy.sparse.coo_matrix((s.A_data * y[A[:, 0]], (A[:, 1], A[:, 2])), shape=(N, N))
    return (s.C + M).toarray()


def _run(s, N, iters, q0, lib, precision="fp64", exec_mode="perop", ctx=None):
    m = s.b.size
    c = QAP
    args = (s, N, m, c["trace_factor"], c["trace_factor"], c["rho"], c["beta"], c["k_curr"],
            c["k_past"], iters, 1e9, -1.0, -1.0, -1.0, min(N, 32), c["restarts"], c["eps"],
            c["sub_iters"], 1.0, None, None, None)
    log = []
    if lib:
        usbs(*args, precision=precision, exec_mode=exec_mode, q0=q0, log=log, context=ctx)
    else:
        ousbs.usbs(*args, q0, log=log)
    return log


@pytest.mark.parametrize("exec_mode", ["perop", "persistent"])
def test_qap_lanczos_full_size_matches_dense(gpu_ctx, exec_mode):
    """l = 48 (N = 2305, 2.0 M constraint triples, 1.54 M cost entries): the
    two smallest eigenvalues of C + A*(y) for a random dual y agree with
    numpy's dense eigensolver to 1e-9 relative."""
    s, N = _state(48, 0)
    y = np.random.default_rng(3).standard_normal(s.b.size)
    q0 = problems.unit_start_vector(N, 4)
    print("case: QAP l=48 eigsh nd=2 exec=%s" % exec_mode, flush=True)
    ev, V, info = eigsh_smallest(N, s.C, s.A_data, s.A_indices, y, q0, 2, 32, 100, 1e-7,
                                 exec_mode=exec_mode, return_info=True, context=gpu_ctx)
    dense = np.linalg.eigvalsh(_operator(s, N, y))
    np.testing.assert_allclose(ev[0], dense[0], rtol=1e-9)
    assert ev[1] >= dense[1] - 1e-9 * abs(dense[1])
    np.testing.assert_allclose(V.T @ V, np.eye(2), atol=1e-10)


def test_qap_lanczos_matches_oracle(gpu_ctx):
    """l = 8: the same Ritz values as the oracle's thick-restart Lanczos."""
    s, N = _state(8, 1)
    y = np.random.default_rng(5).standard_normal(s.b.size)
    q0 = problems.unit_start_vector(N, 6)
    ev_o, _ = olz.eigsh_smallest(N, s.C, s.A_data, s.A_indices, y, q0, 4, 32, 100, 1e-7)
    ev_g, _ = eigsh_smallest(N, s.C, s.A_data, s.A_indices, y, q0, 4, 32, 100, 1e-7,
                             context=gpu_ctx)
    np.testing.assert_allclose(ev_g, ev_o, rtol=1e-9, atol=1e-9 * np.abs(ev_o).max())


@pytest.mark.parametrize("l,exec_mode", [(5, "perop"), (5, "persistent"), (8, "perop")])
def test_qap_usbs_matches_oracle(gpu_ctx, l, exec_mode):
    """Small lifted QAPs, built by the oracle's own entry-by-entry builder
    for the oracle run and by the library's builder for the GPU run: the
    first iterations agree to 1e-6 relative, or to 20x the oracle's response
    to a rounding-level perturbation of q0 when that is larger."""
    so, N = _state(l, 0, lib=False)
    sg, _ = _state(l, 0, lib=True)
    q0 = problems.unit_start_vector(N, 2)
    q1 = q0 + 1e-15 * np.random.default_rng(9).standard_normal(N)
    print("case: QAP l=%d usbs 4 its exec=%s" % (l, exec_mode), flush=True)
    log_o = _run(so, N, 4, q0, False)
    log_p = _run(so, N, 4, q1 / np.linalg.norm(q1), False)
    log_g = _run(sg, N, 4, q0, True, exec_mode=exec_mode, ctx=gpu_ctx)
    assert len(log_g) == len(log_o) == 5
    for a, b, c in zip(log_o[1:], log_g[1:], log_p[1:]):
        for key in ("primal_obj", "obj_ub", "obj_gap", "infeas_gap", "max_infeas"):
            tol = max(1e-6 * abs(a[key]), 20.0 * abs(a[key] - c[key]), 1e-12)
            assert abs(b[key] - a[key]) <= tol, (key, int(a["t"]), b[key], a[key], tol)


def test_qap_usbs_full_size_matches_oracle(gpu_ctx):
    """l = 48 (the 2305-order lifted QAP): two FP64 iterations on the GPU
    against the oracle (same tolerance rule as above), and the mixed
    persistent path within 1e-4 of the FP64 one."""
    s, N = _state(48, 0)
    q0 = problems.unit_start_vector(N, 2)
    q1 = q0 + 1e-15 * np.random.default_rng(9).standard_normal(N)
    print("case: QAP l=48 usbs 2 its fp64 perop / mixed persistent", flush=True)
    log_g = _run(s, N, 2, q0, True, ctx=gpu_ctx)
    log_m = _run(s, N, 2, q0, True, precision="mixed", exec_mode="persistent", ctx=gpu_ctx)
    log_o = _run(s, N, 2, q0, False)
    log_p = _run(s, N, 2, q1 / np.linalg.norm(q1), False)
    for a, b, c, d in zip(log_o[1:], log_g[1:], log_p[1:], log_m[1:]):
        for key in ("primal_obj", "obj_ub", "obj_gap", "infeas_gap", "max_infeas"):
            tol = max(1e-6 * abs(a[key]), 20.0 * abs(a[key] - c[key]), 1e-12)
            assert abs(b[key] - a[key]) <= tol, (key, int(a["t"]), b[key], a[key], tol)
            assert abs(d[key] - b[key]) <= max(1e-4 * abs(b[key]), tol), (key, d[key], b[key])
