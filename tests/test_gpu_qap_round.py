"""QAP rounding (utils/qap_helpers.py:625-653) on the GPU: usbs_qap_round
(Nystrom factor on the device, Hungarian algorithm on the host) against the
oracle's restatement, and usbs() on lifted QAPs with the rounding callback at
every condition check against the oracle's loop with the oracle's rounding.
QAP objectives of integer W, D are exact integers in both, so a rounded value
either equals the oracle's or comes from a different assignment; a candidate
is required to equal the oracle's whenever the oracle's own value for it does
not change when P is perturbed by 1e-9 (relative), comparing each column's
pair of candidates (a Nystrom column's sign is arbitrary)."""
import numpy as np
import pytest

from oracle import qap as oq, usbs as ousbs
from usbs_amd import _lib, problems, qap_helpers
from usbs_amd.usbs import usbs

pytestmark = pytest.mark.gpu


def _oracle_fixed(P, Om, l, D, W):
    best, vals = oq.qap_round(P, Om, l, D, W)
    _, vals_p = oq.qap_round(P * (1.0 + 1e-9 * np.random.default_rng(3).standard_normal(P.shape)),
                             Om, l, D, W)
    return best, vals, vals == vals_p


def test_qap_round_matches_oracle(gpu_ctx):
    """A sketch of sum_k lambda_k x_k x_k^T over l lifted random assignments
    x_k with distinct weights: every Nystrom column is determined."""
    import ctypes
    l = 8
    N = l * l + 1
    D, W = problems.qap_instance(l, 5)
    rng = np.random.default_rng(6)
    X = np.zeros((N, N))
    for k, lam in enumerate(np.linspace(3.0, 0.5, l)):
        x = oq.lifted_assignment(rng.permutation(l))
        X += lam * np.outer(x, x)
    Om = rng.standard_normal((N, l))
    P = X @ Om
    print("case: qap_round l=%d, rank-%d sketch" % (l, l), flush=True)
    vals = np.zeros(2 * l)
    perm = np.zeros(l, dtype=np.int32)
    best = ctypes.c_double(0.0)
    gpu_ctx.check(_lib.lib().usbs_qap_round(gpu_ctx.h, l, l, _lib.dptr(np.ascontiguousarray(P)),
                                            _lib.dptr(Om), _lib.dptr(W), _lib.dptr(D),
                                            _lib.dptr(vals), _lib.iptr(perm), ctypes.byref(best)))
    best_o, vals_o, fixed = _oracle_fixed(P, Om, l, D, W)
    print("fixed candidates: %d of %d" % (fixed.sum(), fixed.size), flush=True)
    assert fixed.sum() >= l
    # a Nystrom column's sign is arbitrary (here and in the oracle), which
    # swaps its positive and negative candidates: compare each column's pair
    for i in range(l):
        if fixed[i] and fixed[l + i]:
            assert sorted([vals[i], vals[l + i]]) == sorted([vals_o[i], vals_o[l + i]]), i
    mp, mn = vals[:l].min(), vals[l:].min()
    assert best.value == (mp if mp < mn else mn)
    assert best.value == qap_helpers.qap_objective(D, W, perm)      # the returned assignment
    if fixed.all():
        assert best.value == best_o
    assert qap_helpers.qap_round(P, Om, None, {"W": W, "D": D}, context=gpu_ctx) == best.value


@pytest.mark.parametrize("l", [5, 8])
def test_qap_usbs_with_rounding_matches_oracle(gpu_ctx, l):
    D, W = problems.qap_instance(l, 0)
    N = l * l + 1
    Om = problems.gaussian_matrix(N, l, 1)
    so = oq.initialize_state(D, W, l, Om)[0]
    sg = qap_helpers.initialize_state(qap_helpers.build_objective_matrix(D, W), l, Omega=Om)
    sg = sg._replace(C=sg.C.to_scipy())
    q0 = problems.unit_start_vector(N, 2)
    q1 = q0 + 1e-15 * np.random.default_rng(9).standard_normal(N)
    m = so.b.size
    nonstatic = {"D": D, "W": W}

    def ocb(P, Omega, static, ns):
        return oq.qap_round(P, Omega, l, ns["D"], ns["W"])[0]

    def run_oracle(q):
        log = []
        ousbs.usbs(so, N, m, 2.0, 2.0, 0.1, 0.25, 2, 0, 4, 1e9, -1.0, -1.0, -1.0, min(N, 32), 100,
                   1e-7, 15, 1.0, ocb, None, nonstatic, q, log=log)
        return log

    print("case: QAP l=%d usbs 4 its with rounding" % l, flush=True)
    log_o, log_p = run_oracle(q0), run_oracle(q1 / np.linalg.norm(q1))
    log_g = []
    out = usbs(sg, N, m, 2.0, 2.0, 0.1, 0.25, 2, 0, 4, 1e9, -1.0, -1.0, -1.0, min(N, 32), 100, 1e-7,
               15, 1.0, qap_helpers.qap_round, None, nonstatic, q0=q0, log=log_g, context=gpu_ctx)
    assert len(log_o) == len(log_g) == 5
    for a, b, c in zip(log_o, log_g, log_p):
        if a["t"] > 0:
            for key in ("primal_obj", "obj_ub", "obj_gap", "infeas_gap"):
                tol = max(1e-6 * abs(a[key]), 20.0 * abs(a[key] - c[key]), 1e-12)
                assert abs(b[key] - a[key]) <= tol, (key, int(a["t"]), b[key], a[key], tol)
        # intermediate roundings: the sketches agree to rounding, but the
        # assignment on undetermined Nystrom columns can still differ; they are
        # reported here and checked at the sketch level below
        print("t %d: rounded QAP value gpu %g oracle %g (perturbed oracle %g)"
              % (a["t"], b["callback_val"], a["callback_val"], c["callback_val"]), flush=True)
    # the last record's value is the library's rounding of the solve's own
    # final sketch, and that rounding equals the oracle's on every column the
    # sketch fixes (as in test_qap_round_matches_oracle)
    import ctypes
    vals = np.zeros(2 * l)
    best = ctypes.c_double(0.0)
    P = np.ascontiguousarray(out.P)
    gpu_ctx.check(_lib.lib().usbs_qap_round(gpu_ctx.h, l, l, _lib.dptr(P), _lib.dptr(Om),
                                            _lib.dptr(W), _lib.dptr(D), _lib.dptr(vals), None,
                                            ctypes.byref(best)))
    assert log_g[-1]["callback_val"] == best.value
    _, vals_o, fixed = _oracle_fixed(P, Om, l, D, W)
    print("final sketch: %d of %d candidates fixed" % (fixed.sum(), fixed.size), flush=True)
    for i in range(l):
        if fixed[i] and fixed[l + i]:
            assert sorted([vals[i], vals[l + i]]) == sorted([vals_o[i], vals_o[l + i]]), i
