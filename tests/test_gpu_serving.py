"""The serving loop: requests admitted while others run, warm starts of
follow-up requests from device-resident state, and every result the same
alone and in the mix (bitwise: the library's reductions have a fixed order
and a warm start from the host copy of a state equals one from the device)."""
import numpy as np
import pytest

from usbs_amd import problems, serving

pytestmark = pytest.mark.gpu


def _mix():
    sizes = [300, 1200, 700, 1500, 500]
    reqs = [serving.Request(i, problems.erdos_renyi(n, 10.0 / (n - 1), seed=40 + i, signed=True),
                            0.0) for i, n in enumerate(sizes)]
    for rid, parent in ((5, 1), (6, 3), (7, 5)):
        W = problems.flip_edge_signs(reqs[parent].W, 0.01, 90 + rid)
        reqs.append(serving.Request(rid, W, 0.0, parent))
    return reqs


def test_serving_mix_equals_alone():
    reqs = _mix()
    print("case: serving mix of %d requests, 3 slots vs 1" % len(reqs), flush=True)
    mix, summary = serving.SolveServer(slots=3, max_iters=6).run(reqs, realtime=False)
    alone, _ = serving.SolveServer(slots=1, max_iters=6).run(reqs, realtime=False)
    assert summary["requests"] == len(reqs) and summary["p99_latency_s"] >= summary["p50_latency_s"]
    assert any(r.warm == "device" for r in mix + alone)
    for a, b in zip(mix, alone):
        assert a.rid == b.rid
        assert a.primal_obj == b.primal_obj and a.obj_gap == b.obj_gap and a.cut == b.cut
        assert np.array_equal(a.y, b.y) and np.array_equal(a.P, b.P)


def test_warm_start_continues_the_parent():
    """A follow-up on the same graph warm-started (the reference's implicit
    transform, which rescales to SCALE_X = 1/n, SCALE_C = 1/|C|_F) from the
    device-resident state begins where its parent ended: its first record
    (t = 0) carries the parent's final primal objective (to the rounding of
    the rescaling) and, rounded from the same sketch, the parent's final cut
    value."""
    from usbs_amd import maxcut_helpers
    from usbs_amd.usbs import usbs
    reqs = _mix()
    srv = serving.SolveServer(slots=1, max_iters=6)
    parent = srv.solve(reqs[1], 0, None, 0.0)
    W = reqs[1].W
    n = W.shape[0]
    st = srv._state(serving.Request(99, W, 0.0, 1), parent)
    assert st.SCALE_X == 1.0 / n
    log = []
    usbs(st, n, n, 2.0 * n * st.SCALE_X, 2.0, 0.01, 0.25, 4, 1, 2, float("inf"), -float("inf"),
         -float("inf"), -float("inf"), 32, 10, 1e-7, 100, 1.0, maxcut_helpers.compute_max_cut, None,
         st.C / st.SCALE_C, precision="mixed", exec_mode="perop", q0=np.ones(n) / np.sqrt(n),
         log=log, context=srv.ctxs[0], warm_start=True)
    assert abs(log[0]["primal_obj"] - parent.primal_obj) <= 1e-12 * abs(parent.primal_obj)
    assert log[0]["callback_val"] == parent.cut
