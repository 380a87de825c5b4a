"""usbs_batch (usbs_solve_batch): several solves in lockstep, sharing each
iteration's persistent Lanczos launches, give bitwise the results and records
of their own single persistent solves (solver/usbs.py:436-789 per item)."""
import numpy as np
import pytest

from usbs_amd import _lib, maxcut_helpers, problems
from usbs_amd.usbs import usbs, usbs_batch

pytestmark = pytest.mark.gpu

TIMES = ("time", "curr_time")


def _kw(W, ctx, seed, iters, precision, rho=1.0, gap_eps=-1.0, cond_exp_base=1.0, r=6):
    n = W.shape[0]
    st = maxcut_helpers.initialize_state(W, r, Omega=problems.gaussian_matrix(n, r, seed))
    return dict(sdp_state=st, n=n, m=n, trace_ub=2.0 * n, trace_factor=2.0, rho=rho, beta=0.25,
                k_curr=6, k_past=1, max_iters=iters, max_time=1e9, obj_gap_eps=gap_eps,
                infeas_gap_eps=gap_eps, max_infeas_eps=gap_eps, lanczos_inner_iterations=24,
                lanczos_max_restarts=10, subprob_eps=1e-7, subprob_max_iters=100,
                cond_exp_base=cond_exp_base, callback_fn=maxcut_helpers.compute_max_cut,
                callback_static_args=None, callback_nonstatic_args=st.C, precision=precision,
                q0=problems.unit_start_vector(n, seed + 7), context=ctx)


def _same(a, b):
    assert np.array_equal(a.y, b.y) and np.array_equal(a.z, b.z)
    assert np.array_equal(a.P, b.P)
    assert a.tr_X == b.tr_X and a.primal_obj == b.primal_obj


def _same_logs(la, lb):
    assert len(la) == len(lb)
    for ra, rb in zip(la, lb):
        for key, v in ra.items():
            if key in TIMES:
                continue
            assert np.array_equal(v, rb[key], equal_nan=True), (key, int(ra["t"]), v, rb[key])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_single_persistent(precision):
    """Four solves of different sizes and schedules: one stops at its first
    check (loose tolerances), one runs 12 iterations, one checks on a
    geometric schedule, one has max_iters = 0 (initial bundle only).  Each
    batched result and log equals its single persistent solve bitwise."""
    graphs = [problems.torus(20, 30, seed=1), problems.erdos_renyi(300, 0.06, seed=3),
              problems.torus(7, 9, seed=4), problems.erdos_renyi(150, 0.1, seed=6)]
    ctxs = [_lib.Context(0) for _ in graphs]
    try:
        specs = [dict(iters=8, gap_eps=1e12), dict(iters=12), dict(iters=15, cond_exp_base=1.5),
                 dict(iters=0)]
        print("case: batch of 4 (n = %s) %s" % ([W.shape[0] for W in graphs], precision), flush=True)
        calls = [_kw(W, ctx, 10 + i, precision=precision, **sp)
                 for i, (W, ctx, sp) in enumerate(zip(graphs, ctxs, specs))]
        logs_b = [[] for _ in calls]
        for kw, lg in zip(calls, logs_b):
            kw["log"] = lg
        out_b = usbs_batch(calls)
        assert [len(lg) for lg in logs_b] == [2, 13, 7, 1]
        for i, (W, ctx, sp) in enumerate(zip(graphs, ctxs, specs)):
            kw = _kw(W, ctx, 10 + i, precision=precision, **sp)
            log_s = []
            out_s = usbs(**kw, exec_mode="persistent", log=log_s)
            _same(out_b[i], out_s)
            _same_logs(logs_b[i], log_s)
    finally:
        for ctx in ctxs:
            ctx.close()


def test_batch_rejects_mismatched_items():
    """A shared context, differing bundle sizes or the per-op form are
    refused before anything runs, with the reason in every context."""
    W = problems.torus(6, 8, seed=0)
    ctxs = [_lib.Context(0), _lib.Context(0)]
    try:
        a = _kw(W, ctxs[0], 1, 2, "fp64")
        b = _kw(W, ctxs[1], 2, 2, "fp64")
        with pytest.raises(_lib.USBSError, match="twice"):
            usbs_batch([a, dict(a)])
        with pytest.raises(_lib.USBSError, match="k_curr"):
            usbs_batch([a, dict(b, k_curr=5)])
        with pytest.raises(_lib.USBSError, match="persistent"):
            usbs_batch([a, dict(b, exec_mode="perop")])
        # the contexts still solve afterwards
        out = usbs_batch([a, b])
        assert np.isfinite(out[0].primal_obj) and np.isfinite(out[1].primal_obj)
    finally:
        for ctx in ctxs:
            ctx.close()


def test_batch_torus_scale_equals_single():
    """Four torus 100 x 200 solves (the benchmark's size: 79 row tiles each,
    all workgroups of the batched launch co-resident), 150 iterations in
    lockstep: every record and result equals the single persistent solve
    bitwise.  The step kernel's device-wide barrier had a rare coherence race
    (another wavefront's stores missing the L2 writeback, or its loads beating
    the invalidation) that made such runs differ now and then."""
    graphs = [problems.torus(100, 200, seed=s) for s in range(4)]
    ctxs = [_lib.Context(0) for _ in graphs]
    try:
        print("case: batch of 4 torus 100x200, 150 iterations, mixed", flush=True)
        calls = [_kw(W, ctx, 20 + i, 150, "mixed", rho=3.0, r=10)
                 for i, (W, ctx) in enumerate(zip(graphs, ctxs))]
        logs_b = [[] for _ in calls]
        for kw, lg in zip(calls, logs_b):
            kw["log"] = lg
        out_b = usbs_batch(calls)
        for i, (W, ctx) in enumerate(zip(graphs, ctxs)):
            log_s = []
            out_s = usbs(**_kw(W, ctx, 20 + i, 150, "mixed", rho=3.0, r=10), exec_mode="persistent",
                         log=log_s)
            _same(out_b[i], out_s)
            _same_logs(logs_b[i], log_s)
    finally:
        for ctx in ctxs:
            ctx.close()
