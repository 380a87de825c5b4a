#!/usr/bin/env python3
"""Benchmark of the USBS hot path on one MI355X.

Default workload: the max-cut SDP of a 2-D toroidal grid 100 x 200
(n = 20 000 vertices, 40 000 edges, weights uniform on {-1, +1} from seed 0)
solved by usbs() with k_curr = 10, k_past = 1, sketch_dim = 10,
trace_factor = 2, beta = 0.25, Lanczos 32 inner iterations x 10 restarts at
tolerance 1e-7, subproblem eps 1e-7 / 100 iterations.

The headline ("value") is the metric itself: seconds until the relative
objective gap (solver/usbs.py:646-647) first reaches 1e-5, for the
mixed-precision form, on a context whose problem already sits in device
memory, at rho = 3 (the torus does not reach 1e-5 within minutes at the
reference's rho = 0.01; rho = 3 reaches it in about 3 000 iterations).  The
same run reports:

* plain FP64 (per-op) from the identical start, to the same gap, and the ratio;
* a batch of torus problems (seeds 0..B-1) solved together (one context and
  host thread each, launches overlapping on the device) to a 1e-3 gap,
  against plain FP64 solving them one after another;
* p50 / p99 latency and solves/s of a short serving stream;
* throughput (USBS iterations/s) over K iterations at rho = 0.01, with the
  roofline of the Lanczos step kernels measured with HIP events on the
  solver's stream, and the gap both forms reach in those iterations;
* the FP64 CPU restatement (the oracle) on a bounded sample, with all host
  cores and with one.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "time-to-1e-5 rel gap: mixed/batched vs plain FP64 on GPU; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "lz_steps_coop_torus100x200_traffic.txt")
TRAFFIC_FILE_PEROP = os.path.join(ROOT, "profiles", "lz_perop_steps_torus100x200_traffic.txt")
TRAFFIC_FILE_RGG = os.path.join(ROOT, "profiles", "lz_perop_steps_rgg2e20_rcm_traffic.txt")
KRON_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_kron_bf16_qap48_8192vec_traffic.txt")


def progress(msg):
    """A line on stderr as each leg of a run ends (the JSON result goes to stdout)."""
    print("[bench] %s" % msg, file=sys.stderr, flush=True)


def traffic_from(path):
    """traffic_bytes_per_launch recorded in a counter summary, or None."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1])
    return None

SOLVER = dict(trace_factor=2.0, rho=0.01, beta=0.25, k_curr=10, k_past=1, lanczos_inner=32,
              lanczos_restarts=10, subprob_eps=1e-7, subprob_max_iters=100, sketch_dim=10)
# time-to-gap legs: the same settings at rho = 3
GAP_RHO = 3.0
GAP_TARGET = 1e-5
BATCH_GAP_TARGET = 1e-3
# The lifted QAP with the QAP script's settings (k_curr=2, k_past=0,
# sketch_dim=l, rho=0.1, 100 restarts, 15 subproblem iterations) and
# trace_factor 2 (trace_ub = 2 after SCALE_X = 1/(l+1)).
QAP_L = 48
QAP_SOLVER = dict(trace_factor=2.0, rho=0.1, beta=0.25, k_curr=2, k_past=0, lanczos_inner=32,
                  lanczos_restarts=100, subprob_eps=1e-7, subprob_max_iters=15, sketch_dim=QAP_L)


WORKLOADS = {
    "torus": "max-cut SDP, torus 100x200 (n=20000, 40000 edges, w in {-1,+1}, seed 0)",
    "rgg": "max-cut SDP, random geometric graph n=2^20 in the unit square, radius 1.3504e-3 "
           "(mean degree 6), unit weights, seed 0",
    "qap": "lifted QAP SDP, l=48 (X of order 2305): flow F symmetric integers 0-100 with 70% "
           "zeros, distance D from 48 random points on the 100x100 grid (rounded Euclidean, "
           "clipped at 100), seed 0",
}


def workload(name="torus", seed=0):
    from usbs_amd import problems
    if name == "rgg":
        W = problems.random_geometric(1 << 20, 1.3504e-3, 0)
    else:
        W = problems.torus(100, 200, seed=seed, signed=True)
    n = W.shape[0]
    Om = problems.gaussian_matrix(n, SOLVER["sketch_dim"], 1 + 10 * seed)
    q0 = problems.unit_start_vector(n, 2 + 10 * seed)
    return W, n, Om, q0


def make_problem(name, oracle=False, seed=0, reorder="none"):
    """The named workload as (state, n, m, q0, settings, trace_ub, callback):
    max-cut states come from the graph, the QAP state from its builder
    (library host code; ``oracle`` selects the oracle's max-cut builder and
    rounding for the CPU baseline)."""
    from usbs_amd import problems
    if name == "qap":
        from usbs_amd import qap_helpers
        l = QAP_L
        D, F = problems.qap_instance(l, 0)
        N = l * l + 1
        st = qap_helpers.initialize_state(qap_helpers.build_objective_matrix(D, F), l,
                                          Omega=problems.gaussian_matrix(N, l, 1))
        if oracle:
            st = st._replace(C=st.C.to_scipy())
        s = QAP_SOLVER
        return (st, N, st.b.size, problems.unit_start_vector(N, 2), s,
                s["trace_factor"] * (l + 1) * st.SCALE_X, None)
    W, n, Om, q0 = workload(name, seed)
    s = SOLVER
    if oracle:
        from oracle import maxcut as omc
        return (omc.initialize_state(W, s["sketch_dim"], Om), n, n, q0, s, s["trace_factor"] * n,
                omc.compute_max_cut)
    from usbs_amd import maxcut_helpers
    st = maxcut_helpers.initialize_state(W, s["sketch_dim"], Omega=Om)
    if reorder == "rcm":
        # node renumbering for SpMV locality (host setup, outside the timed region)
        from usbs_amd.reorder import permute_state
        st, q0, _ = permute_state(st, q0)
    return st, n, n, q0, s, s["trace_factor"] * n, maxcut_helpers.compute_max_cut


def run_gpu(ctx, prob, iters, precision, exec_mode, timing=False, rho=None, gap_eps=None,
            max_time=float("inf")):
    """usbs() on the resident problem: `iters` iterations, or until the
    objective gap reaches gap_eps (infeasibility not part of the stop test)."""
    from usbs_amd.usbs import usbs
    state, n, m, q0, s, tu, cb = prob
    log = []
    if timing:
        ctx.set_timing(True)
    ge = -float("inf") if gap_eps is None else gap_eps
    other = -float("inf") if gap_eps is None else float("inf")
    t0 = time.perf_counter()
    usbs(state, n, m, tu, s["trace_factor"], rho if rho is not None else s["rho"], s["beta"],
         s["k_curr"], s["k_past"], iters, max_time, ge, other, other,
         s["lanczos_inner"], s["lanczos_restarts"], s["subprob_eps"], s["subprob_max_iters"], 1.0,
         cb, None, state.C, precision=precision, exec_mode=exec_mode,
         q0=q0, log=log, context=ctx, upload=False)
    el = time.perf_counter() - t0
    stats = ctx.stats() if timing else None
    if timing:
        ctx.set_timing(False)
    return el, log, stats


def gap_marks(log, thresholds=(1e-1, 1e-2, 1e-3, 1e-4, 1e-5)):
    marks = {}
    for thr in thresholds:
        hit = [r for r in log if r["t"] > 0 and r["obj_gap"] <= thr]
        marks["%g" % thr] = ({"iteration": int(hit[0]["t"]), "seconds": round(hit[0]["time"], 4),
                              "infeas_gap": hit[0]["infeas_gap"]} if hit else None)
    return marks


def roofline(stats, rgg_rcm=False):
    """Algorithmic bytes per launch of the Lanczos step kernels (the
    persistent kernel, or one restart's per-op launches): per step the CSR
    SpMV nnz*(s+4) + 4(n+1) + 2 n s plus CGS2 over c valid columns
    (4c + 6) n s, with s the bytes per stored value."""
    L = stats["lz_launches"]
    if not L or stats["lz_kernel_ms"] <= 0:
        return None
    n, nnz, s = stats["n"], stats["nnz"], stats["value_bytes"]
    steps, sumc = stats["lz_steps"], stats["lz_sum_cols"]
    spmv = nnz * (s + 4) + 4 * (n + 1) + 2 * n * s
    total = steps * spmv + (4 * sumc + 6 * steps) * n * s
    per_launch = total / L
    avg_s = stats["lz_kernel_ms"] / L / 1e3
    achieved = per_launch / avg_s / 1e9
    # counters were taken on the torus workload (both forms) and on the RGG in
    # reverse Cuthill-McKee order (per-op), in FP32 storage; others report None
    persistent = stats.get("lz_form", 1) == 1
    traffic = None
    if (n, nnz, s) == (20000, 100000, 4):
        traffic = traffic_from(TRAFFIC_FILE if persistent else TRAFFIC_FILE_PEROP)
    elif rgg_rcm and n == (1 << 20) and s == 4 and not persistent:
        traffic = traffic_from(TRAFFIC_FILE_RGG)
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "kernel": ("k_lz_steps_coop (all Lanczos steps of one restart)" if persistent else
                       "per-op Lanczos step kernels k_lz_A/k_lz_BC of one restart "
                       "(launch gaps included)"),
            "avg_launch_us": round(avg_s * 1e6, 2), "launches": L,
            "steps_per_launch": round(steps / L, 3), "algorithmic_bytes_per_launch": int(per_launch)}


def cpu_baseline(iters, name="torus", iters_single=0):
    """The FP64 CPU restatement (oracle) on the same workload: `iters`
    iterations including the initial Lanczos call with BLAS on every core of
    this process's affinity, and `iters_single` iterations on one core."""
    import threadpoolctl
    from oracle import usbs as ousbs
    # every core this process may use, within the thread budget the host sets
    # (OMP_NUM_THREADS: the GPU's share of the host's cores)
    cores = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))

    def timed(n_iter, threads):
        with threadpoolctl.threadpool_limits(limits=threads):
            st, n, m, q0, s, tu, cb = make_problem(name, oracle=True)
            log = []
            t0 = time.perf_counter()
            ousbs.usbs(st, n, m, tu, s["trace_factor"], s["rho"], s["beta"],
                       s["k_curr"], s["k_past"], n_iter, 1e300, -1e300, -1e300, -1e300,
                       s["lanczos_inner"], s["lanczos_restarts"], s["subprob_eps"],
                       s["subprob_max_iters"], 1.0, cb, None, st.C, q0, log=log)
            return time.perf_counter() - t0

    el = timed(iters, cores)
    out = {"value": round(iters / el, 4), "unit": "USBS iterations/s", "cores": cores,
           "kind": "port", "host_cpus": os.cpu_count(),
           "affinity_cpus": len(os.sched_getaffinity(0)),
           "blas": [dict(api=d.get("internal_api"), threads=d.get("num_threads"))
                    for d in threadpoolctl.threadpool_info()],
           "sample": "FP64 NumPy/SciPy restatement, %s, %d iterations incl. the initial "
                     "Lanczos call (%.1f s), BLAS threads %d" % (name, iters, el, cores)}
    if iters_single > 0:
        el1 = timed(iters_single, 1)
        single = {"value": round(iters_single / el1, 4), "unit": "USBS iterations/s", "cores": 1,
                  "sample": "%d iterations (%.1f s)" % (iters_single, el1)}
        if single["value"] > out["value"]:
            # BLAS threads do not pay at this size: the baseline is the faster
            # single-core figure, the all-core one is kept beside it
            multi = {k: out[k] for k in ("value", "cores", "sample")}
            out.update(value=single["value"], cores=1,
                       sample="FP64 NumPy/SciPy restatement, %s, %s, BLAS threads 1"
                              % (name, single["sample"]))
            out["all_cores"] = multi
        else:
            out["single_thread"] = single
    return out


MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak


def kron_bench(args):
    """The lifted QAP's cost operator C = W (x) D (l = 48) on a batch of
    vectors: U_b = W Y_b D^T as BF16 matrix-core GEMMs with FP32 accumulation
    against the FP64 kernel, inputs resident in device memory."""
    from usbs_amd import _lib, problems, qap_helpers
    l, nvec, reps = QAP_L, args.kron_vectors, max(1, args.steps)
    D, F = problems.qap_instance(l, 0)
    Y = np.random.default_rng(0).standard_normal((nvec, l, l))
    ctx = _lib.Context(0)
    U16, ms16 = qap_helpers.kron_matvec(D, F, Y, "bf16", reps, ctx, return_time=True)
    U64, ms64 = qap_helpers.kron_matvec(D, F, Y, "fp64", reps, ctx, return_time=True)
    flops = 4.0 * l ** 3 * nvec
    bytes16 = nvec * l * l * (2 + 4)  # bf16 Y in, fp32 U out (W, D stay in LDS)
    achieved = bytes16 / (ms16 * 1e-3) / 1e9
    tflops = flops / (ms16 * 1e-3) / 1e12
    rel = float(np.linalg.norm(U16 - U64) / np.linalg.norm(U64))
    out = {"metric": METRIC, "value": round(nvec / (ms16 * 1e-3), 1), "unit": "vectors/s",
           "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": round(ms16, 5),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16 operands / f32 accumulate", "data": "synthetic (seeded F, D and vectors)",
           "config": {"workload": "QAP Kronecker matvec U = W Y D^T, l=48, batch of %d vectors "
                                  "(F, D as the l=48 lifted QAP)" % nvec,
                      "l": l, "vectors": nvec, "step": "one launch over the batch"},
           "plain_fp64": {"ms_per_step": round(ms64, 5), "speedup_of_value": round(ms64 / ms16, 3),
                          "rel_frobenius_diff": rel},
           "mfma": {"achieved": round(tflops, 2), "peak": MFMA_BF16_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 5)},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                        "traffic": traffic_from(KRON_TRAFFIC_FILE) if nvec == 8192 else None,
                        "kernel": "k_kron_bf16 (W Y D^T per vector)",
                        "avg_launch_us": round(ms16 * 1e3, 3),
                        "algorithmic_bytes_per_launch": int(bytes16)}}
    out["cpu_baseline"] = None
    if args.cpu_iters > 0:
        import threadpoolctl
        from oracle import qap as oq
        cores = min(16, len(os.sched_getaffinity(0)))
        rows, cols, vals = oq.build_objective_matrix(D.tolist(), F.tolist())
        import scipy.sparse
        N = l * l + 1
        C = scipy.sparse.csr_matrix((vals, (rows, cols)), shape=(N, N))
        ns = min(nvec, 512)
        V = np.zeros((N, ns))
        V[1:] = Y[:ns].reshape(ns, -1).T
        with threadpoolctl.threadpool_limits(limits=cores):
            t0 = time.perf_counter()
            it = 0
            while time.perf_counter() - t0 < 10.0:
                C @ V
                it += 1
            el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(it * ns / el, 1), "unit": "vectors/s", "cores": 1,
                               "kind": "port",
                               "sample": "the oracle's assembled C (scipy CSR, %d entries) times "
                                         "%d vectors, repeated for %.1f s" % (C.nnz, ns, el)}
    print(json.dumps(out), flush=True)


def batch_bench(args):
    """B torus 100x200 max-cut SDPs (seeds 0..B-1), K iterations each: the
    mixed-precision form with the B solves in flight together (a thread and
    a context each, so their launches overlap on the device) against the
    plain FP64 per-op form solving them one after another.  Also the mixed
    form one after another, which separates the gain of batching from the
    gain of precision."""
    import threading
    # a hardware queue per stream (two per solve in flight, and a few
    # spare); HIP reads this when it starts and nothing here has touched the
    # device yet
    os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(4, 2 * args.batch + 8)))
    from usbs_amd import _lib
    from usbs_amd.usbs import upload_problem
    B, K, W0 = args.batch, args.steps, args.warmup
    probs = [make_problem("torus", seed=b) for b in range(B)]

    def concurrent(precision, exec_mode):
        ctxs = [_lib.Context(0) for _ in range(B)]
        for ctx, pr in zip(ctxs, probs):
            upload_problem(ctx, pr[0], pr[1], pr[2])
        gate = threading.Barrier(B)
        spans, errs = [None] * B, []

        def work(b):
            try:
                if W0 > 0:
                    run_gpu(ctxs[b], probs[b], W0, precision, exec_mode)
                gate.wait()
                t0 = time.perf_counter()
                _, log, _ = run_gpu(ctxs[b], probs[b], K, precision, exec_mode)
                spans[b] = (t0, time.perf_counter(), int(log[-1]["t"]))
            except Exception as e:  # noqa: BLE001 - reported below
                errs.append(e)
                gate.abort()

        th = [threading.Thread(target=work, args=(b,)) for b in range(B)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        if errs:
            raise errs[0]
        for ctx in ctxs:
            ctx.close()
        wall = max(s[1] for s in spans) - min(s[0] for s in spans)
        return sum(s[2] for s in spans) / wall, wall

    def sequential(precision, exec_mode):
        ctx = _lib.Context(0)
        its, wall = 0, 0.0
        for pr in probs:
            upload_problem(ctx, pr[0], pr[1], pr[2])
            if W0 > 0:
                run_gpu(ctx, pr, W0, precision, exec_mode)
            el, log, _ = run_gpu(ctx, pr, K, precision, exec_mode)
            its += int(log[-1]["t"])
            wall += el
        ctx.close()
        return its / wall, wall

    # the solves in flight use the per-op form (what tune picks at this size);
    # persistent launches would compete for co-residency across the contexts
    rate_b, wall_b = concurrent("mixed", "perop")
    rate_s, wall_s = sequential("mixed", "perop")
    rate_64, wall_64 = sequential("fp64", "perop")
    out = {"metric": METRIC, "value": round(rate_b, 3), "unit": "USBS iterations/s (all solves)",
           "n_gpus": 1, "steps": B * K, "warmup": W0, "ms_per_step": round(1e3 / rate_b, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 storage / f64 accumulate",
           "data": "synthetic (seeded graphs, seeded Omega and q0)",
           "config": {"workload": "batch of %d max-cut SDPs, torus 100x200 (n=20000, w in {-1,+1}, "
                                  "seeds 0..%d), %d iterations each" % (B, B - 1, K),
                      "solves_in_flight": B, "precision": "mixed", "exec_mode": "perop",
                      **SOLVER},
           "mixed_sequential": {"value": round(rate_s, 3), "wall_s": round(wall_s, 3)},
           "plain_fp64": {"exec_mode": "perop", "order": "one solve after another",
                          "value": round(rate_64, 3), "wall_s": round(wall_64, 3),
                          "speedup_of_value": round(rate_b / rate_64, 3)},
           "wall_s": round(wall_b, 3), "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def serving_bench(args):
    """The serving workload: Poisson arrivals, sizes log-uniform in
    [1e3, 1e5], ER mean degree 10 with +-1 weights, 30 % follow-ups flipping
    1 % of an earlier graph's edge signs (warm-started)."""
    from usbs_amd import serving
    reqs = serving.request_stream(args.requests, args.rate, 1000, 100000, 10.0, 0.3, 0.01,
                                  seed=0)
    srv = serving.SolveServer(slots=args.slots, precision=args.precision, exec_mode="perop",
                              max_iters=args.solve_iters)
    # warm-up: one small solve per slot (module load, first kernel launches)
    warm = serving.request_stream(args.slots, 1e6, 1000, 1001, seed=1)
    srv.run(warm, realtime=False)
    srv.held = [-1] * args.slots
    results, summ = srv.run(reqs, realtime=True)
    out = {"metric": METRIC, "value": round(summ["solves_per_s"], 4), "unit": "solves/s",
           "n_gpus": 1, "steps": len(results), "warmup": args.slots,
           "ms_per_step": round(1e3 * summ["makespan_s"] / max(1, len(results)), 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 storage / f64 accumulate" if args.precision == "mixed" else "f64",
           "data": "synthetic (seeded Erdos-Renyi graphs, seeded arrivals)",
           "config": {"workload": "serving stream: Poisson %.2f/s, n log-uniform 1e3..1e5, ER "
                                  "degree 10, +-1 weights, 30%% follow-ups with 1%% sign flips"
                                  % args.rate,
                      "requests": len(results), "slots": args.slots,
                      "iterations_per_solve": args.solve_iters, "k_curr": 4, "k_past": 1,
                      "sketch_dim": 10},
           "p50_latency_s": round(summ["p50_latency_s"], 4),
           "p99_latency_s": round(summ["p99_latency_s"], 4),
           "warm_starts": {"device": summ["warm_device"], "host": summ["warm_host"]},
           "sizes": [r.n for r in results],
           "latencies_s": [round(r.latency, 4) for r in results],
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def gap_leg(ctx, prob, precision, exec_mode, target, max_iters, max_time):
    """One solve from the problem's start until the objective gap reaches
    `target` (or max_iters / max_time): the wall time of the usbs() call."""
    el, log, _ = run_gpu(ctx, prob, max_iters, precision, exec_mode, rho=GAP_RHO, gap_eps=target,
                         max_time=max_time)
    last = log[-1]
    reached = bool(last["obj_gap"] <= target)
    t = int(last["t"])
    return {"precision": precision, "exec_mode": exec_mode, "reached": reached,
            "seconds": round(el, 4) if reached else None, "wall_s": round(el, 4), "iterations": t,
            "ms_per_iteration": round(1e3 * el / max(t, 1), 4), "obj_gap": last["obj_gap"],
            "infeas_gap": last["infeas_gap"], "max_infeas": last["max_infeas"],
            "primal_obj": last["primal_obj"], "obj_ub": last["obj_ub"],
            "cut_value": None if last["callback_val"] != last["callback_val"] else last["callback_val"],
            "marks": gap_marks(log)}


def gap_batch_leg(B, target, max_iters, max_time):
    """B torus problems (seeds 0..B-1) to `target`: the mixed form with every
    solve in flight together (a context and host thread each), the mixed
    form in lockstep (usbs_solve_batch: one persistent Lanczos launch per
    restart for all solves), and plain FP64 per-op solving them one after
    another."""
    import threading
    from usbs_amd import _lib
    from usbs_amd.usbs import upload_problem
    probs = [make_problem("torus", seed=b) for b in range(B)]
    ctxs = [_lib.Context(0) for _ in range(B)]
    try:
        for ctx, pr in zip(ctxs, probs):
            upload_problem(ctx, pr[0], pr[1], pr[2])
            run_gpu(ctx, pr, 3, "mixed", "perop", rho=GAP_RHO)   # module load, first launches
            run_gpu(ctx, pr, 3, "fp64", "perop", rho=GAP_RHO)
        gate = threading.Barrier(B)
        res, errs = [None] * B, []

        def work(b):
            try:
                gate.wait()
                t0 = time.perf_counter()
                el, log, _ = run_gpu(ctxs[b], probs[b], max_iters, "mixed", "perop", rho=GAP_RHO,
                                     gap_eps=target, max_time=max_time)
                res[b] = (t0, time.perf_counter(), int(log[-1]["t"]), log[-1]["obj_gap"] <= target)
            except Exception as e:  # noqa: BLE001 - reported below
                errs.append(e)
                gate.abort()

        th = [threading.Thread(target=work, args=(b,)) for b in range(B)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        if errs:
            raise errs[0]
        wall_b = max(r[1] for r in res) - min(r[0] for r in res)
        progress("batched leg: %d mixed solves in flight, %.2f s" % (B, wall_b))
        lock = lockstep_leg(ctxs, probs, target, max_iters, max_time)
        progress("batched leg: %d mixed solves in lockstep, %.2f s" % (B, lock["wall_s"]))
        seq = []
        for b in range(B):
            el, log, _ = run_gpu(ctxs[b], probs[b], max_iters, "fp64", "perop", rho=GAP_RHO,
                                 gap_eps=target, max_time=max_time)
            seq.append((el, int(log[-1]["t"]), log[-1]["obj_gap"] <= target))
            progress("batched leg: fp64 solve %d of %d in %.2f s" % (b + 1, B, el))
    finally:
        for ctx in ctxs:
            ctx.close()
    wall_s = sum(x[0] for x in seq)
    return {"problems": B, "target_gap": target, "rho": GAP_RHO,
            "workload": "torus 100x200, w in {-1,+1}, seeds 0..%d" % (B - 1),
            "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
            "mixed_in_flight": {"wall_s": round(wall_b, 3), "iterations": [r[2] for r in res],
                                "all_reached": all(r[3] for r in res), "exec_mode": "perop"},
            "mixed_lockstep": lock,
            "plain_fp64_one_after_another": {"wall_s": round(wall_s, 3),
                                             "iterations": [x[1] for x in seq],
                                             "all_reached": all(x[2] for x in seq),
                                             "exec_mode": "perop"},
            "speedup": round(wall_s / wall_b, 3),
            "speedup_lockstep": round(wall_s / lock["wall_s"], 3)}


def lockstep_leg(ctxs, probs, target, max_iters, max_time):
    """usbs_batch over the resident problems (mixed, persistent form) until
    each reaches `target`; a short untimed batch first loads the modules."""
    from usbs_amd.usbs import usbs_batch

    def calls(iters, gap_eps, logs):
        out = []
        for ctx, (state, n, m, q0, s, tu, cb), log in zip(ctxs, probs, logs):
            other = -float("inf") if gap_eps is None else float("inf")
            out.append(dict(sdp_state=state, n=n, m=m, trace_ub=tu, trace_factor=s["trace_factor"],
                            rho=GAP_RHO, beta=s["beta"], k_curr=s["k_curr"], k_past=s["k_past"],
                            max_iters=iters, max_time=max_time,
                            obj_gap_eps=-float("inf") if gap_eps is None else gap_eps,
                            infeas_gap_eps=other, max_infeas_eps=other,
                            lanczos_inner_iterations=s["lanczos_inner"],
                            lanczos_max_restarts=s["lanczos_restarts"],
                            subprob_eps=s["subprob_eps"], subprob_max_iters=s["subprob_max_iters"],
                            cond_exp_base=1.0, callback_fn=cb, callback_static_args=None,
                            callback_nonstatic_args=state.C, precision="mixed",
                            exec_mode="persistent", q0=q0, log=log, context=ctx, upload=False))
        return out
    usbs_batch(calls(3, None, [[] for _ in probs]))
    logs = [[] for _ in probs]
    t0 = time.perf_counter()
    usbs_batch(calls(max_iters, target, logs))
    wall = time.perf_counter() - t0
    return {"wall_s": round(wall, 3), "iterations": [int(lg[-1]["t"]) for lg in logs],
            "all_reached": all(lg[-1]["obj_gap"] <= target for lg in logs),
            "exec_mode": "persistent (usbs_solve_batch)"}


def serving_leg(requests, rate, slots, iters):
    from usbs_amd import serving
    reqs = serving.request_stream(requests, rate, 1000, 100000, 10.0, 0.3, 0.01, seed=0)
    srv = serving.SolveServer(slots=slots, precision="mixed", exec_mode="perop", max_iters=iters)
    try:
        srv.run(serving.request_stream(slots, 1e6, 1000, 1001, seed=1), realtime=False)
        srv.held = [-1] * slots
        results, summ = srv.run(reqs, realtime=True)
    finally:
        for c in srv.ctxs:
            c.close()
    return {"requests": len(results), "rate_per_s": rate, "slots": slots,
            "iterations_per_solve": iters, "solves_per_s": round(summ["solves_per_s"], 4),
            "p50_latency_s": round(summ["p50_latency_s"], 4),
            "p99_latency_s": round(summ["p99_latency_s"], 4),
            "warm_starts": {"device": summ["warm_device"], "host": summ["warm_host"]},
            "sizes": [r.n for r in results]}


def gap_bench(args):
    """The default run: time to a 1e-5 objective gap (mixed vs plain FP64),
    the batched leg, the serving leg, throughput with the roofline, and the
    CPU baseline."""
    from usbs_amd import _lib
    from usbs_amd.usbs import upload_problem
    prob = make_problem("torus")
    state, n, m = prob[0], prob[1], prob[2]
    ctx = _lib.Context(0)
    try:
        upload_problem(ctx, state, n, m)
        # warm-up: module load, tune's choice of the Lanczos form, first launches
        if args.warmup > 0:
            run_gpu(ctx, prob, args.warmup, args.precision, args.exec_mode, rho=GAP_RHO)
            run_gpu(ctx, prob, args.warmup, "fp64", "perop", rho=GAP_RHO)
        progress("warm-up done")
        mixed = gap_leg(ctx, prob, args.precision, args.exec_mode, GAP_TARGET, args.gap_max_iters,
                        args.gap_max_time)
        progress("mixed to gap %g: %s s, %d iterations" % (GAP_TARGET, mixed["seconds"],
                                                           mixed["iterations"]))
        fp64 = gap_leg(ctx, prob, "fp64", "perop", GAP_TARGET, args.gap_max_iters,
                       args.gap_max_time)
        progress("fp64 to gap %g: %s s, %d iterations" % (GAP_TARGET, fp64["seconds"],
                                                          fp64["iterations"]))
        # throughput at the reference's rho = 0.01, Lanczos step kernels timed
        el, log, stats = run_gpu(ctx, prob, args.steps, args.precision, args.exec_mode,
                                 timing=(args.exec_mode != "perop"))
        el64, log64, _ = run_gpu(ctx, prob, args.steps, "fp64", "perop")
        progress("throughput legs done")
    finally:
        ctx.close()
    K, K64 = int(log[-1]["t"]), int(log64[-1]["t"])
    thr = {"rho": SOLVER["rho"], "steps": K, "value": round(K / el, 3), "unit": "USBS iterations/s",
           "ms_per_step": round(1e3 * el / max(K, 1), 4),
           "plain_fp64_ms_per_step": round(1e3 * el64 / max(K64, 1), 4),
           "obj_gap_end": log[-1]["obj_gap"], "plain_fp64_obj_gap_end": log64[-1]["obj_gap"],
           "marks": gap_marks(log), "plain_fp64_marks": gap_marks(log64)}
    batch = gap_batch_leg(args.gap_batch, BATCH_GAP_TARGET, args.gap_max_iters,
                          args.gap_max_time) if args.gap_batch > 0 else None
    progress("batched leg done")
    serve = serving_leg(args.requests, args.rate, args.slots, args.solve_iters) \
        if args.requests > 0 else None
    progress("serving leg done")
    t_m, t_f = mixed["seconds"], fp64["seconds"]
    out = {"metric": METRIC, "value": t_m, "unit": "s to 1e-5 relative objective gap",
           "n_gpus": 1, "steps": mixed["iterations"], "warmup": args.warmup,
           "ms_per_step": mixed["ms_per_iteration"], "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 storage / f64 accumulate" if args.precision == "mixed" else "f64",
           "data": "synthetic (seeded graph, seeded Omega and q0)",
           "config": {"workload": WORKLOADS["torus"], "n": n, "m": m, "precision": args.precision,
                      "exec_mode": args.exec_mode, "step": "one USBS iteration; value = wall time "
                      "of the usbs() call until obj_gap <= 1e-5 (solver/usbs.py:646-647)",
                      **dict(SOLVER, rho=GAP_RHO), "target_gap": GAP_TARGET},
           "time_to_gap": {"mixed": mixed, "plain_fp64": fp64,
                           "speedup": (round(t_f / t_m, 3) if t_m and t_f else None)},
           "batched": batch, "serving": serve, "throughput": thr}
    out["roofline"] = roofline(stats) if stats else None
    if args.cpu_iters > 0:
        cb = cpu_baseline(args.cpu_iters, "torus", args.cpu_iters_single)
        progress("cpu baseline done")
        # extrapolated: the mixed run's iteration count at the CPU's iteration rate
        cb["time_to_gap_estimate_s"] = round(mixed["iterations"] / cb["value"], 1)
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300, help="timed USBS iterations")
    ap.add_argument("--warmup", type=int, default=20, help="untimed USBS iterations")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--exec-mode", default="tune", choices=["tune", "auto", "persistent", "perop"],
                    help="tune: time the first Lanczos call in both forms and keep the faster; "
                         "auto: persistent where row tiles stay in LDS and fit one per CU, else per-op")
    ap.add_argument("--no-fp64-leg", action="store_true", help="skip the plain FP64 comparison")
    ap.add_argument("--throughput-only", action="store_true",
                    help="torus: only the throughput leg with its roofline (no time-to-gap, "
                         "batched or serving legs; the form profiled under rocprofv3)")
    ap.add_argument("--workload", default="torus",
                    choices=["torus", "rgg", "qap", "qap-kron", "serving", "torus-batch"],
                    help="torus 100x200 (default), RGG 2^20, QAP l=48, the QAP Kronecker "
                         "matvec alone, the serving stream, or a batch of torus solves in flight "
                         "together against plain FP64 one after another")
    ap.add_argument("--batch", type=int, default=16, help="torus-batch: solves in the batch")
    ap.add_argument("--reorder", default="auto", choices=["auto", "rcm", "none"],
                    help="node renumbering before upload: reverse Cuthill-McKee (auto: for the "
                         "randomly numbered RGG only)")
    ap.add_argument("--kron-vectors", type=int, default=8192,
                    help="qap-kron: vectors per launch")
    ap.add_argument("--requests", type=int, default=12,
                    help="serving: number of requests (0 skips the serving leg of the default run)")
    ap.add_argument("--rate", type=float, default=4.0, help="serving: arrivals per second")
    ap.add_argument("--slots", type=int, default=4, help="serving: concurrent solver slots")
    ap.add_argument("--solve-iters", type=int, default=10, help="serving: USBS iterations per solve")
    ap.add_argument("--cpu-iters", type=int, default=15, help="oracle iterations (0 = skip)")
    ap.add_argument("--cpu-iters-single", type=int, default=4,
                    help="oracle iterations on one core (torus default run)")
    ap.add_argument("--gap-max-iters", type=int, default=12000,
                    help="time-to-gap legs: iteration cap")
    ap.add_argument("--gap-max-time", type=float, default=60.0,
                    help="time-to-gap legs: seconds cap per solve")
    ap.add_argument("--gap-batch", type=int, default=8,
                    help="torus default run: problems in the batched time-to-gap leg (0 = skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 20 device-to-device copies of MIB MiB (the reference kernel "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy > 0:
        import torch
        nbytes = args.calibrate_copy << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(20):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "copies": 20}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py measures one GPU: the USBS hot path here runs one solve per GPU and "
              "no multi-GPU form is built (requested --gpus %d, WORLD_SIZE %d)"
              % (args.gpus, world), file=sys.stderr)
        sys.exit(2)

    if args.workload == "serving":
        serving_bench(args)
        return
    if args.workload == "torus-batch":
        batch_bench(args)
        return
    if args.workload == "qap-kron":
        kron_bench(args)
        return
    if args.workload == "torus" and not args.throughput_only:
        # 24 hardware queues (HIP's default is 4): the lockstep leg runs two
        # streams per solve and a batch stream (17 for eight solves), and
        # streams that share a queue serialise (eight solves in lockstep:
        # 235 it/s with 4 queues, 471 with 16, 519 with 24).  Set before the
        # first HIP call of the process.
        os.environ["GPU_MAX_HW_QUEUES"] = "24"
        gap_bench(args)
        return

    from usbs_amd import _lib
    from usbs_amd.usbs import upload_problem

    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except ImportError:
        sync = lambda: None  # noqa: E731

    reorder = args.reorder if args.reorder != "auto" else ("rcm" if args.workload == "rgg" else "none")
    prob = make_problem(args.workload, reorder=reorder)
    state, n, m = prob[0], prob[1], prob[2]
    ctx = _lib.Context(0)
    upload_problem(ctx, state, n, m)

    if args.warmup > 0:
        run_gpu(ctx, prob, args.warmup, args.precision, args.exec_mode)
    sync()
    el, log, stats = run_gpu(ctx, prob, args.steps, args.precision, args.exec_mode,
                             timing=(args.exec_mode != "perop"))
    sync()
    K = int(log[-1]["t"])
    ms = 1e3 * el / max(K, 1)
    out = {"metric": METRIC, "value": round(K / el, 3), "unit": "USBS iterations/s",
           "n_gpus": 1, "steps": K, "warmup": args.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 storage / f64 accumulate" if args.precision == "mixed" else "f64",
           "data": "synthetic (seeded graph, seeded Omega and q0)",
           "config": {"workload": WORKLOADS[args.workload], "n": n, "m": m,
                      "precision": args.precision, "exec_mode": args.exec_mode,
                      "node_order": "reverse Cuthill-McKee" if reorder == "rcm" else "as generated",
                      "step": "one USBS iteration", **prob[4]}}
    last = log[-1]
    out["gap"] = {"obj_gap_end": last["obj_gap"], "infeas_gap_end": last["infeas_gap"],
                  "max_infeas_end": last["max_infeas"],
                  "cut_value_end": (None if last["callback_val"] != last["callback_val"]
                                    else last["callback_val"]),
                  "time_to_gap": gap_marks(log), "target_1e-5_reached":
                      gap_marks(log)["1e-05"] is not None}
    if not args.no_fp64_leg:
        el64, log64, _ = run_gpu(ctx, prob, args.steps, "fp64", "perop")
        K64 = int(log64[-1]["t"])
        ms64 = 1e3 * el64 / max(K64, 1)
        out["plain_fp64"] = {"exec_mode": "perop", "ms_per_step": round(ms64, 4),
                             "speedup_of_value": round(ms64 / ms, 3),
                             "obj_gap_end": log64[-1]["obj_gap"],
                             "primal_rel_diff_end": abs(log64[-1]["primal_obj"] - last["primal_obj"])
                             / max(1.0, abs(log64[-1]["primal_obj"]))}
    out["roofline"] = (roofline(stats, rgg_rcm=(args.workload == "rgg" and reorder == "rcm"))
                       if stats else None)
    out["cpu_baseline"] = (cpu_baseline(args.cpu_iters, args.workload) if args.cpu_iters > 0
                           else None)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
