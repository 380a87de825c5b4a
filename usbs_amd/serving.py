"""A request loop that serves max-cut SDP solves on one GPU.

Requests (a graph, an arrival time, optionally the id of an earlier request
whose graph it perturbs) are admitted continuously: whenever a solver slot is
free and the request's device-memory footprint fits the remaining budget,
the request starts at once, while others are still running.  Each slot is
its own context (device memory and HIP stream), so the kernels of concurrent
solves overlap on the device.  A follow-up request warm-starts from its
parent's final y, z and sketch P:

* when the parent ran on a free slot whose context still holds that state on
  the device, the follow-up runs there and the state never leaves the device;
* otherwise it starts from the host copy of the same arrays (the same values,
  so the result is identical).

The warm start is the reference's implicit warm start
(utils/maxcut_helpers.py:95-161, ``maxcut_helpers.get_implicit_warm_start_state``):
y, z and P carry over, the primal objective is corrected by the change of C's
diagonal, p' = p - <diag C_old, z> + <diag C_new, z>, and the follow-up is
solved in the scaling SCALE_X = 1/n, SCALE_C = 1/|C|_F (a cold request keeps
SCALE_C = SCALE_X = 1, as utils/maxcut_helpers.py:36-92).  On the device
path the library rescales the kept z and P from the parent's SCALE_X with the
same operations, so both paths give the same start bit for bit.

Every solve runs the same fixed work (``max_iters`` USBS iterations), so a
result does not depend on what else was on the device: the library's
reductions have a fixed order.  ``run`` returns per-request results and a
summary with p50 / p99 latency (arrival to completion) and solves per second.
"""
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib, maxcut_helpers, problems
from .usbs import usbs


@dataclass
class Request:
    rid: int
    W: object                 # scipy.sparse weighted adjacency (n x n)
    arrival: float            # seconds after the start of the run
    parent: int = -1          # rid of the request this one perturbs (warm start)


@dataclass
class Result:
    rid: int
    n: int
    slot: int
    warm: str                 # "cold", "device" or "host"
    arrival: float
    start: float
    end: float
    primal_obj: float
    obj_gap: float
    infeas_gap: float
    cut: float
    y: np.ndarray = field(repr=False, default=None)
    z: np.ndarray = field(repr=False, default=None)
    P: np.ndarray = field(repr=False, default=None)
    tr_X: float = 0.0
    state: object = field(repr=False, default=None)   # the solve's final SDPState

    @property
    def latency(self):
        return self.end - self.arrival


def footprint_bytes(n, nnz, k=5, r=10, m_lanczos=32, precision="mixed"):
    """Device memory of one usbs() solve of an n-vertex max-cut problem with
    nnz operator entries: the merged CSR (values in FP64 and, mixed, FP32),
    the Lanczos basis (n x (m+1)), bundle, sketch and work vectors, plus the
    per-tile partial sums; rounded up by 25 % for allocator slack."""
    s = 4 if precision == "mixed" else 8
    ld = (n + 63) // 64 * 64
    b = nnz * (4 + 8 + (4 if precision == "mixed" else 0) + 8 + 4)
    b += ld * s * ((m_lanczos + 1) + 2 + 3 * (k + 1) + k + 3 * r) + ld * 8 * (k + 8)
    b += 256 * 8 * (k * (k + 1) // 2) ** 2
    return int(1.25 * b)


class SolveServer:
    """``slots`` concurrent solver contexts on device ``device``; solve
    settings are usbs()'s (k_curr, k_past, rho, ..., max_iters)."""

    def __init__(self, slots=4, device=0, precision="mixed", exec_mode="perop", mem_budget=None,
                 k_curr=4, k_past=1, sketch_dim=10, trace_factor=2.0, rho=0.01, beta=0.25,
                 max_iters=20, lanczos_inner=32, lanczos_restarts=10, subprob_eps=1e-7,
                 subprob_max_iters=100, q0_seed=2):
        if slots > 1 and exec_mode != "perop":
            # a persistent (cooperative) launch is sized to the whole device; two
            # of them from concurrent slots could each wait for workgroups the
            # other holds, so the library admits one at a time and concurrent
            # slots use per-op launches
            raise ValueError("exec_mode=%r cannot run in %d concurrent slots: concurrent solves "
                             "use exec_mode='perop'" % (exec_mode, slots))
        self.ctxs = [_lib.Context(device) for _ in range(slots)]
        self.held = [-1] * slots          # rid whose final state each context keeps
        self.precision, self.exec_mode = precision, exec_mode
        self.cfg = dict(k_curr=k_curr, k_past=k_past, r=sketch_dim, tf=trace_factor, rho=rho,
                        beta=beta, iters=max_iters, mL=lanczos_inner, restarts=lanczos_restarts,
                        eps=subprob_eps, sub_iters=subprob_max_iters, q0_seed=q0_seed)
        if mem_budget is None:
            free, _ = self.ctxs[0].device_memory()
            mem_budget = int(0.6 * free)
        self.mem_budget = mem_budget

    def footprint(self, req):
        W = req.W
        return footprint_bytes(W.shape[0], W.nnz + W.shape[0], self.cfg["k_curr"] + self.cfg["k_past"],
                               self.cfg["r"], self.cfg["mL"], self.precision)

    # -- one solve --------------------------------------------------------
    def _state(self, req, parent_res):
        W = req.W
        n = W.shape[0]
        if parent_res is None:
            Om = problems.gaussian_matrix(n, self.cfg["r"], n)
            return maxcut_helpers.initialize_state(W, self.cfg["r"], Omega=Om)
        # a follow-up on a graph of the parent's size: no new rows of Omega
        return maxcut_helpers.get_implicit_warm_start_state(
            parent_res.state, W, self.cfg["r"], Omega_pad=np.zeros((0, self.cfg["r"])))

    def solve(self, req, slot, parent_res, t0):
        """Run one request on ``slot`` (blocking); returns its Result."""
        ctx = self.ctxs[slot]
        start = time.perf_counter() - t0
        st = self._state(req, parent_res)
        n = req.W.shape[0]
        warm = "cold"
        if parent_res is not None:
            warm = "device" if self.held[slot] == parent_res.rid else "host"
        c = self.cfg
        log = []
        out = usbs(st, n, n, c["tf"] * n * st.SCALE_X, c["tf"], c["rho"], c["beta"], c["k_curr"],
                   c["k_past"], c["iters"], float("inf"), -float("inf"), -float("inf"),
                   -float("inf"), c["mL"], c["restarts"], c["eps"], c["sub_iters"], 1.0,
                   maxcut_helpers.compute_max_cut, None, st.C / st.SCALE_C,
                   precision=self.precision, exec_mode=self.exec_mode,
                   q0=problems.unit_start_vector(n, c["q0_seed"]), log=log, context=ctx,
                   warm_start=(warm == "device"))
        self.held[slot] = req.rid
        end = time.perf_counter() - t0
        last = log[-1]
        return Result(req.rid, n, slot, warm, req.arrival, start, end, last["primal_obj"],
                      last["obj_gap"], last["infeas_gap"], last["callback_val"], y=out.y, z=out.z,
                      P=out.P, tr_X=out.tr_X, state=out)

    # -- the request loop -------------------------------------------------
    def run(self, requests, realtime=True):
        """Serve ``requests`` (arrival times in seconds from now).  With
        realtime=False every request is available at once (arrival 0)."""
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        pending = list(reqs)
        results = {}
        lock = threading.Condition()
        busy = [False] * len(self.ctxs)
        in_use = [0]
        t0 = time.perf_counter()
        errors = []

        def worker(req, slot, parent_res, fp):
            try:
                res = self.solve(req, slot, parent_res, t0)
            except Exception as e:  # surfaced after the loop
                errors.append(e)
                res = None
            with lock:
                if res is not None:
                    results[req.rid] = res
                busy[slot] = False
                in_use[0] -= fp
                lock.notify_all()

        threads = []
        with lock:
            while pending and not errors:
                now = time.perf_counter() - t0
                started = False
                for req in list(pending):
                    if realtime and req.arrival > now:
                        break
                    if req.parent >= 0 and req.parent not in results:
                        continue  # the parent is still running
                    fp = self.footprint(req)
                    if in_use[0] + fp > self.mem_budget and in_use[0] > 0:
                        continue
                    parent_res = results.get(req.parent) if req.parent >= 0 else None
                    free = [s for s in range(len(busy)) if not busy[s]]
                    if not free:
                        break
                    slot = free[0]
                    if parent_res is not None:
                        for s in free:
                            if self.held[s] == parent_res.rid:
                                slot = s
                                break
                    busy[slot] = True
                    in_use[0] += fp
                    pending.remove(req)
                    th = threading.Thread(target=worker, args=(req, slot, parent_res, fp))
                    th.start()
                    threads.append(th)
                    started = True
                if pending and not started:
                    nxt = min((r.arrival for r in pending), default=now)
                    wait = max(0.0, nxt - now) if realtime else 0.05
                    lock.wait(timeout=min(max(wait, 0.001), 0.05))
        for th in threads:
            th.join()
        if errors:
            raise errors[0]
        makespan = time.perf_counter() - t0
        out = [results[r.rid] for r in reqs]
        lat = np.array([r.latency for r in out])
        summary = dict(requests=len(out), makespan_s=makespan, solves_per_s=len(out) / makespan,
                       p50_latency_s=float(np.percentile(lat, 50)),
                       p99_latency_s=float(np.percentile(lat, 99)),
                       warm_device=sum(r.warm == "device" for r in out),
                       warm_host=sum(r.warm == "host" for r in out))
        return out, summary


def request_stream(count, rate, n_min=1000, n_max=100000, degree=10.0, follow_frac=0.3,
                   flip_frac=0.01, seed=0):
    """The serving workload: Poisson arrivals at ``rate`` per second, sizes
    log-uniform in [n_min, n_max], Erdos-Renyi graphs of mean degree
    ``degree`` with weights +-1; a ``follow_frac`` share of requests flips the
    signs of ``flip_frac`` of an earlier graph's edges and warm-starts."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for rid in range(count):
        t += rng.exponential(1.0 / rate)
        if reqs and rng.random() < follow_frac:
            parent = reqs[int(rng.integers(len(reqs)))]
            W = problems.flip_edge_signs(parent.W, flip_frac, int(rng.integers(1 << 31)))
            reqs.append(Request(rid, W, t, parent.rid))
        else:
            n = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
            W = problems.erdos_renyi(n, degree / (n - 1), seed=int(rng.integers(1 << 31)), signed=True)
            reqs.append(Request(rid, W, t))
    return reqs
