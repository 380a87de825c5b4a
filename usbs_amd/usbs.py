"""usbs() with the reference's signature, run by the HIP library.

Mirrors solver/usbs.py:436-789.  The whole loop (prox-bundle IPM, sketch,
objective, Lanczos, lower bound, descent test, QR of the bundle, gaps,
condition checks and the max-cut rounding callback) runs on the device; the
host only reads one record per condition check.  Extra keyword arguments
after the reference's select the execution:

  backend   "hip" (the only backend in this package)
  precision "fp64" (plain FP64) or "mixed" (FP32 Lanczos / sketch storage,
            FP64 accumulation, Ritz solve, IPM, duals and convergence test)
  exec_mode "perop" (a kernel per Lanczos phase), "persistent" (one
            cooperative launch per Lanczos restart), "auto" (persistent
            where every row tile stays in LDS for a restart, else per-op) or
            "tune" (time the first Lanczos call in both forms, keep the
            faster; the results are the same either way)
  q0        Lanczos start vector (default: seeded normal, normalised; the
            reference draws it from jax.random with PRNGKey(0))
  log       a list that receives one dict per condition check (with
            curr_time, the wall clock of the check, as the reference prints)
  print_log print each condition check's line in the reference's format
            (solver/usbs.py:514-530; usbs_amd.loaders.format_log_line)

usbs_batch(calls) runs several such calls (one context each) in lockstep on
one device, sharing each iteration's persistent Lanczos launches.
"""
import ctypes
import time

import numpy as np
import scipy.sparse

from . import _lib, problems
from .common import SDPState
from .sparse import BCOO, coo_parts


class usbs_params(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int32) for name in
                ("k_curr", "k_past", "max_iters", "lanczos_inner_iterations",
                 "lanczos_max_restarts", "subprob_max_iters", "sketch_dim", "rounding",
                 "precision", "exec_mode", "warm", "reserved")] + \
               [(name, ctypes.c_double) for name in
                ("trace_ub", "trace_factor", "rho", "beta", "max_time", "obj_gap_eps",
                 "infeas_gap_eps", "max_infeas_eps", "subprob_eps", "cond_exp_base", "SCALE_C",
                 "SCALE_X", "cut_scale", "lanczos_rtol")]


RECORD_FIELDS = ("t", "time", "pen_dual_obj", "cand_pen_dual_obj", "lb_spec_est", "primal_obj",
                 "dual_obj", "obj_ub", "obj_gap", "infeas_gap", "max_infeas", "callback_val",
                 "ipm_iters", "lb_iters", "lanczos_restarts", "descent")


class usbs_record(ctypes.Structure):
    _fields_ = [(f, ctypes.c_double) for f in RECORD_FIELDS]


class usbs_batch_item(ctypes.Structure):
    _fields_ = [("ctx", _lib._vp), ("params", ctypes.POINTER(usbs_params)),
                ("y0", _lib._dp), ("z0", _lib._dp), ("tr_X0", ctypes.c_double),
                ("primal_obj0", ctypes.c_double), ("Omega", _lib._dp), ("P0", _lib._dp),
                ("q0", _lib._dp), ("y_out", _lib._dp), ("z_out", _lib._dp), ("P_out", _lib._dp),
                ("scal_out", _lib._dp), ("log", ctypes.POINTER(usbs_record)),
                ("log_cap", ctypes.c_int32), ("log_len", ctypes.POINTER(ctypes.c_int32))]


_lib.EXTRA_SIGNATURES["usbs_solve_batch"] = [ctypes.POINTER(usbs_batch_item), ctypes.c_int32]
_lib.EXTRA_SIGNATURES["usbs_solve"] = [
    _lib._vp, ctypes.POINTER(usbs_params), _lib._dp, _lib._dp, ctypes.c_double, ctypes.c_double,
    _lib._dp, _lib._dp, _lib._dp, _lib._dp, _lib._dp, _lib._dp, _lib._dp,
    ctypes.POINTER(usbs_record), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]


def _finite(x, default):
    """x as a finite double for the C ABI: +-inf become +-1e300 (so an eps of
    +inf places no requirement and -inf can never be met, as in the
    reference's comparisons); NaN becomes `default`."""
    x = float(x)
    if x != x:
        return default
    return max(-1e300, min(1e300, x))


def cut_scale_of(sdp_state, callback_nonstatic_args):
    """The factor s with sdp_state.C = s * callback_nonstatic_args, i.e. the
    cost the rounding callback prices cuts with (the reference's scripts pass
    C / SCALE_C, scripts/warm_start_maxcut_usbs.py:121).  None means SCALE_C.
    The native rounding works on the uploaded C, so any other matrix raises."""
    if callback_nonstatic_args is None:
        return float(sdp_state.SCALE_C)
    C, B = sdp_state.C, callback_nonstatic_args
    if B is C:
        return 1.0
    if isinstance(C, BCOO) and isinstance(B, BCOO) and C.shape == B.shape and \
            C.indices.shape == B.indices.shape and np.array_equal(C.indices, B.indices):
        # the same entries in the same order (a scaled copy of C, as the
        # reference's scripts pass): compare the value arrays
        amax = float(np.abs(C.data).max()) if C.data.size else 0.0
        bmax = float(np.abs(B.data).max()) if B.data.size else 0.0
        if amax == 0.0 or bmax == 0.0:
            return float(sdp_state.SCALE_C)
        s = amax / bmax
        if float(np.abs(C.data - s * B.data).max()) <= 1e-12 * amax:
            return float(s)
    r1, c1, v1, sh1 = coo_parts(C)
    r2, c2, v2, sh2 = coo_parts(B)
    if sh1 != sh2:
        raise ValueError("callback_nonstatic_args has shape %s, C has %s" % (sh2, sh1))
    A = scipy.sparse.csr_matrix((v1, (r1, c1)), shape=sh1)
    B = scipy.sparse.csr_matrix((v2, (r2, c2)), shape=sh2)
    A.sum_duplicates()
    B.sum_duplicates()
    amax = abs(A).max() if A.nnz else 0.0
    bmax = abs(B).max() if B.nnz else 0.0
    if amax == 0.0 or bmax == 0.0:
        return float(sdp_state.SCALE_C)
    s = amax / bmax
    if abs(A - s * B).max() > 1e-12 * amax:
        raise NotImplementedError("the native max-cut rounding prices cuts with the uploaded C; "
                                  "callback_nonstatic_args must be C / s for a constant s")
    return float(s)


def upload_problem(ctx, sdp_state, n, m):
    rows, cols, vals, shape = coo_parts(sdp_state.C)
    ctx.set_problem(n, m, rows, cols, vals, sdp_state.A_indices, sdp_state.A_data, sdp_state.b,
                    sdp_state.b_ineq_mask, np.broadcast_to(np.asarray(sdp_state.SCALE_A, float), (m,)))


class _Call:
    """One usbs() call with its arguments converted for the library: params,
    input arrays, output buffers and the record array."""

    def __init__(self, sdp_state, n, m, trace_ub, trace_factor, rho, beta, k_curr, k_past,
                 max_iters, max_time, obj_gap_eps, infeas_gap_eps, max_infeas_eps,
                 lanczos_inner_iterations, lanczos_max_restarts, subprob_eps, subprob_max_iters,
                 cond_exp_base, callback_fn, callback_static_args, callback_nonstatic_args,
                 backend="hip", precision="fp64", exec_mode="perop", q0=None, log=None, device=0,
                 context=None, upload=True, warm_start=False, print_log=False, lanczos_rtol=None):
        if backend != "hip":
            raise ValueError("backend=%r is not available in usbs_amd (only 'hip')" % (backend,))
        from .maxcut_helpers import compute_max_cut
        from . import qap_helpers
        self.state, self.n, self.log, self.print_log = sdp_state, n, log, print_log
        dense = self.dense = sdp_state.X is not None
        rounding = 0
        if callback_fn is not None and sdp_state.Omega is not None:
            if callback_fn is compute_max_cut:
                rounding = 1
            elif callback_fn is qap_helpers.qap_round:
                rounding = 2
            else:
                raise NotImplementedError("only the max-cut and QAP rounding callbacks run natively")
        self.own_ctx = context is None
        ctx = self.ctx = context if context is not None else _lib.Context(device)
        if upload:
            upload_problem(ctx, sdp_state, n, m)
        if rounding == 2:
            qap_helpers.set_rounding(ctx, callback_nonstatic_args)
        r = self.r = (-1 if dense else 0 if sdp_state.Omega is None
                      else int(np.asarray(sdp_state.Omega).shape[1]))
        self.prm = usbs_params(
            k_curr=k_curr, k_past=k_past, max_iters=int(max_iters),
            lanczos_inner_iterations=lanczos_inner_iterations,
            lanczos_max_restarts=lanczos_max_restarts, subprob_max_iters=subprob_max_iters,
            sketch_dim=r, rounding=rounding, precision=_lib.PRECISION[precision],
            exec_mode=_lib.EXEC_MODE[exec_mode], warm=1 if warm_start else 0, reserved=0,
            trace_ub=trace_ub, trace_factor=trace_factor, rho=rho, beta=beta,
            max_time=_finite(max_time, 1e300), obj_gap_eps=_finite(obj_gap_eps, -1e300),
            infeas_gap_eps=_finite(infeas_gap_eps, -1e300),
            max_infeas_eps=_finite(max_infeas_eps, -1e300), subprob_eps=subprob_eps,
            cond_exp_base=cond_exp_base, SCALE_C=float(sdp_state.SCALE_C),
            SCALE_X=float(sdp_state.SCALE_X),
            cut_scale=(cut_scale_of(sdp_state, callback_nonstatic_args) if rounding == 1 else 0.0),
            lanczos_rtol=0.0 if lanczos_rtol is None else float(lanczos_rtol))
        self.y0 = _lib.as_f64(sdp_state.y)
        self.z0 = _lib.as_f64(sdp_state.z)
        self.q0 = _lib.as_f64(problems.unit_start_vector(n, 0) if q0 is None else q0)
        self.Om = None if r <= 0 else _lib.as_f64(sdp_state.Omega)
        if dense:  # X0 in, X out (n x n) through the P arguments
            self.P0 = _lib.as_f64(sdp_state.X)
            if self.P0.shape != (n, n):
                raise ValueError("X must be n x n")
        else:
            self.P0 = None if sdp_state.P is None else _lib.as_f64(sdp_state.P)
        self.y_out, self.z_out = np.zeros(m), np.zeros(m)
        self.P_out = np.zeros((n, n)) if dense else np.zeros((n, r)) if r > 0 else np.zeros(n)
        self.scal = np.zeros(6)
        self.cap = 1 << 16
        self.recs = (usbs_record * self.cap)()
        self.nlog = ctypes.c_int(0)
        self.start_epoch = None

    def solve(self):
        self.start_epoch = time.time()
        st = self.state
        self.ctx.check(_lib.lib().usbs_solve(
            self.ctx.h, ctypes.byref(self.prm), _lib.dptr(self.y0), _lib.dptr(self.z0),
            float(st.tr_X), float(st.primal_obj), _lib.dptr(self.Om), _lib.dptr(self.P0),
            _lib.dptr(self.q0), _lib.dptr(self.y_out), _lib.dptr(self.z_out),
            _lib.dptr(self.P_out), _lib.dptr(self.scal), self.recs, self.cap,
            ctypes.byref(self.nlog)))

    def item(self):
        st = self.state
        return usbs_batch_item(
            ctx=self.ctx.h, params=ctypes.pointer(self.prm), y0=_lib.dptr(self.y0),
            z0=_lib.dptr(self.z0), tr_X0=float(st.tr_X), primal_obj0=float(st.primal_obj),
            Omega=_lib.dptr(self.Om), P0=_lib.dptr(self.P0), q0=_lib.dptr(self.q0),
            y_out=_lib.dptr(self.y_out), z_out=_lib.dptr(self.z_out),
            P_out=_lib.dptr(self.P_out), scal_out=_lib.dptr(self.scal), log=self.recs,
            log_cap=self.cap, log_len=ctypes.pointer(self.nlog))

    def result(self):
        records = []
        for i in range(min(self.nlog.value, self.cap)):
            rec = {f: getattr(self.recs[i], f) for f in RECORD_FIELDS}
            rec["curr_time"] = self.start_epoch + rec["time"]
            records.append(rec)
        if self.log is not None:
            self.log.extend(records)
        if self.print_log:
            from .loaders import format_log_line
            for rec in records:
                print(format_log_line(rec, self.start_epoch), flush=True)
        st, dense, r = self.state, self.dense, self.r
        P_ret = None if dense else self.P_out if st.P is not None or r > 0 else None
        return SDPState(C=st.C, A_indices=st.A_indices, A_data=st.A_data, b=st.b,
                        b_ineq_mask=st.b_ineq_mask, X=self.P_out if dense else None, P=P_ret,
                        Omega=st.Omega, y=self.y_out, z=self.z_out, tr_X=self.scal[0],
                        primal_obj=self.scal[1], SCALE_C=st.SCALE_C, SCALE_X=st.SCALE_X,
                        SCALE_A=st.SCALE_A)


def usbs(sdp_state, n, m, trace_ub, trace_factor, rho, beta, k_curr, k_past, max_iters, max_time,
         obj_gap_eps, infeas_gap_eps, max_infeas_eps, lanczos_inner_iterations,
         lanczos_max_restarts, subprob_eps, subprob_max_iters, cond_exp_base, callback_fn,
         callback_static_args, callback_nonstatic_args, backend="hip", precision="fp64",
         exec_mode="perop", q0=None, log=None, device=0, context=None, upload=True,
         warm_start=False, print_log=False, lanczos_rtol=None):
    """Run USBS on the GPU and return the final SDPState (scaled units).

    warm_start=True starts from the y, z and sketch P that ``context`` kept on
    the device from its previous solve (same sizes), instead of the state's
    y, z, P; tr_X and primal_obj still come from ``sdp_state`` (the scalars a
    warm-start transform recomputes, utils/maxcut_helpers.py:95-161)."""
    call = _Call(sdp_state, n, m, trace_ub, trace_factor, rho, beta, k_curr, k_past, max_iters,
                 max_time, obj_gap_eps, infeas_gap_eps, max_infeas_eps, lanczos_inner_iterations,
                 lanczos_max_restarts, subprob_eps, subprob_max_iters, cond_exp_base, callback_fn,
                 callback_static_args, callback_nonstatic_args, backend=backend,
                 precision=precision, exec_mode=exec_mode, q0=q0, log=log, device=device,
                 context=context, upload=upload, warm_start=warm_start, print_log=print_log,
                 lanczos_rtol=lanczos_rtol)
    call.solve()
    return call.result()


def usbs_batch(calls):
    """Several usbs() calls solved together on one device; returns their
    final SDPStates in order.  Each entry of ``calls`` is a dict of usbs()'s
    keyword arguments and must name its own ``context`` (distinct, one
    device).  The calls share precision, k_curr, k_past,
    lanczos_inner_iterations and lanczos_max_restarts, and run the persistent
    Lanczos form (exec_mode "persistent" or "auto"); each result and log is
    bitwise that of usbs(..., exec_mode="persistent") on its own (record
    times aside)."""
    prepared = []
    for kw in calls:
        if kw.get("context") is None:
            raise ValueError("each batched call needs its own context")
        kw = dict(kw)
        kw.setdefault("exec_mode", "persistent")
        prepared.append(_Call(**kw))
    items = (usbs_batch_item * len(prepared))(*[c.item() for c in prepared])
    t0 = time.time()
    for c in prepared:
        c.start_epoch = t0
    prepared[0].ctx.check(_lib.lib().usbs_solve_batch(items, len(prepared)))
    return [c.result() for c in prepared]
