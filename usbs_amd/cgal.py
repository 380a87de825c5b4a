"""cgal() with the reference's signature, run by the HIP library.

Mirrors solver/cgal.py:21-243.  The whole loop (Lanczos eigen-oracle on
-C + A*(y + beta (z - proj_b)), rank-one sketch update, primal / dual
updates, gaps and the max-cut rounding callback) runs on the device through
``usbs_cgal_solve``.  Extra keyword arguments select the execution, as for
usbs(): backend ("hip"), precision ("fp64" | "mixed"), exec_mode ("perop" |
"persistent" | "auto"), q0 (the Lanczos start vector used at every iteration; the
reference draws it from jax.random with PRNGKey(0), cgal.py:108-109), log.
"""
import ctypes

import numpy as np

from . import _lib, problems
from .common import SDPState
from .usbs import RECORD_FIELDS, _finite, cut_scale_of, upload_problem, usbs_record


class usbs_cgal_params(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int32) for name in
                ("max_iters", "lanczos_inner_iterations", "lanczos_max_restarts", "sketch_dim",
                 "rounding", "precision", "exec_mode", "reserved")] + \
               [(name, ctypes.c_double) for name in
                ("trace_ub", "beta0", "max_time", "obj_gap_eps", "infeas_gap_eps",
                 "max_infeas_eps", "subprob_eps", "cond_exp_base", "SCALE_C", "SCALE_X",
                 "cut_scale")]


_lib.EXTRA_SIGNATURES["usbs_cgal_solve"] = [
    _lib._vp, ctypes.POINTER(usbs_cgal_params), _lib._dp, _lib._dp, ctypes.c_double,
    ctypes.c_double, _lib._dp, _lib._dp, _lib._dp, _lib._dp, _lib._dp, _lib._dp, _lib._dp,
    ctypes.POINTER(usbs_record), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]

CGAL_LOG_FIELDS = ("t", "time", "primal_obj", "obj_gap", "infeas_gap", "max_infeas",
                   "callback_val")


def cgal(sdp_state, n, m, trace_ub, beta0, max_iters, max_time, obj_gap_eps, infeas_gap_eps,
         max_infeas_eps, lanczos_inner_iterations, lanczos_max_restarts, subprob_eps,
         cond_exp_base, callback_fn, callback_static_args, callback_nonstatic_args,
         backend="hip", precision="fp64", exec_mode="perop", q0=None, log=None, device=0,
         context=None, upload=True):
    """Run CGAL on the GPU and return the final SDPState (scaled units)."""
    if backend != "hip":
        raise ValueError("backend=%r is not available in usbs_amd (only 'hip')" % (backend,))
    from .maxcut_helpers import compute_max_cut
    if sdp_state.X is not None or sdp_state.Omega is None:
        raise NotImplementedError("the HIP CGAL keeps a Nystrom sketch; dense-X storage is not "
                                  "provided")
    rounding = 0
    if callback_fn is not None:
        if callback_fn is not compute_max_cut:
            raise NotImplementedError("only the max-cut rounding callback runs natively")
        rounding = 1
    ctx = context if context is not None else _lib.Context(device)
    if upload:
        upload_problem(ctx, sdp_state, n, m)
    r = int(np.asarray(sdp_state.Omega).shape[1])
    prm = usbs_cgal_params(max_iters=int(max_iters), lanczos_inner_iterations=lanczos_inner_iterations,
                           lanczos_max_restarts=lanczos_max_restarts, sketch_dim=r,
                           rounding=rounding, precision=_lib.PRECISION[precision],
                           exec_mode=_lib.EXEC_MODE[exec_mode], reserved=0, trace_ub=trace_ub,
                           beta0=beta0, max_time=_finite(max_time, 1e300),
                           obj_gap_eps=_finite(obj_gap_eps, -1e300),
                           infeas_gap_eps=_finite(infeas_gap_eps, -1e300),
                           max_infeas_eps=_finite(max_infeas_eps, -1e300), subprob_eps=subprob_eps,
                           cond_exp_base=cond_exp_base, SCALE_C=float(sdp_state.SCALE_C),
                           SCALE_X=float(sdp_state.SCALE_X),
                           cut_scale=(cut_scale_of(sdp_state, callback_nonstatic_args) if rounding
                                      else 0.0))
    y0, z0 = _lib.as_f64(sdp_state.y), _lib.as_f64(sdp_state.z)
    if q0 is None:
        q0 = problems.unit_start_vector(n, 0)
    q0 = _lib.as_f64(q0)
    Om = _lib.as_f64(sdp_state.Omega)
    P0 = None if sdp_state.P is None else _lib.as_f64(sdp_state.P)
    y_out, z_out, P_out, scal = np.zeros(m), np.zeros(m), np.zeros((n, r)), np.zeros(6)
    cap = 1 << 16
    recs = (usbs_record * cap)()
    nlog = ctypes.c_int(0)
    ctx.check(_lib.lib().usbs_cgal_solve(ctx.h, ctypes.byref(prm), _lib.dptr(y0), _lib.dptr(z0),
                                         float(sdp_state.tr_X), float(sdp_state.primal_obj),
                                         _lib.dptr(Om), _lib.dptr(P0), _lib.dptr(q0),
                                         _lib.dptr(y_out), _lib.dptr(z_out), _lib.dptr(P_out),
                                         _lib.dptr(scal), recs, cap, ctypes.byref(nlog)))
    if log is not None:
        for i in range(min(nlog.value, cap)):
            log.append({f: getattr(recs[i], f) for f in CGAL_LOG_FIELDS})
    return SDPState(C=sdp_state.C, A_indices=sdp_state.A_indices, A_data=sdp_state.A_data,
                    b=sdp_state.b, b_ineq_mask=sdp_state.b_ineq_mask, X=None, P=P_out,
                    Omega=sdp_state.Omega, y=y_out, z=z_out, tr_X=scal[0], primal_obj=scal[1],
                    SCALE_C=sdp_state.SCALE_C, SCALE_X=sdp_state.SCALE_X, SCALE_A=sdp_state.SCALE_A)
