"""The USBS outer loop in FP64 NumPy (oracle code).

Restates solver/usbs.py:436-789.  ``q0`` is injected (the reference draws it
from jax.random with PRNGKey(0) and normalises it, usbs.py:691-692).  Every
condition check appends a record to ``log`` with the fields the reference
prints at usbs.py:514-530.
"""
import time
from collections import namedtuple

import numpy as np

from .common import SDPState
from .ipm import solve_step_subprob, compute_lb_spec_est_ipm
from .lanczos import eigsh_smallest
from .loop import while_loop
from .operators import apply_A_operator_batched, create_svec_matrix

State = namedtuple("State", ["t", "upsilon", "X", "tr_X", "X_bar", "tr_X_bar", "P", "P_bar", "z",
                             "z_bar", "y", "V", "start_time", "curr_time", "obj_gap",
                             "infeas_gap", "max_infeas", "primal_obj", "bar_primal_obj",
                             "pen_dual_obj", "cand_pen_dual_obj", "lb_spec_est", "obj_ub"])


def usbs(sdp_state, n, m, trace_ub, trace_factor, rho, beta, k_curr, k_past, max_iters, max_time,
         obj_gap_eps, infeas_gap_eps, max_infeas_eps, lanczos_inner_iterations,
         lanczos_max_restarts, subprob_eps, subprob_max_iters, cond_exp_base, callback_fn,
         callback_static_args, callback_nonstatic_args, q0, log=None, iter_hook=None):
    k = k_curr + k_past
    U = create_svec_matrix(k)
    SCALE_C, SCALE_X, SCALE_A = sdp_state.SCALE_C, sdp_state.SCALE_X, sdp_state.SCALE_A
    C, A_data, A_indices = sdp_state.C, sdp_state.A_data, np.asarray(sdp_state.A_indices)
    b, mask, Omega = sdp_state.b, sdp_state.b_ineq_mask, sdp_state.Omega
    negC = (-C).tocsr() if hasattr(C, "tocsr") else -C

    def cond(st):
        if Omega is not None and callback_fn is not None:
            cb = callback_fn(st.P, Omega, callback_static_args, callback_nonstatic_args)
        else:
            cb = None
        if log is not None:
            log.append(dict(t=st.t, curr_time=st.curr_time, pen_dual_obj=st.pen_dual_obj,
                            cand_pen_dual_obj=st.cand_pen_dual_obj, lb_spec_est=st.lb_spec_est,
                            primal_obj=st.primal_obj / (SCALE_C * SCALE_X),
                            dual_obj=float(np.dot(b, st.y)), obj_ub=st.obj_ub,
                            obj_gap=st.obj_gap, infeas_gap=st.infeas_gap,
                            max_infeas=st.max_infeas, callback_val=cb))
        return st.t == 0 or (st.curr_time - st.start_time < max_time
                             and (st.obj_gap > obj_gap_eps or st.infeas_gap > infeas_gap_eps
                                  or st.max_infeas > max_infeas_eps))

    def body(st):
        eta, S, ups_n = solve_step_subprob(C, A_data, A_indices, U, b, mask, st.upsilon, trace_ub,
                                           rho, st.bar_primal_obj, st.z_bar, st.tr_X_bar, st.y,
                                           st.V, n, m, k, subprob_eps, subprob_max_iters)
        lam, E = np.linalg.eigh(S)
        lam = np.clip(lam, 0, None)
        F = (st.V @ E) * np.sqrt(lam).reshape(1, -1)
        AF = apply_A_operator_batched(m, A_data, A_indices, F)
        if Omega is None and st.X is None:
            X_n, P_n = None, eta * st.P_bar + np.sum(F * F, axis=1)
        elif st.X_bar is not None:
            X_n, P_n = eta * st.X_bar + st.V @ S @ st.V.T, None
        else:
            X_n, P_n = None, eta * st.P_bar + F @ (F.T @ Omega)
        tr_X_n = eta * st.tr_X_bar + np.trace(S)
        z_n = eta * st.z_bar + AF
        y_cand = st.y - (1.0 / rho) * (b - z_n - ups_n)
        primal_n = eta * st.bar_primal_obj + np.trace(F.T @ (C @ F))

        cvals, cvecs = eigsh_smallest(n, negC, A_data, A_indices, y_cand, q0, k_curr,
                                      lanczos_inner_iterations, lanczos_max_restarts, subprob_eps)
        cvals = -cvals
        cand_pen = np.dot(b, y_cand) + trace_ub * max(cvals[0], 0.0)
        obj_ub = min(np.dot(b, y_cand) + (trace_ub / trace_factor) * max(cvals[0], 0.0), st.obj_ub)

        lb = compute_lb_spec_est_ipm(C, A_data, A_indices, U, b, trace_ub, st.bar_primal_obj,
                                     st.z_bar, st.tr_X_bar, y_cand, st.V, n, k, subprob_eps,
                                     subprob_max_iters)
        descent = beta * (st.pen_dual_obj - lb) <= st.pen_dual_obj - cand_pen
        y_n = y_cand if descent else st.y
        pen_n = cand_pen if descent else st.pen_dual_obj

        Fc = (st.V @ E[:, :k_curr]) * np.sqrt(lam[:k_curr]).reshape(1, -1)
        if Omega is None and st.X is None:
            Xb_n, Pb_n = None, eta * st.P_bar + np.sum(Fc * Fc, axis=1)
        elif st.X_bar is not None:
            Xb_n, Pb_n = eta * st.X_bar + Fc @ Fc.T, None
        else:
            Xb_n, Pb_n = None, eta * st.P_bar + Fc @ (Fc.T @ Omega)
        trXb_n = eta * st.tr_X_bar + np.sum(lam[:k_curr])
        zb_n = eta * st.z_bar + apply_A_operator_batched(m, A_data, A_indices, Fc)
        V_n, _ = np.linalg.qr(np.concatenate([st.V @ E[:, k_curr:], cvecs], axis=1))
        pb_n = eta * st.bar_primal_obj + np.trace(Fc.T @ (C @ Fc))

        obj_gap = np.abs(obj_ub - primal_n) / (SCALE_C * SCALE_X)
        obj_gap /= 1.0 + max(np.abs(primal_n), np.abs(obj_ub)) / (SCALE_C * SCALE_X)
        infeas = np.clip(z_n - b, 0.0, None) + (1.0 - mask) * np.clip(b - z_n, 0.0, None)
        infeas_gap = np.linalg.norm(infeas / SCALE_A) / SCALE_X
        infeas_gap /= 1.0 + np.linalg.norm((b / SCALE_A) / SCALE_X)
        max_infeas = np.max(np.abs(infeas) / SCALE_A) / SCALE_X
        st_n = State(st.t + 1, ups_n, X_n, tr_X_n, Xb_n, trXb_n, P_n, Pb_n, z_n, zb_n, y_n, V_n,
                     st.start_time, time.time(), obj_gap, infeas_gap, max_infeas, primal_n, pb_n,
                     pen_n, cand_pen, lb, obj_ub)
        if iter_hook is not None:
            iter_hook(st_n)
        return st_n

    ivals, ivecs = eigsh_smallest(n, negC, A_data, A_indices, sdp_state.y, q0, k,
                                  lanczos_inner_iterations, lanczos_max_restarts, subprob_eps)
    ivals = -ivals
    init_pen = np.dot(b, sdp_state.y) + trace_ub * max(ivals[0], 0.0)
    t0 = time.time()
    init = State(0, np.zeros(m), sdp_state.X, sdp_state.tr_X, sdp_state.X, sdp_state.tr_X,
                 sdp_state.P, sdp_state.P, sdp_state.z, sdp_state.z, sdp_state.y, ivecs, t0, t0,
                 1.1 * obj_gap_eps, 1.1 * infeas_gap_eps, 1.1 * max_infeas_eps,
                 sdp_state.primal_obj, sdp_state.primal_obj, init_pen, init_pen, 0.0, np.inf)
    final = while_loop(cond, body, init, max_iters, cond_exp_base=cond_exp_base)
    return sdp_state._replace(X=final.X, P=final.P, y=final.y, z=final.z, tr_X=final.tr_X,
                              primal_obj=final.primal_obj), final
