"""The CGAL loop in FP64 NumPy (oracle code).

Restates solver/cgal.py:21-243: conditional-gradient augmented Lagrangian
with a rank-one Frank-Wolfe step from the smallest eigenvector of
-C + A*(y + beta (z - proj_b)), a Nystrom sketch (or dense X) of the primal,
and the dual step clipped to [0, beta0].  ``q0`` is injected (the reference
draws it from jax.random with PRNGKey(0) and normalises it, cgal.py:108-109).
Every condition check appends a record with the fields the reference prints
at cgal.py:77-88.
"""
import time
from collections import namedtuple

import numpy as np

from .lanczos import eigsh_smallest
from .loop import while_loop
from .operators import apply_A_adjoint_slim, apply_A_operator_slim

State = namedtuple("State", ["t", "X", "P", "y", "z", "tr_X", "start_time", "curr_time",
                             "obj_gap", "infeas_gap", "max_infeas", "primal_obj"])


def cgal(sdp_state, n, m, trace_ub, beta0, max_iters, max_time, obj_gap_eps, infeas_gap_eps,
         max_infeas_eps, lanczos_inner_iterations, lanczos_max_restarts, subprob_eps,
         cond_exp_base, callback_fn, callback_static_args, callback_nonstatic_args, q0, log=None):
    SCALE_C, SCALE_X, SCALE_A = sdp_state.SCALE_C, sdp_state.SCALE_X, sdp_state.SCALE_A
    C, A_data, A_indices = sdp_state.C, sdp_state.A_data, np.asarray(sdp_state.A_indices)
    b, mask, Omega = sdp_state.b, sdp_state.b_ineq_mask, sdp_state.Omega
    negC = (-C).tocsr() if hasattr(C, "tocsr") else -C

    def cond(st):  # cgal.py:67-97
        if Omega is not None and callback_fn is not None:
            cb = callback_fn(st.P, Omega, callback_static_args, callback_nonstatic_args)
        else:
            cb = None
        if log is not None:
            log.append(dict(t=st.t, curr_time=st.curr_time,
                            primal_obj=st.primal_obj / (SCALE_C * SCALE_X), obj_gap=st.obj_gap,
                            infeas_gap=st.infeas_gap, max_infeas=st.max_infeas,
                            callback_val=cb))
        return st.t == 0 or (st.curr_time - st.start_time < max_time
                             and (st.obj_gap > obj_gap_eps or st.infeas_gap > infeas_gap_eps
                                  or st.max_infeas > max_infeas_eps))

    def body(st):  # cgal.py:99-185
        beta = beta0 * np.sqrt(st.t + 1)
        proj_b = (1 - mask) * b + mask * np.minimum(st.z + (1.0 / beta) * st.y, b)
        v = st.y + beta * (st.z - proj_b)
        ev, evec = eigsh_smallest(n, negC, A_data, A_indices, v, q0, 1, lanczos_inner_iterations,
                                  lanczos_max_restarts, subprob_eps)
        u = np.zeros(n) if ev[0] >= 0 else evec[:, 0].copy()
        uCu = float(u @ (C @ u))
        sgap = trace_ub * uCu - st.primal_obj + float(v @ st.z)
        sgap -= trace_ub * float(u @ apply_A_adjoint_slim(n, A_data, A_indices, v, u))
        obj_gap = sgap - float(st.y @ (st.z - proj_b))
        obj_gap -= 0.5 * beta * float(np.linalg.norm(st.z - proj_b)) ** 2
        obj_gap = obj_gap / (SCALE_C * SCALE_X)
        obj_gap /= 1.0 + abs(st.primal_obj) / (SCALE_C * SCALE_X)
        proj_b = (1 - mask) * b + mask * np.minimum(st.z, b)
        infeas_gap = np.linalg.norm((st.z - proj_b) / SCALE_A / SCALE_X)
        infeas_gap /= 1.0 + np.linalg.norm(proj_b / SCALE_A / SCALE_X)
        max_infeas = np.max(np.abs((st.z - proj_b) / SCALE_A)) / SCALE_X
        eta = 2.0 / (st.t + 2.0)
        if Omega is None:
            X_n, P_n = (1 - eta) * st.X + eta * trace_ub * np.outer(u, u), None
        else:
            X_n, P_n = None, (1 - eta) * st.P + eta * trace_ub * np.outer(u, u @ Omega)
        tr_X_n = (1 - eta) * st.tr_X + eta * trace_ub * float(u @ u)
        z_n = (1 - eta) * st.z + eta * trace_ub * apply_A_operator_slim(m, A_data, A_indices, u)
        beta = beta0 * np.sqrt(st.t + 2)
        proj_b = (1 - mask) * b + mask * np.minimum(z_n + (1.0 / beta) * st.y, b)
        step = np.clip(4 * beta * eta ** 2 * trace_ub ** 2 / np.sum(np.square(z_n - proj_b)),
                       0.0, beta0)
        y_n = st.y + step * (z_n - proj_b)
        p_n = (1 - eta) * st.primal_obj + eta * trace_ub * uCu
        return State(st.t + 1, X_n, P_n, y_n, z_n, tr_X_n, st.start_time, time.time(), obj_gap,
                     infeas_gap, max_infeas, p_n)

    t0 = time.time()
    init = State(0, sdp_state.X, sdp_state.P, sdp_state.y, sdp_state.z, sdp_state.tr_X, t0, t0,
                 1.1 * obj_gap_eps, 1.1 * infeas_gap_eps, 1.1 * max_infeas_eps,
                 sdp_state.primal_obj)
    final = while_loop(cond, body, init, max_iters, cond_exp_base=cond_exp_base)
    return sdp_state._replace(X=final.X, P=final.P, y=final.y, z=final.z, tr_X=final.tr_X,
                              primal_obj=final.primal_obj), final
