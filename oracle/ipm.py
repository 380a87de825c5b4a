"""Prox-bundle IPM subproblems of USBS in FP64 NumPy (oracle code).

Restates solver/usbs.py:25-433.  The line search takes the largest 0.9^j
(j < ipm_max_iters) keeping S and T positive definite, exactly as the
reference's vmapped ``eigvalsh(...)[0] > 0`` test does (usbs.py:135-154).
"""
from collections import namedtuple

import numpy as np

from .operators import (apply_A_adjoint_batched, apply_A_operator_batched, create_Q_base)
from .loop import while_loop


def _min_eig_positive_steps(M, dM, steps):
    """eigvalsh(M + s dM)[0] > 0 for every candidate step s (batched)."""
    return np.linalg.eigvalsh(M[None, :, :] + steps[:, None, None] * dM[None, :, :])[:, 0] > 0


def quad_subprob_data(C, A_data, A_indices, U, b, upsilon, trace_ub, rho, bar_primal_obj, z_bar,
                      tr_X_bar, y, V, n, m, k):
    """Constants of the prox QP (solver/usbs.py:53-61): the subproblem is
    min 1/2 [s; eta]^T [[Q11, q12], [q12^T, q22]] [s; eta] + h1^T s + h2 eta
    over s = svec(S), S >= 0, eta >= 0, trace(S) + eta <= 1."""
    svec = lambda mx: U @ mx.reshape(-1)
    tr_X_bar = tr_X_bar if tr_X_bar > 0.0 else trace_ub
    Q_11 = (trace_ub ** 2 / rho) * create_Q_base(m, k, U, A_data, A_indices, V)
    q_12 = (trace_ub ** 2 / (rho * tr_X_bar)) * svec(
        V.T @ apply_A_adjoint_batched(n, A_data, A_indices, z_bar, V))
    q_22 = (trace_ub ** 2 / (rho * tr_X_bar ** 2)) * np.dot(z_bar, z_bar)
    h_1 = trace_ub * svec(V.T @ (apply_A_adjoint_batched(n, A_data, A_indices, y, V)
                                 - C @ V
                                 - apply_A_adjoint_batched(n, A_data, A_indices, b - upsilon, V) / rho))
    h_2 = (trace_ub / tr_X_bar) * (np.dot(z_bar, y) - bar_primal_obj - (np.dot(z_bar, b - upsilon) / rho))
    return Q_11, q_12, q_22, h_1, h_2, tr_X_bar


IPMState = namedtuple("IPMState", ["i", "S", "eta", "T", "zeta", "omega", "mu", "step_size"])


def solve_quad_subprob_ipm(C, A_data, A_indices, U, b, upsilon, trace_ub, rho, bar_primal_obj,
                           z_bar, tr_X_bar, y, V, n, m, k, eta_init, S_init, ipm_eps, ipm_max_iters,
                           return_info=False):
    """Quadratic prox subproblem (solver/usbs.py:25-192).  Returns (eta, S)."""
    svec = lambda mx: U @ mx.reshape(-1)
    svec_inv = lambda vec: (U.T @ vec).reshape(k, k)

    Q_11, q_12, q_22, h_1, h_2, tr_X_bar = quad_subprob_data(
        C, A_data, A_indices, U, b, upsilon, trace_ub, rho, bar_primal_obj, z_bar, tr_X_bar, y, V,
        n, m, k)
    svec_I = svec(np.eye(k))

    eta0 = (tr_X_bar / trace_ub) * eta_init
    S0 = S_init / trace_ub
    T0 = svec_inv(Q_11 @ svec(S0) + eta0 * q_12 + h_1)
    zeta0 = np.dot(q_12, svec(S0)) + eta0 * q_22 + h_2
    dual_infeas_vals = np.append(np.diag(T0), zeta0)
    width = max(np.max(dual_infeas_vals) - np.min(dual_infeas_vals), 100.0)
    omega0 = max(-1.01 * np.min(dual_infeas_vals), width)
    T0 = T0 + omega0 * np.eye(k)
    zeta0 = zeta0 + omega0
    mu0 = (np.dot(svec(S0), svec(T0)) + eta0 * zeta0
           + omega0 * (1 - np.dot(svec_I, svec(S0)) - eta0)) / (2 * k + 4.0)

    possible_step_sizes = 0.9 ** np.arange(ipm_max_iters)

    def body(st):
        S, eta, T, zeta, omega, mu = st.S, st.eta, st.T, st.zeta, st.omega, st.mu
        kappa_1 = (1 - np.dot(svec_I, svec(S)) - eta) / omega
        kappa_2 = (zeta / eta) + q_22
        F_1 = Q_11 @ svec(S) + eta * q_12 + h_1
        F_1 = F_1 - svec(T) + omega * svec_I
        F_2 = np.dot(q_12, svec(S)) + eta * q_22 + h_2
        F_2 = F_2 - zeta + omega

        S_inv = np.linalg.inv(S)
        T_sym_kron_S_inv = 0.5 * U @ (np.kron(T, S_inv) + np.kron(S_inv, T)) @ U.T
        coeff = Q_11 + T_sym_kron_S_inv
        denom = kappa_1 * kappa_2 + 1
        coeff = coeff - (np.outer(q_12, kappa_1 * q_12 + svec_I)
                         + np.outer(svec_I, q_12 - kappa_2 * svec_I)) / denom
        trS = np.dot(svec_I, svec(S))
        ordv = -F_1 + mu * svec(S_inv) - svec(T)
        ordv = ordv + ((mu / omega + trS + eta - 1 + kappa_1 * (F_2 - mu / eta + zeta)) / denom) * q_12
        ordv = ordv + ((F_2 - mu / eta + zeta - kappa_2 * (mu / omega + trS + eta - 1)) / denom) * svec_I
        d_sS = np.linalg.solve(coeff, ordv)

        d_eta = (-np.dot(kappa_1 * q_12 + svec_I, d_sS) - (mu / omega) - trS - eta + 1
                 + kappa_1 * (-F_2 + mu / eta - zeta)) / denom
        d_omega = (-F_2 + mu / eta - zeta + kappa_2 * (mu / omega + trS + eta - 1)
                   - np.dot(q_12 - kappa_2 * svec_I, d_sS)) / denom
        d_sT = Q_11 @ d_sS + d_eta * q_12 + d_omega * svec_I + F_1
        d_zeta = np.dot(q_12, d_sS) + d_eta * q_22 + d_omega + F_2

        dS = svec_inv(d_sS)
        dT = svec_inv(d_sT)
        ratios = np.array([d_eta / eta, d_zeta / zeta, d_omega / omega])
        step = min(-1.0 / np.min(ratios), 1.0)
        step = 0.99 * (1.0 if step <= 0.0 else step)

        okS = _min_eig_positive_steps(S, dS, possible_step_sizes)
        okT = _min_eig_positive_steps(T, dT, possible_step_sizes)
        cS = np.max(possible_step_sizes * okS)
        cT = np.max(possible_step_sizes * okT)
        cand = cS if cS < cT else cT
        step = step if step < cand else cand

        S_n = S + step * dS
        eta_n = eta + step * d_eta
        T_n = T + step * dT
        zeta_n = zeta + step * d_zeta
        omega_n = omega + step * d_omega
        mu_n = (np.dot(svec(S_n), svec(T_n)) + eta_n * zeta_n
                + omega_n * (1 - np.dot(svec_I, svec(S_n)) - eta_n)) / (2 * k + 4.0)
        mu_n *= (0.5 - 0.4 * step ** 2) if step > 0.2 else 1.0
        mu_n = min(mu_n, mu)
        return IPMState(st.i + 1, S_n, eta_n, T_n, zeta_n, omega_n, mu_n, step)

    init = IPMState(0, S0, eta0, T0, zeta0, omega0, mu0, 1.0)
    final = while_loop(lambda st: st.mu > ipm_eps and st.step_size > ipm_eps ** 2, body, init,
                       ipm_max_iters)
    out = ((trace_ub / tr_X_bar) * final.eta, trace_ub * final.S)
    if return_info:
        return out, final
    return out


LBState = namedtuple("LBState", ["i", "S", "eta", "T", "zeta", "omega", "mu"])


def compute_lb_spec_est_ipm(C, A_data, A_indices, U, b, trace_ub, bar_primal_obj, z_bar, tr_X_bar,
                            y, V, n, k, ipm_eps, ipm_max_iters, return_info=False):
    """Spectral lower bound IPM (solver/usbs.py:195-356)."""
    svec = lambda mx: U @ mx.reshape(-1)
    svec_inv = lambda vec: (U.T @ vec).reshape(k, k)

    tr_X_bar = tr_X_bar if tr_X_bar > 0.0 else trace_ub
    g_1 = trace_ub * svec(V.T @ apply_A_adjoint_batched(n, A_data, A_indices, y, V) - V.T @ (C @ V))
    g_2 = (trace_ub / tr_X_bar) * (np.dot(z_bar, y) - bar_primal_obj)
    svec_I = svec(np.eye(k))

    S0 = 0.99 * (np.eye(k) / (k + 1.0))
    eta0 = 0.00001 if bar_primal_obj == 0.0 else 0.99 * (1.0 / (k + 1.0))
    T0 = svec_inv(g_1)
    zeta0 = g_2
    vals = np.append(np.diag(T0), zeta0)
    width = max(np.max(vals) - np.min(vals), 100.0)
    omega0 = max(-1.01 * np.min(vals), width)
    T0 = T0 + omega0 * np.eye(k)
    zeta0 = zeta0 + omega0
    mu0 = (np.dot(svec(S0), svec(T0)) + eta0 * zeta0
           + omega0 * (1 - np.dot(svec_I, svec(S0)) - eta0)) / (2 * k + 4.0)
    possible_step_sizes = 0.9 ** np.arange(ipm_max_iters)

    def body(st):
        S, eta, T, zeta, omega, mu = st.S, st.eta, st.T, st.zeta, st.omega, st.mu
        kappa_1 = (1 - np.dot(svec_I, svec(S)) - eta) / omega
        S_inv = np.linalg.inv(S)
        coeff = 0.5 * U @ (np.kron(T, S_inv) + np.kron(S_inv, T)) @ U.T
        coeff = coeff + ((zeta / eta) / (kappa_1 * zeta / eta + 1)) * np.outer(svec_I, svec_I)
        trS = np.dot(svec_I, svec(S))
        ordv = zeta / eta * (-kappa_1 * (mu / eta - g_2 - omega) + mu / omega + trS + eta - 1) \
            / (-kappa_1 * zeta / eta - 1) * svec_I
        ordv = ordv + (g_2 - mu / eta) * svec_I - g_1
        ordv = ordv + mu * svec(S_inv)
        d_sS = np.linalg.solve(coeff, ordv)

        d_eta = (np.dot(svec_I, d_sS) - kappa_1 * (mu / eta - g_2 - omega) + mu / omega + trS + eta - 1) \
            / (-kappa_1 * zeta / eta - 1)
        d_omega = -zeta / eta * d_eta + mu / eta
        d_omega = d_omega - g_2 - omega
        d_sT = d_omega * svec_I - svec(T) + g_1 + omega * svec_I
        d_zeta = d_omega - zeta + g_2 + omega

        dS = svec_inv(d_sS)
        dT = svec_inv(d_sT)
        ratios = np.array([d_eta / eta, d_zeta / zeta, d_omega / omega])
        step = min(-1.0 / np.min(ratios), 1.0)
        step = 0.99 * (1.0 if step <= 0.0 else step)
        okS = _min_eig_positive_steps(S, dS, possible_step_sizes)
        okT = _min_eig_positive_steps(T, dT, possible_step_sizes)
        cS = np.max(possible_step_sizes * okS)
        cT = np.max(possible_step_sizes * okT)
        cand = cS if cS < cT else cT
        step = step if step < cand else cand

        S_n = S + step * dS
        eta_n = eta + step * d_eta
        T_n = T + step * dT
        zeta_n = zeta + step * d_zeta
        omega_n = omega + step * d_omega
        mu_n = (np.dot(svec(S_n), svec(T_n)) + eta_n * zeta_n
                + omega_n * (1 - np.dot(svec_I, svec(S_n)) - eta_n)) / (2 * k + 4.0)
        mu_n *= (0.5 - 0.4 * step ** 2) if step > 0.2 else 1.0
        mu_n = min(mu_n, mu)
        return LBState(st.i + 1, S_n, eta_n, T_n, zeta_n, omega_n, mu_n)

    init = LBState(0, S0, eta0, T0, zeta0, omega0, mu0)
    final = while_loop(lambda st: st.mu > ipm_eps, body, init, ipm_max_iters)
    lb = np.dot(b, y) - np.dot(g_1, svec(final.S)) - final.eta * g_2
    if return_info:
        return lb, final
    return lb


def solve_step_subprob(C, A_data, A_indices, U, b, b_ineq_mask, upsilon, trace_ub, rho,
                       bar_primal_obj, z_bar, tr_X_bar, y, V, n, m, k, subprob_eps,
                       subprob_max_iters):
    """Alternating upsilon / IPM loop (solver/usbs.py:359-433)."""
    Sub = namedtuple("Sub", ["i", "eta", "S", "upsilon", "upsilon_gap"])

    def body(st):
        eta_n, S_n = solve_quad_subprob_ipm(C, A_data, A_indices, U, b, st.upsilon, trace_ub, rho,
                                            bar_primal_obj, z_bar, tr_X_bar, y, V, n, m, k,
                                            st.eta, st.S, subprob_eps, subprob_max_iters)
        lam, E = np.linalg.eigh(S_n)
        lam = np.clip(lam, 0, None)
        F = (V @ E) * np.sqrt(lam).reshape(1, -1)
        z_n = eta_n * z_bar + apply_A_operator_batched(m, A_data, A_indices, F)
        ups_n = b_ineq_mask * np.clip(b - z_n - (rho * y), 0.0, None)
        gap = np.max(np.abs(st.upsilon - ups_n))
        return Sub(st.i + 1, eta_n, S_n, ups_n, gap)

    trXb = tr_X_bar if tr_X_bar > 0.0 else trace_ub
    S_init = trace_ub * (0.99 * (np.eye(k) / (k + 1.0)))
    eta_init = (trace_ub / trXb) * (0.00001 if bar_primal_obj == 0.0 else 0.99 * (1.0 / (k + 1.0)))
    final = while_loop(lambda st: st.upsilon_gap > subprob_eps, body,
                       Sub(0, eta_init, S_init, upsilon, 1.0), subprob_max_iters)
    return final.eta, final.S, final.upsilon
