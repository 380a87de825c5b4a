"""Thick-restart Lanczos for the smallest eigenpairs of C + A*(v), FP64 NumPy.

Restates solver/lanczos.py:59-284 step for step (oracle code: see
oracle/__init__.py).  That reference file is Copyright 2020 Google LLC under
the Apache License, Version 2.0; this restatement follows its algorithm and
keeps some of its function names (_iterative_classical_gram_schmidt,
_build_arrowhead_matrix, _pick_saved_ritz_values) so the two read side by
side.  Differences that are forced, not chosen:

* the Krylov-breakdown vector (solver/lanczos.py:108-111) comes from
  ``breakdown_normal`` below instead of jax.random.normal(PRNGKey(i+1)); the
  HIP library uses the same counter-based generator, so both sides agree;
* eigenvector signs follow numpy's LAPACK, not XLA's; every quantity the
  tests compare is sign-invariant.
"""
from collections import namedtuple

import numpy as np

from .operators import apply_A_adjoint_slim

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) on a uint64 array."""
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def breakdown_normal(key, n):
    """n standard normals from a counter-based stream: Box-Muller (1958) on two
    SplitMix64 uniforms per entry, counter = key * 2^32 + 2*idx (+1)."""
    idx = np.arange(n, dtype=np.uint64)
    base = np.uint64(key) << np.uint64(32)
    with np.errstate(over="ignore"):
        a = _splitmix64(base + np.uint64(2) * idx)
        b = _splitmix64(base + np.uint64(2) * idx + np.uint64(1))
    u1 = (a >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u2 = (b >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def _iterative_classical_gram_schmidt(Q, x):
    """CGS2, "twice is enough" (solver/lanczos.py:59-77)."""
    q = x
    h = Q.T @ q
    q = q - Q @ h
    r = h
    h = Q.T @ q
    q = q - Q @ h
    r = r + h
    return r, q


def _lanczos_restart(n, C, A_data, A_indices, adjoint_left_vec, k, m, Q, alpha, beta, tolerance):
    """Inner steps i = k .. m-1 (solver/lanczos.py:80-125)."""
    col = np.arange(m + 1)
    for i in range(k, m):
        q_prev = Q[:, i]
        q = C @ q_prev + apply_A_adjoint_slim(n, A_data, A_indices, adjoint_left_vec, q_prev)
        Q_valid = (i >= col) * Q
        r, q = _iterative_classical_gram_schmidt(Q_valid, q)
        a = r[i]
        b = np.linalg.norm(q)
        if b < 1e-15:
            _, _q = _iterative_classical_gram_schmidt(Q_valid, breakdown_normal(i + 1, n))
            q = _q / np.linalg.norm(_q)
        else:
            q = q / b
        Q[:, i + 1] = q
        alpha[i] = a
        beta[i] = b
    return Q, alpha, beta


def _build_arrowhead_matrix(alpha, beta, k):
    """Arrowhead T (solver/lanczos.py:128-135): arrow column at k holds
    beta[:k], then the tridiagonal band beta[j] at (j, j+1) for j >= k."""
    m = alpha.size
    ar = np.arange(m)
    upper = np.outer(beta * (k > ar), (k == ar)) + np.diag(beta[:-1] * (k <= np.arange(m - 1)), k=1)
    return upper + upper.T + np.diag(alpha)


def _pick_saved_ritz_values(lambda_, num_converged, num_desired):
    """Gamma-ratio restart choice (solver/lanczos.py:138-168).  Indexing past
    the end clamps to the last entry, as XLA's gather does."""
    m = lambda_.size
    delta = min(m - num_desired, 2 * (m - num_converged) // 5)
    k_range = np.arange(m)
    kr = k_range + delta
    lambda_target = lambda_[min(num_converged, m - 1)]
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = (lambda_ - lambda_target) / (lambda_[np.minimum(kr, m - 1)] - lambda_target)
    gamma = np.where(kr <= m, ratio, 0.0)
    k_left = int(np.argmax(gamma))
    k_right = k_left + delta
    return (k_left >= k_range) | (k_right < k_range), gamma


ThickRestartState = namedtuple("ThickRestartState", "Q alpha beta restart num_saved num_converged")


def _thick_restart_lanczos(n, C, A_data, A_indices, adjoint_left_vec, q0, num_desired,
                           inner_iterations, max_restarts, tolerance):
    """Restart loop (solver/lanczos.py:177-236); at least one restart runs."""
    m = inner_iterations
    Q = np.zeros((n, m + 1))
    Q[:, 0] = q0
    alpha = np.zeros(m)
    beta = np.zeros(m)
    k = 0
    num_converged = 0
    restart = 0
    for _ in range(max_restarts):
        if not num_converged < num_desired:
            break
        Q, alpha, beta = _lanczos_restart(n, C, A_data, A_indices, adjoint_left_vec, k, m, Q,
                                          alpha, beta, tolerance)
        T = _build_arrowhead_matrix(alpha, beta, k)
        ritz_values, Y = np.linalg.eigh(T)
        residual_norm = beta[-1] * np.abs(Y[-1, :])
        converged = residual_norm < tolerance
        num_converged = m if converged.all() else int(np.argmin(converged))
        saved, _ = _pick_saved_ritz_values(ritz_values, num_converged, num_desired)
        k_next = int(np.sum(saved))
        saved_to_front = np.argsort(~saved, kind="stable")
        ritz_values = ritz_values[saved_to_front]
        Y = Y[:, saved_to_front]
        updated = k_next > np.arange(m)
        alpha = np.where(updated, ritz_values, 0.0)
        beta_hat = np.where(updated, beta[-1] * Y[-1, :], 0.0)
        Q_hat = Q.copy()
        Q_hat[:, :-1] = np.where(updated, Q[:, :-1] @ Y, Q[:, :-1])
        Q_hat[:, k_next] = Q[:, -1]
        Q, beta, k = Q_hat, beta_hat, k_next
        restart += 1
    return ThickRestartState(Q, alpha, beta, restart, k, num_converged)


def eigsh_smallest(n, C, A_data, A_indices, adjoint_left_vec, q0, num_desired,
                   inner_iterations, max_restarts, tolerance, return_state=False):
    """num_desired smallest eigenpairs of C + A*(v) (solver/lanczos.py:239-284).
    C is anything supporting ``C @ vector`` (scipy sparse or dense)."""
    state = _thick_restart_lanczos(n, C, np.asarray(A_data, dtype=np.float64), np.asarray(A_indices),
                                   np.asarray(adjoint_left_vec, dtype=np.float64),
                                   np.asarray(q0, dtype=np.float64), num_desired,
                                   inner_iterations, max_restarts, tolerance)
    vals = state.alpha[:num_desired].copy()
    vecs = state.Q[:, :num_desired].copy()
    if return_state:
        return vals, vecs, state
    return vals, vecs
