"""Certified SDP values of benchmark max-cut instances (golden fixture).

The max-cut SDP max <C, X> s.t. diag(X) = 1, X PSD (C = L/4, the cost of
utils/maxcut_helpers.py:36-42) is solved here independently of USBS, by the
low-rank factorisation of Burer and Monteiro (Math. Program. 95, 2003):
X = V V^T with unit rows, rank p >= sqrt(2 n), maximised with L-BFGS.  Two
certificates pin the value:

* lower bound: <C, V V^T> of the final V (a feasible X);
* upper bound: for y_i = ((C V) V^T)_ii, weak duality gives
  OPT <= sum(y) + n max(0, lambda_max(C - diag(y))) (the shift that makes
  diag(y) - C PSD), computed with a dense eigensolver.

The fixture also records where the oracle's USBS loop (the FP64
restatement of solver/usbs.py) stops on ER(800, 0.06) seed 0 at rho = 0.1
with obj_gap <= 1e-5 and infeas_gap <= 1e-3 (Omega seed 1, q0 seed 2, the
reference's other max-cut settings): its iteration count, primal objective,
bound and gaps, so that a GPU test can hold its own run to the same
criterion without the oracle's minutes of CPU time.

Run: python tests/golden/make_maxcut_sdp_values.py  (writes
maxcut_sdp_values.json beside this script; several minutes on 8 cores).
"""
import json
import os
import sys

import numpy as np
import scipy.optimize

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from usbs_amd import problems  # noqa: E402  (graph generators only)
from usbs_amd.maxcut_helpers import laplacian_cost  # noqa: E402


def burer_monteiro(C, p, seed, iters=20000):
    n = C.shape[0]
    rng = np.random.default_rng(seed)

    def f(u):
        U = u.reshape(n, p)
        nr = np.linalg.norm(U, axis=1, keepdims=True)
        V = U / nr
        CV = C @ V
        val = np.sum(V * CV)
        G = 2.0 * CV
        G = (G - np.sum(G * V, axis=1, keepdims=True) * V) / nr
        return -val, -G.ravel()

    res = scipy.optimize.minimize(f, rng.standard_normal(n * p), jac=True, method="L-BFGS-B",
                                  options=dict(maxiter=iters, maxcor=30, gtol=1e-12, ftol=1e-16))
    U = res.x.reshape(n, p)
    return U / np.linalg.norm(U, axis=1, keepdims=True)


def certify(C, V):
    n = C.shape[0]
    CV = C @ V
    lower = float(np.sum(V * CV))
    y = np.sum(CV * V, axis=1)
    Cd = C.toarray() - np.diag(y)
    lmax = float(np.linalg.eigvalsh(Cd)[-1])
    upper = float(np.sum(y) + n * max(0.0, lmax))
    return lower, upper


def oracle_converged(W):
    from oracle import maxcut as omc, usbs as ousbs
    n = W.shape[0]
    Om = problems.gaussian_matrix(n, 10, 1)
    q0 = problems.unit_start_vector(n, 2)
    so = omc.initialize_state(W, 10, Om)
    log = []
    out, _ = ousbs.usbs(so, n, n, 2.0 * n, 2.0, 0.1, 0.25, 10, 1, 20000, 1e9, 1e-5, 1e-3, 1e300, 32,
                        10, 1e-7, 100, 1.0, omc.compute_max_cut, None, so.C, q0, log=log)
    last = log[-1]
    rec = {k: float(last[k]) for k in ("t", "primal_obj", "obj_ub", "obj_gap", "infeas_gap",
                                       "max_infeas", "callback_val")}
    rec["settings"] = ("k_curr 10, k_past 1, sketch 10, trace_factor 2, rho 0.1, beta 0.25, "
                       "Lanczos 32 x 10, eps 1e-7, subproblem 100 iterations; stop at obj_gap "
                       "<= 1e-5 and infeas_gap <= 1e-3")
    print("oracle", rec, flush=True)
    return rec


def main():
    out = {}
    cases = {"er800_p0.06_seed0": problems.erdos_renyi(800, 0.06, seed=0)}
    for name, W in cases.items():
        C = laplacian_cost(W)
        n = C.shape[0]
        p = int(np.ceil(np.sqrt(2 * n))) + 8
        best = None
        for seed in range(3):
            V = burer_monteiro(C, p, seed)
            lo, hi = certify(C, V)
            if best is None or hi - lo < best[1] - best[0]:
                best = (lo, hi)
            print(name, seed, lo, hi, hi - lo, flush=True)
        out[name] = dict(n=n, lower=best[0], upper=best[1], C_abs_sum=float(np.abs(C).sum()),
                         method="Burer-Monteiro rank %d, L-BFGS; dual certificate by dense eigvalsh"
                                % p)
    out["er800_p0.06_seed0"]["oracle_usbs_rho0.1"] = oracle_converged(cases["er800_p0.06_seed0"])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "maxcut_sdp_values.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
