"""FP64 CPU oracle for the USBS hot path (TEST INFRASTRUCTURE ONLY).

This package is a NumPy/SciPy restatement of the reference's JAX code for the
path that the HIP library accelerates.  It exists to check the HIP path, and
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  Nothing in ``usbs_amd`` imports it: the product
path fails loudly when the HIP library is missing instead of falling back here.

Module map (reference file:line each function follows is in its docstring):

* ``operators``  -- solver/utils.py:11-65 (A(.), A*(.), svec, Q_base)
* ``lanczos``    -- solver/lanczos.py:59-284 (thick-restart Lanczos)
* ``loop``       -- utils/loop.py:21-115 (bounded while-loop semantics)
* ``ipm``        -- solver/usbs.py:25-433 (prox-bundle IPM subproblems)
* ``usbs``       -- solver/usbs.py:436-789 (the USBS outer loop)
* ``cgal``       -- solver/cgal.py:21-243 (the CGAL loop)
* ``qap``        -- utils/qap_helpers.py:157-381, 625-653; utils/munkres.py:14-120
* ``common``     -- utils/common.py:11-121 (SDPState, scaling, Nystrom)
* ``maxcut``     -- utils/maxcut_helpers.py:26-92, 95-289 (warm starts), 292-302

Pinning status of each function (what checks it besides its own formula) is
listed in ``PINS`` below; tests/test_oracle_*.py implement those checks.

Things the oracle cannot reproduce because they live in JAX (absent here):
the threefry streams behind ``jax.random.normal`` (q0, Omega, the Krylov
breakdown vector, warm-start padding) and LAPACK's sign conventions for
eigenvectors / QR / SVD.  Random inputs are therefore injected explicitly and
results are compared through sign- and basis-invariant quantities.
"""

PINS = {
    "operators.apply_A_adjoint_slim": "dense A_c tensors (brute force), tests/test_oracle_operators.py",
    "operators.apply_A_operator_slim": "dense trace <A_c, uu^T> (brute force)",
    "operators.create_svec_matrix": "isometry <svec A, svec B> = <A, B> for symmetric A, B",
    "operators.create_Q_base": "brute-force sum_c svec(V^T A_c V) svec(V^T A_c V)^T",
    "lanczos.eigsh_smallest": "numpy.linalg.eigh of the dense operator; torus closed form",
    "lanczos._build_arrowhead_matrix": "hand-worked m<=6 arrowhead cases",
    "lanczos._pick_saved_ritz_values": "hand-worked cases of the gamma-ratio rule",
    "ipm.solve_quad_subprob_ipm": "scipy SLSQP on the same convex QP over (S>=0, eta>=0, trS+eta<=1), k=2",
    "ipm.compute_lb_spec_est_ipm": "closed form: lower bound of a linear objective over the spectraplex-with-eta",
    "usbs.usbs": "SDP closed forms: bipartite |E|, K_n n^2/4, odd cycle (n/2)(1+cos(pi/n))",
    "cgal.cgal": "SDP closed forms: bipartite |E| (exact), K_n n^2/4 (1e-3, first-order method)",
    "common.reconstruct_from_sketch": "exact recovery of a rank<=r PSD matrix (Tropp et al. 2017)",
    "maxcut.compute_max_cut": "brute force over all 2^n cuts on n<=14; integer edge-sum identity",
    "maxcut.implicit_warm_start_state": "unscaling round trip on the same graph; diagonal "
                                        "correction and scales on a grown graph",
    "maxcut.explicit_warm_start_state": "sketch of an exactly rank-r PSD X: trace kept, "
                                        "z = diag(X + delta E E^T) to the Nystrom shift",
    "maxcut.dual_only_warm_start_state": "y padded, primal quantities zero",
    "qap.munkres": "brute force over all n! assignments (the reference's 7 matrices, random n<=7)",
    "qap.qap_round": "sketch of the lifted brute-force optimal assignment rounds to the optimal value",
}
