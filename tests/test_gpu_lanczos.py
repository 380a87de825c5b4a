"""Parity of the HIP thick-restart Lanczos with the FP64 oracle
(solver/lanczos.py:239-284 restated in oracle/lanczos.py)."""
import numpy as np
import pytest
import scipy.sparse

from oracle import lanczos as olz, maxcut as omc
from usbs_amd import problems
from usbs_amd.lanczos import eigsh_smallest

pytestmark = pytest.mark.gpu


def _maxcut_operator(W, seed):
    C = omc.laplacian_cost(W)
    n = C.shape[0]
    A_data, A_indices, _, _ = omc.get_all_problem_data(n)
    y = np.random.default_rng(seed).standard_normal(n)
    q0 = problems.unit_start_vector(n, seed + 1)
    return (-C).tocsr(), A_data, A_indices, y, q0


@pytest.mark.parametrize("exec_mode", ["perop", "persistent"])
def test_lanczos_fp64_matches_oracle(gpu_ctx, exec_mode):
    W = problems.erdos_renyi(300, 0.06, seed=3)
    negC, A_data, A_indices, y, q0 = _maxcut_operator(W, 7)
    n = negC.shape[0]
    print("case: ER(300,0.06) nd=10 m=32 exec=%s" % exec_mode, flush=True)
    ev_o, V_o, st = olz.eigsh_smallest(n, negC, A_data, A_indices, y, q0, 10, 32, 10, 1e-7,
                                       return_state=True)
    ev_g, V_g, info = eigsh_smallest(n, negC, A_data, A_indices, y, q0, 10, 32, 10, 1e-7,
                                     exec_mode=exec_mode, return_info=True, context=gpu_ctx)
    assert info["restarts"] == st.restart
    np.testing.assert_allclose(ev_g, ev_o, rtol=1e-9, atol=1e-9 * np.abs(ev_o).max())
    # eigenvectors agree up to sign
    for j in range(10):
        s = np.sign(V_g[:, j] @ V_o[:, j])
        np.testing.assert_allclose(s * V_g[:, j], V_o[:, j], atol=1e-6)
    # against a dense eigensolver: the smallest Ritz value has converged, and
    # Ritz values never lie below the true eigenvalues (Cauchy interlacing)
    dense = np.linalg.eigvalsh((negC + scipy.sparse.diags(y)).toarray())
    np.testing.assert_allclose(ev_g[0], dense[0], rtol=1e-9)
    assert np.all(np.sort(ev_g) >= dense[:10] - 1e-9)
    # orthonormal basis
    np.testing.assert_allclose(V_g.T @ V_g, np.eye(10), atol=1e-10)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_persistent_bitwise_equals_perop(gpu_ctx, precision):
    """The persistent launch (FP32 basis: tiles cached in LDS; FP64: read
    from global memory) reproduces the per-op kernels bit for bit."""
    W = problems.torus(20, 30, seed=1)
    negC, A_data, A_indices, y, q0 = _maxcut_operator(W, 11)
    n = negC.shape[0]
    print("case: torus 20x30 bitwise perop == persistent, %s" % precision, flush=True)
    a = eigsh_smallest(n, negC, A_data, A_indices, y, q0, 10, 32, 10, 1e-7, exec_mode="perop",
                       precision=precision, context=gpu_ctx)
    b = eigsh_smallest(n, negC, A_data, A_indices, y, q0, 10, 32, 10, 1e-7,
                       exec_mode="persistent", precision=precision, context=gpu_ctx)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("exec_mode", ["perop", "persistent"])
def test_lanczos_mixed_precision_error_stated(gpu_ctx, exec_mode):
    """FP32 basis/operator, FP64 accumulation + Ritz: eigenvalues within
    2e-6 relative of the FP64 oracle (stated bound for the mixed form)."""
    W = problems.torus(40, 50, seed=2)
    negC, A_data, A_indices, y, q0 = _maxcut_operator(W, 5)
    n = negC.shape[0]
    dense = np.linalg.eigvalsh((negC + scipy.sparse.diags(y)).toarray())
    ev_g, V_g = eigsh_smallest(n, negC, A_data, A_indices, y, q0, 10, 32, 10, 1e-5,
                               precision="mixed", exec_mode=exec_mode, context=gpu_ctx)
    scale = np.abs(dense).max()
    assert np.abs(ev_g[0] - dense[0]) <= 2e-6 * scale


@pytest.mark.parametrize("exec_mode", ["persistent", "perop"])
def test_krylov_breakdown_small_n(gpu_ctx, exec_mode):
    """n <= 32: the Krylov space is exhausted inside the first restart, which
    forces the random-restart branch (lanczos.py:107-117).  The per-op form
    meets it in a step's phase A, stops the restart there and resumes after
    building the restart vector."""
    W = problems.erdos_renyi(24, 0.3, seed=4)
    negC, A_data, A_indices, y, q0 = _maxcut_operator(W, 9)
    n = negC.shape[0]
    print("case: ER(24,0.3) breakdown exec=%s" % exec_mode, flush=True)
    ev_o, _ = olz.eigsh_smallest(n, negC, A_data, A_indices, y, q0, 4, n, 3, 1e-7)
    ev_g, V_g, info = eigsh_smallest(n, negC, A_data, A_indices, y, q0, 4, n, 3, 1e-7,
                                     exec_mode=exec_mode, return_info=True, context=gpu_ctx)
    dense = np.linalg.eigvalsh((negC + scipy.sparse.diags(y)).toarray())
    np.testing.assert_allclose(ev_g, dense[:4], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(ev_o, dense[:4], rtol=1e-9, atol=1e-9)
    assert info["breakdowns"] >= 1


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_krylov_breakdown_forms_bitwise(gpu_ctx, precision):
    """Breakdowns in both forms: the per-op resume after a flagged step gives
    bitwise the persistent form's Ritz pairs, breakdown and step counts.  (In
    FP32 storage the exhausted Krylov space leaves residuals at rounding level,
    above the 1e-15 breakdown threshold, so only FP64 breaks down.)"""
    out = {}
    for n, p, seed in ((24, 0.3, 4), (30, 0.25, 7)):
        W = problems.erdos_renyi(n, p, seed=seed)
        negC, A_data, A_indices, y, q0 = _maxcut_operator(W, 9)
        for mode in ("persistent", "perop"):
            print("case: ER(%d,%g) breakdown %s exec=%s" % (n, p, precision, mode), flush=True)
            out[(n, mode)] = eigsh_smallest(n, negC, A_data, A_indices, y, q0, 4, n, 3, 1e-7,
                                            precision=precision, exec_mode=mode,
                                            return_info=True, context=gpu_ctx)
        ev_a, V_a, info_a = out[(n, "persistent")]
        ev_b, V_b, info_b = out[(n, "perop")]
        assert np.array_equal(ev_a, ev_b) and np.array_equal(V_a, V_b)
        assert info_a == info_b
        assert info_a["breakdowns"] >= (1 if precision == "fp64" else 0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_equals_single(precision):
    """Instances of different sizes advanced in one batched persistent launch
    give bitwise the results of their own single-instance calls."""
    from usbs_amd import _lib
    from usbs_amd.lanczos import eigsh_smallest_batch
    cases = [problems.torus(20, 30, seed=1), problems.erdos_renyi(300, 0.06, seed=3),
             problems.torus(7, 9, seed=4), problems.erdos_renyi(400, 0.12, seed=6)]
    probs = []
    for i, W in enumerate(cases):
        negC, A_data, A_indices, y, q0 = _maxcut_operator(W, 20 + i)
        probs.append((negC.shape[0], negC, A_data, A_indices, y, q0))
    print("case: batch of %d (n = %s) %s" % (len(probs), [p[0] for p in probs], precision), flush=True)
    ctxs = [_lib.Context(0) for _ in probs]
    batch = eigsh_smallest_batch(probs, 6, 24, 10, 1e-7, precision=precision, contexts=ctxs)
    for (n, negC, A_data, A_indices, y, q0), (ev_b, V_b), ctx in zip(probs, batch, ctxs):
        ev_s, V_s = eigsh_smallest(n, negC, A_data, A_indices, y, q0, 6, 24, 10, 1e-7,
                                   precision=precision, exec_mode="persistent", context=ctx)
        assert np.array_equal(ev_b, ev_s) and np.array_equal(V_b, V_s)


def test_lanczos_large_tiles_match_oracle(gpu_ctx):
    """n = 90 000 uses 512-row tiles, whose reductions run row-wise: Ritz
    values match the oracle to 1e-9 and per-op equals persistent bitwise."""
    W = problems.torus(300, 300, seed=5)
    negC, A_data, A_indices, y, q0 = _maxcut_operator(W, 13)
    n = negC.shape[0]
    print("case: torus 300x300 (n=%d) large tiles" % n, flush=True)
    ev_o, _ = olz.eigsh_smallest(n, negC, A_data, A_indices, y, q0, 6, 24, 4, 1e-7)
    a = eigsh_smallest(n, negC, A_data, A_indices, y, q0, 6, 24, 4, 1e-7, exec_mode="perop",
                       context=gpu_ctx)
    b = eigsh_smallest(n, negC, A_data, A_indices, y, q0, 6, 24, 4, 1e-7,
                       exec_mode="persistent", context=gpu_ctx)
    np.testing.assert_allclose(a[0], ev_o, rtol=1e-9, atol=1e-9 * np.abs(ev_o).max())
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lanczos_heavy_rows_match_oracle(gpu_ctx, precision):
    """ER(800, 0.06), about 49 operator entries per row: the SpMV takes a
    wavefront per row over 16-row tiles.  FP64 Ritz values match the oracle
    to 1e-9 (mixed: 2e-6 of the spectrum's scale), and per-op equals
    persistent bitwise."""
    W = problems.erdos_renyi(800, 0.06, seed=0)
    negC, A_data, A_indices, y, q0 = _maxcut_operator(W, 17)
    n = negC.shape[0]
    assert negC.nnz + n >= 32 * n
    print("case: ER(800,0.06) heavy rows %s" % precision, flush=True)
    ev_o, _ = olz.eigsh_smallest(n, negC, A_data, A_indices, y, q0, 10, 32, 10, 1e-7)
    a = eigsh_smallest(n, negC, A_data, A_indices, y, q0, 10, 32, 10, 1e-7, exec_mode="perop",
                       precision=precision, context=gpu_ctx, return_info=True)
    b = eigsh_smallest(n, negC, A_data, A_indices, y, q0, 10, 32, 10, 1e-7,
                       exec_mode="persistent", precision=precision, context=gpu_ctx)
    scale = np.abs(ev_o).max()
    nc = a[2]["num_converged"]
    assert nc >= 1
    if precision == "fp64":
        # converged Ritz values; the unconverged tail depends on rounding
        np.testing.assert_allclose(a[0][:nc], ev_o[:nc], rtol=1e-9, atol=1e-9 * scale)
    else:
        assert np.abs(a[0][0] - ev_o[0]) <= 2e-6 * scale
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
