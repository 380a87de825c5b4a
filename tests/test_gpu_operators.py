"""The constraint operator and its adjoint through the C ABI
(usbs_apply_A_operator / usbs_apply_A_adjoint) against the oracle's
restatement of solver/utils.py:11-37, on max-cut's diagonal triples and on
general triples with repeated constraints, off-diagonal (i, j) and
duplicates."""
import numpy as np
import pytest

from oracle import operators as oops
from usbs_amd import operators

pytestmark = pytest.mark.gpu


def _general_triples(n, m, nA, seed):
    rng = np.random.default_rng(seed)
    A_indices = np.stack([rng.integers(0, m, nA), rng.integers(0, n, nA), rng.integers(0, n, nA)], 1)
    A_indices[:5] = A_indices[5]          # duplicated triples sum
    A_data = rng.standard_normal(nA)
    return A_data, A_indices


@pytest.mark.parametrize("kind", ["maxcut", "general"])
@pytest.mark.parametrize("p", [1, 7])
def test_apply_A_operator_and_adjoint(gpu_ctx, kind, p):
    rng = np.random.default_rng(3 + p)
    if kind == "maxcut":
        n = m = 5000
        A_data, A_indices = np.ones(n), np.stack([np.arange(n)] * 3, 1)
    else:
        n, m = 3000, 1700
        A_data, A_indices = _general_triples(n, m, 20000, 11)
    U = rng.standard_normal((n, p))
    z = rng.standard_normal(m)
    print("case: %s p=%d n=%d m=%d nA=%d" % (kind, p, n, m, A_data.size), flush=True)
    got = operators.apply_A_operator_batched(m, A_data, A_indices, U, context=gpu_ctx)
    ref = oops.apply_A_operator_batched(m, A_data, A_indices, U)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    got = operators.apply_A_adjoint_batched(n, A_data, A_indices, z, U, context=gpu_ctx)
    ref = oops.apply_A_adjoint_batched(n, A_data, A_indices, z, U)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    if p == 1:
        u = U[:, 0]
        got = operators.apply_A_operator_slim(m, A_data, A_indices, u, context=gpu_ctx)
        ref = oops.apply_A_operator_slim(m, A_data, A_indices, u)
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
        got = operators.apply_A_adjoint_slim(n, A_data, A_indices, z, u, context=gpu_ctx)
        ref = oops.apply_A_adjoint_slim(n, A_data, A_indices, z, u)
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_adjoint_is_the_transpose_of_the_operator(gpu_ctx):
    """<A(u u^T), z> = <u, A*(z) u> for symmetric triples (a closed identity,
    independent of the oracle)."""
    n, m = 2000, 900
    A_data, A_indices = _general_triples(n, m, 8000, 5)
    rng = np.random.default_rng(9)
    u, z = rng.standard_normal(n), rng.standard_normal(m)
    lhs = operators.apply_A_operator_slim(m, A_data, A_indices, u, context=gpu_ctx) @ z
    rhs = u @ operators.apply_A_adjoint_slim(n, A_data, A_indices, z, u, context=gpu_ctx)
    assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs))


def test_empty_constraint_set_and_bad_indices(gpu_ctx):
    from usbs_amd._lib import USBSError
    n, m = 10, 4
    out = operators.apply_A_operator_batched(m, np.zeros(0), np.zeros((0, 3), int), np.ones((n, 2)),
                                             context=gpu_ctx)
    assert np.array_equal(out, np.zeros(m))
    with pytest.raises(USBSError):
        operators.apply_A_operator_slim(m, np.ones(1), np.array([[m, 0, 0]]), np.ones(n),
                                        context=gpu_ctx)
