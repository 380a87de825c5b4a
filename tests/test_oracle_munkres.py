"""The Hungarian algorithm and the QAP rounding (utils/munkres.py:14-120,
utils/qap_helpers.py:625-653).

Oracle pins: on the seven cost matrices the reference's own munkres module
runs (utils/munkres.py:125-171; it prints their assignments without expected
values) and on random integer and real matrices up to n = 7, the oracle's
assignment is a permutation whose cost equals the brute-force minimum over
all n! permutations; the rounding of a sketch of x x^T, x the lifted optimal
permutation (found by brute force), returns the optimal QAP value.
The library's host-side Hungarian algorithm (usbs_munkres, no GPU needed)
must return the oracle's assignment exactly, ties included."""
import ctypes
import itertools

import numpy as np
import pytest

from oracle import qap as oq
from usbs_amd import _lib, problems

REFERENCE_MATRICES = [
    [[1, 2, 3], [2, 4, 6], [3, 6, 9]],
    [[7, 6, 2, 9, 2], [6, 2, 1, 3, 9], [5, 6, 8, 9, 5], [6, 8, 5, 8, 6], [9, 5, 6, 4, 7]],
    [[62, 75, 80, 93, 95, 97], [75, 80, 82, 85, 71, 97], [80, 75, 81, 98, 90, 97],
     [78, 82, 84, 80, 50, 98], [90, 85, 85, 80, 85, 99], [65, 75, 80, 75, 68, 96]],
    [[1, 2, 3, 4], [2, 4, 6, 8], [3, 6, 9, 12], [4, 8, 12, 16]],
    [[1, 2, 3], [3, 3, 3], [3, 3, 2]],
    [[7, 4, 3], [3, 1, 2], [3, 0, 0]],
    [[-1, -2, -3], [-3, -3, -3], [-3, -3, -2]],
]


def _brute_min(C):
    n = C.shape[0]
    return min(sum(C[i, p[i]] for i in range(n)) for p in itertools.permutations(range(n)))


def _cases():
    rng = np.random.default_rng(0)
    cases = [np.array(m, dtype=np.float64) for m in REFERENCE_MATRICES]
    for n in (2, 4, 5, 6, 7):
        cases.append(rng.integers(0, 4, (n, n)).astype(np.float64))   # many ties
        cases.append(rng.standard_normal((n, n)))
    return cases


@pytest.mark.parametrize("case", range(len(_cases())))
def test_oracle_munkres_is_optimal(case):
    C = _cases()[case]
    n = C.shape[0]
    mask = oq.munkres(n, C)
    assert (mask.sum(0) == 1).all() and (mask.sum(1) == 1).all() and set(np.unique(mask)) <= {0, 1}
    assert np.isclose(float((C * mask).sum()), _brute_min(C), rtol=0, atol=1e-12)


def test_library_munkres_equals_oracle():
    lib = _lib.lib()
    lib.usbs_munkres.argtypes = [ctypes.c_int32, _lib._dp, _lib._ip]
    rng = np.random.default_rng(1)
    cases = _cases() + [rng.integers(0, 3, (12, 12)).astype(np.float64),
                        rng.standard_normal((30, 30))]
    # plus cost matrices of the rounding's form, max(c) - c of a smooth
    # sketch column, on which the n^2-step bound can end the search early
    E = np.random.default_rng(2).standard_normal((26, 4))
    for i in range(4):
        c = E[1:, i].reshape(5, 5)
        cases.append(c.max() - c)
    for C in cases:
        n = C.shape[0]
        C = np.ascontiguousarray(C)
        out = np.zeros((n, n), dtype=np.int32)
        assert lib.usbs_munkres(n, _lib.dptr(C), _lib.iptr(out)) == 0
        assert np.array_equal(out, oq.munkres(n, C)), C


def test_oracle_qap_round_recovers_the_optimal_assignment():
    l = 5
    D, W = problems.qap_instance(l, 3)
    perms = list(itertools.permutations(range(l)))
    costs = [float(np.sum(W * D[np.ix_(p, p)])) for p in perms]
    best = perms[int(np.argmin(costs))]
    x = oq.lifted_assignment(best)
    N = l * l + 1
    rng = np.random.default_rng(4)
    Om = rng.standard_normal((N, l))
    val, vals = oq.qap_round(np.outer(x, x) @ Om, Om, l, D, W)
    assert val == min(costs)
