"""Host logic of usbs_amd.usbs.cut_scale_of: the factor s with C = s * the
matrix the rounding callback prices cuts with (the reference's scripts pass
C / SCALE_C, scripts/warm_start_maxcut_usbs.py:121).  No GPU."""
from types import SimpleNamespace

import numpy as np
import pytest
import scipy.sparse

from usbs_amd.sparse import BCOO
from usbs_amd.usbs import cut_scale_of


def _state(C, scale_c=0.25):
    return SimpleNamespace(C=C, SCALE_C=scale_c)


def _bcoo(rows, cols, vals, n):
    return BCOO((np.asarray(vals, float), np.stack([rows, cols], axis=1)), shape=(n, n))


def _random_c(n=60, seed=3):
    rng = np.random.default_rng(seed)
    W = scipy.sparse.random(n, n, density=0.1, random_state=seed, format="coo")
    W = scipy.sparse.coo_matrix(W + W.T)
    return W.row.astype(np.int32), W.col.astype(np.int32), rng.integers(1, 5, W.nnz).astype(float), n


def test_none_means_scale_c():
    r, c, v, n = _random_c()
    assert cut_scale_of(_state(_bcoo(r, c, v, n), 0.125), None) == 0.125


def test_same_object_is_one():
    r, c, v, n = _random_c()
    C = _bcoo(r, c, v, n)
    assert cut_scale_of(_state(C), C) == 1.0


def test_scaled_copy_same_order():
    r, c, v, n = _random_c()
    C = _bcoo(r, c, 3.0 * v, n)
    assert cut_scale_of(_state(C), _bcoo(r, c, v, n)) == pytest.approx(3.0, rel=1e-15)


def test_scaled_copy_other_order_and_scipy():
    r, c, v, n = _random_c()
    C = _bcoo(r, c, 0.5 * v, n)
    perm = np.random.default_rng(1).permutation(v.size)
    other = scipy.sparse.coo_matrix((v[perm], (r[perm], c[perm])), shape=(n, n))
    assert cut_scale_of(_state(C), other) == pytest.approx(0.5, rel=1e-15)


def test_not_proportional_raises():
    r, c, v, n = _random_c()
    w = v.copy()
    w[0] += 1.0
    with pytest.raises(NotImplementedError):
        cut_scale_of(_state(_bcoo(r, c, v, n)), _bcoo(r, c, w, n))
