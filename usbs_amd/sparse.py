"""A minimal stand-in for jax.experimental.sparse.BCOO on the host.

The reference passes the cost matrix around as a BCOO with ``data`` (nse,)
and ``indices`` (nse, 2) (utils/common.py:11-26).  This class keeps that
layout so code written against SDPState reads the same, and converts from
and to scipy.sparse.
"""
import numpy as np
import scipy.sparse


class BCOO:
    def __init__(self, args, shape):
        data, indices = args
        self.data = np.asarray(data, dtype=np.float64)
        self.indices = np.asarray(indices, dtype=np.int64).reshape(-1, 2)
        self.shape = tuple(shape)

    @classmethod
    def from_scipy_sparse(cls, M):
        M = scipy.sparse.coo_matrix(M)
        return cls((M.data, np.stack([M.row, M.col], axis=1)), shape=M.shape)

    def to_scipy(self):
        return scipy.sparse.csr_matrix((self.data, (self.indices[:, 0], self.indices[:, 1])),
                                       shape=self.shape)

    def __neg__(self):
        return BCOO((-self.data, self.indices), self.shape)

    def __mul__(self, s):
        return BCOO((self.data * s, self.indices), self.shape)

    __rmul__ = __mul__

    def __truediv__(self, s):
        return BCOO((self.data / s, self.indices), self.shape)

    def __matmul__(self, x):
        return self.to_scipy() @ x

    @property
    def nse(self):
        return self.data.size


def coo_parts(C):
    """(rows int32, cols int32, vals f64, shape) of a BCOO-like or scipy matrix."""
    if isinstance(C, BCOO):
        return (C.indices[:, 0].astype(np.int32), C.indices[:, 1].astype(np.int32),
                C.data.astype(np.float64), C.shape)
    if scipy.sparse.issparse(C):
        M = scipy.sparse.coo_matrix(C)
        return M.row.astype(np.int32), M.col.astype(np.int32), M.data.astype(np.float64), M.shape
    if hasattr(C, "indices") and hasattr(C, "data"):
        ind = np.asarray(C.indices)
        return (ind[:, 0].astype(np.int32), ind[:, 1].astype(np.int32),
                np.asarray(C.data, dtype=np.float64), tuple(C.shape))
    raise TypeError("C must be a BCOO-like (data, indices) matrix or scipy.sparse matrix")
