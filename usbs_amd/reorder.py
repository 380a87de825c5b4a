"""Locality reordering of the node space (host-side setup, no counterpart in
the reference).

The Lanczos SpMV gathers the basis column at the operator's column indices.
For a graph whose nodes are numbered in random order (the random geometric
graph's points, for instance) those gathers scatter over the whole vector and
miss the caches.  Renumbering the nodes with reverse Cuthill-McKee on the
pattern of C - A*(y) puts the neighbours of a row next to it.  A permutation
of the node space changes nothing in the SDP: the objective, the constraint
values and the cut values are invariant; only the summation order of the
reductions, hence the rounding, differs.  Constraints keep their order.

    st2, q02, perm = permute_state(st, q0)     # before upload
    out = usbs(st2, ..., q0=q02)               # as usual
    P = unpermute_rows(out.P, perm)            # sketch rows back in input order
"""
import numpy as np
import scipy.sparse
from scipy.sparse.csgraph import reverse_cuthill_mckee

from .sparse import BCOO


def rcm_permutation(C, A_indices, n):
    """perm with perm[new] = old: reverse Cuthill-McKee order of the
    symmetric pattern of C and of the constraint triples' (i, j) pairs."""
    ci = np.asarray(C.indices).reshape(-1, 2)
    ai = np.asarray(A_indices).reshape(-1, 3)[:, 1:3]
    rows = np.concatenate([ci[:, 0], ci[:, 1], ai[:, 0], ai[:, 1]])
    cols = np.concatenate([ci[:, 1], ci[:, 0], ai[:, 1], ai[:, 0]])
    G = scipy.sparse.csr_matrix((np.ones(rows.size, dtype=np.int8), (rows, cols)), shape=(n, n))
    return np.asarray(reverse_cuthill_mckee(G, symmetric_mode=True), dtype=np.int64)


def permute_state(state, q0=None, perm=None):
    """The same SDP with node i renamed pos[i] (pos = inverse of perm): C's
    indices, the triples' (i, j), and the rows of Omega, P and q0.  Returns
    (state, q0, perm); perm defaults to rcm_permutation."""
    n = state.C.shape[0]
    if perm is None:
        perm = rcm_permutation(state.C, state.A_indices, n)
    perm = np.asarray(perm, dtype=np.int64)
    if perm.shape != (n,) or not np.array_equal(np.sort(perm), np.arange(n)):
        raise ValueError("perm must be a permutation of range(n)")
    pos = np.empty(n, dtype=np.int64)
    pos[perm] = np.arange(n)
    C = BCOO((state.C.data, pos[np.asarray(state.C.indices)]), shape=state.C.shape)
    Ai = np.array(state.A_indices, dtype=np.int64).reshape(-1, 3)
    Ai[:, 1:3] = pos[Ai[:, 1:3]]
    rows = lambda a: None if a is None else np.asarray(a)[perm]  # noqa: E731
    st = state._replace(C=C, A_indices=Ai, Omega=rows(state.Omega), P=rows(state.P))
    return st, rows(q0), perm


def unpermute_rows(X, perm):
    """Rows of a node-space array of the permuted problem back in the input
    order (row perm[k] of the result is row k of X)."""
    if X is None:
        return None
    X = np.asarray(X)
    out = np.empty_like(X)
    out[np.asarray(perm)] = X
    return out
