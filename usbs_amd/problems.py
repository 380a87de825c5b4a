"""Seeded synthetic problem generators for the max-cut and QAP workloads.

All randomness comes from numpy's PCG64 with an explicit seed so that the HIP
path, the CPU oracle and the benchmark see identical inputs.  Graphs are
returned as symmetric scipy CSR adjacency matrices (the object the
reference's max-cut script loads from a .mat file before calling
``initialize_state``, scripts/warm_start_maxcut_usbs.py:82-99).
"""
import math

import numpy as np
import scipy.sparse


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _sym_from_edges(n, i, j, w):
    A = scipy.sparse.coo_matrix((np.concatenate([w, w]), (np.concatenate([i, j]),
                                                           np.concatenate([j, i]))), shape=(n, n))
    A = A.tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return A


def erdos_renyi(n, p, seed, signed=False):
    """G(n, p): each unordered pair {i<j} is an edge with probability p.
    Weights are 1, or uniform on {-1, +1} when ``signed``.  Small n draws one
    Bernoulli per pair (upper triangle, row-major); large n draws the edge
    count from Binomial(n(n-1)/2, p) and then distinct uniform pairs."""
    rng = _rng(seed)
    if n <= 4096:
        iu, ju = np.triu_indices(n, 1)
        keep = rng.random(iu.size) < p
        i, j = iu[keep], ju[keep]
    else:
        npairs = n * (n - 1) // 2
        target = int(rng.binomial(npairs, p))
        keys = np.empty(0, dtype=np.int64)
        while keys.size < target:
            need = target - keys.size
            a = rng.integers(0, n, size=int(need * 1.1) + 16)
            b = rng.integers(0, n, size=a.size)
            lo, hi = np.minimum(a, b), np.maximum(a, b)
            ok = lo != hi
            cand = lo[ok].astype(np.int64) * n + hi[ok]
            keys = np.unique(np.concatenate([keys, cand]))
        keys = keys[rng.permutation(keys.size)[:target]]
        keys.sort()
        i, j = keys // n, keys % n
    w = rng.choice(np.array([-1.0, 1.0]), size=i.size) if signed else np.ones(i.size)
    return _sym_from_edges(n, i, j, w)


def torus(rows, cols, seed, signed=True):
    """2-D toroidal grid rows x cols: vertex r*cols+c joined to its right and
    down neighbours with wrap-around (2*rows*cols edges), weights uniform on
    {-1, +1} by seed (or 1)."""
    rng = _rng(seed)
    r, c = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    v = (r * cols + c).ravel()
    right = (r * cols + (c + 1) % cols).ravel()
    down = (((r + 1) % rows) * cols + c).ravel()
    i = np.concatenate([v, v])
    j = np.concatenate([right, down])
    w = rng.choice(np.array([-1.0, 1.0]), size=i.size) if signed else np.ones(i.size)
    return _sym_from_edges(rows * cols, i, j, w)


def rgg_radius(n, mean_degree):
    """Radius giving the requested mean degree in the unit square, corrected
    for the boundary: E[deg] = (n-1) (pi r^2 - 8 r^3/3 + r^4/2)."""
    lo, hi = 0.0, 1.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        d = (n - 1) * (math.pi * mid ** 2 - 8.0 * mid ** 3 / 3.0 + mid ** 4 / 2.0)
        lo, hi = (mid, hi) if d < mean_degree else (lo, mid)
    return 0.5 * (lo + hi)


def qap_instance(l, seed):
    """QAP data: flow F symmetric integers 0-100 with ~70% zeros; distance D
    from l random points on the 100x100 integer grid (Euclidean, rounded,
    clipped at 100).  F is the sparser matrix, called W in the reference's
    builder (utils/qap_helpers.py:35-37)."""
    rng = _rng(seed)
    iu = np.triu_indices(l, 1)
    vals = rng.integers(1, 101, size=iu[0].size).astype(np.float64)
    vals[rng.random(iu[0].size) < 0.7] = 0.0
    F = np.zeros((l, l))
    F[iu] = vals
    F = F + F.T
    pts = rng.integers(0, 101, size=(l, 2)).astype(np.float64)
    D = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    D = np.minimum(np.round(D), 100.0)
    return D, F


def gaussian_matrix(n, r, seed):
    """Standard normal n x r matrix (the sketch test matrix Omega)."""
    return _rng(seed).standard_normal((n, r))


def unit_start_vector(n, seed):
    """Normalised standard normal start vector q0 (cf. solver/usbs.py:691-692)."""
    q = _rng(seed).standard_normal(n)
    return q / np.linalg.norm(q)


def flip_edge_signs(W, frac, seed):
    """Copy of the symmetric weighted adjacency W with the signs of a random
    ``frac`` share of its edges flipped (both (i, j) and (j, i))."""
    import scipy.sparse
    U = scipy.sparse.triu(W, k=1).tocoo()
    rng = _rng(seed)
    flip = rng.random(U.nnz) < frac
    data = np.where(flip, -U.data, U.data)
    T = scipy.sparse.coo_matrix((data, (U.row, U.col)), shape=W.shape)
    return (T + T.T).tocsr()


def random_geometric(n, radius, seed):
    """Random geometric graph: n points uniform in the unit square (seeded),
    an edge between every pair closer than ``radius``, unit weights.  Pairs
    come from scipy's k-d tree, so n = 2^20 takes seconds."""
    import scipy.sparse
    import scipy.spatial
    pts = _rng(seed).random((n, 2))
    pairs = scipy.spatial.cKDTree(pts).query_pairs(radius, output_type="ndarray")
    i, j = pairs[:, 0], pairs[:, 1]
    W = scipy.sparse.coo_matrix((np.ones(i.size), (i, j)), shape=(n, n))
    return (W + W.T).tocsr()
