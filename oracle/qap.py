"""The lifted QAP SDP built entry by entry in plain Python (oracle code).

Restates utils/qap_helpers.py:157-381.  Each constraint family is written as
the explicit loop the reference gives beside its vectorised JAX code
(utils/qap_helpers.py:186-318), with the coefficients and right-hand sides of
the code that runs (the loops in its comments say -0.5/-0.5/1 and b = 1
where the code uses -1/-1/2 and b = 2; the code is followed).  The spectral
norm used for SCALE_A is taken from a dense SVD (numpy's 2-norm), so this
module is for small l only; Omega is injected.
"""
import numpy as np
import scipy.sparse

from .common import SDPState, scale_sdp_state


def build_objective_matrix(D, W):
    """Entries (1 + l*a + i, 1 + l*b + j, W[a,b] D[i,j]) for the nonzeros of
    W (row-major) and, within each, of D (row-major), as three lists
    (utils/qap_helpers.py:157-172)."""
    l = len(D)
    rows, cols, vals = [], [], []
    for a in range(l):
        for b_ in range(l):
            if W[a][b_] == 0:
                continue
            for i in range(l):
                for j in range(l):
                    if D[i][j] == 0:
                        continue
                    rows.append(1 + l * a + i)
                    cols.append(1 + l * b_ + j)
                    vals.append(float(W[a][b_]) * float(D[i][j]))
    return rows, cols, vals


def get_all_problem_data(N, C_rows, C_cols):
    """Constraint triples, data, b and inequality mask in the reference's row
    order (utils/qap_helpers.py:175-318)."""
    l = int(round((N - 1) ** 0.5))
    assert l * l + 1 == N
    trip, data, b, mask = [], [], [], []
    c = 0
    trip.append((0, 0, 0)); data.append(1.0); b.append(1.0); mask.append(0.0)
    c += 1
    for j in range(1, N):                       # diag(Y) = vec(B)
        trip += [(c, j, 0), (c, 0, j), (c, j, j)]
        data += [-1.0, -1.0, 2.0]
        b.append(0.0); mask.append(0.0)
        c += 1
    for j1 in range(l):                         # B 1 = 1
        for j2 in range(l):
            v = j1 * l + j2 + 1
            trip += [(c, v, 0), (c, 0, v)]
            data += [1.0, 1.0]
        b.append(2.0); mask.append(0.0)
        c += 1
    for j1 in range(l):                         # 1'B = 1'
        for j2 in range(l):
            v = j1 + j2 * l + 1
            trip += [(c, v, 0), (c, 0, v)]
            data += [1.0, 1.0]
        b.append(2.0); mask.append(0.0)
        c += 1
    for coords in (lambda j1, j2, d: (j1 + d * l + 1, j2 + d * l + 1),     # tr_1(Y) = I
                   lambda j1, j2, d: (j1 * l + d + 1, j2 * l + d + 1)):    # tr_2(Y) = I
        first, second = [], []
        for j1 in range(l):
            for j2 in range(l):
                for d in range(l):
                    p, q = coords(j1, j2, d)
                    first.append((c, p, q))
                    second.append((c, q, p))
                b.append(2.0 if j1 == j2 else 0.0); mask.append(0.0)
                c += 1
        trip += first + second
        data += [1.0] * (len(first) + len(second))
    first, second = [], []
    for r, s in zip(C_rows, C_cols):            # -Y <= 0 on C's upper triangle
        if r <= s:
            first.append((c, r, s))
            second.append((c, s, r))
            b.append(0.0); mask.append(1.0)
            c += 1
    trip += first + second
    data += [-0.5] * (len(first) + len(second))
    return (np.array(data), np.array(trip, dtype=np.int64).reshape(-1, 3), np.array(b),
            np.array(mask))


def constraint_matrix(A_data, A_indices, m, N):
    """Dense m x N^2 matrix of the constraint operator (duplicates summed)."""
    A = np.zeros((m, N * N))
    for (c, i, j), v in zip(A_indices, A_data):
        A[c, i * N + j] += v
    return A


def initialize_state(D, W, sketch_dim, Omega):
    """Cold start (utils/qap_helpers.py:321-381) with C = -(W (x) D) padded."""
    l = len(D)
    N = l * l + 1
    rows, cols, vals = build_objective_matrix(D, W)
    vals = [-v for v in vals]
    A_data, A_indices, b, mask = get_all_problem_data(N, rows, cols)
    m = b.shape[0]
    SCALE_X = 1.0 / float(l + 1)
    SCALE_C = 1.0 / float(np.sqrt(np.sum(np.square(vals))))
    sq = np.zeros(m)
    for (c, _, _), v in zip(A_indices, A_data):
        sq[c] += v * v
    SCALE_A = 1.0 / np.sqrt(sq)
    A = constraint_matrix(SCALE_A[A_indices[:, 0]] * A_data, A_indices, m, N)
    SCALE_A = SCALE_A / np.linalg.norm(A, 2)
    C = scipy.sparse.csr_matrix((vals, (rows, cols)), shape=(N, N))
    P = np.zeros_like(Omega)
    s = SDPState(C=C, A_indices=A_indices, A_data=A_data, b=b, b_ineq_mask=mask, X=None, P=P,
                 Omega=Omega, y=np.zeros(m), z=np.zeros(m), tr_X=0.0, primal_obj=0.0,
                 SCALE_C=SCALE_C, SCALE_X=SCALE_X, SCALE_A=SCALE_A)
    return scale_sdp_state(s), (rows, cols, vals)


def lifted_assignment(perm):
    """x = [1; vec(P)] with P[a, i] = 1 iff perm[a] = i, at index 1 + l*a + i."""
    l = len(perm)
    x = np.zeros(l * l + 1)
    x[0] = 1.0
    for a, i in enumerate(perm):
        x[1 + l * a + i] = 1.0
    return x


def munkres(n, cost):
    """Minimum-cost assignment mask (n x n, 1 at the chosen entries) by the
    steps of utils/munkres.py:14-120 in the same order: row reduction;
    zeros starred in row-major order; then, until n stars, the first
    uncovered zero (row-major) is primed and either covers its row (a star
    in the row: the star's column is uncovered) or starts an augmenting
    path; with no uncovered zero, the smallest uncovered value (covered
    entries read as the matrix maximum) is subtracted from the uncovered
    columns and added to the covered rows."""
    M = np.array(cost, dtype=np.float64)
    M = M - M.min(axis=1).reshape(-1, 1)
    mask = np.zeros((n, n), dtype=np.int64)
    rc = np.zeros(n, dtype=bool)
    cc = np.zeros(n, dtype=bool)
    for idx in range(n * n):
        r, c = divmod(idx, n)
        if M[r, c] == 0 and not rc[r] and not cc[c]:
            mask[r, c] = 1
            rc[r] = True
            cc[c] = True
    rc[:] = False
    cc = (mask == 1).any(axis=0)
    for _ in range(n * n):
        if (mask == 1).sum() >= n:
            break
        cover = rc[:, None] | cc[None, :]
        zeros = (M == 0) & ~cover
        if not zeros.any():
            v = np.where(~cover, M, M.max()).min()
            M = np.where(~cc[None, :], M - v, M)
            M = np.where(rc[:, None], M + v, M)
            continue
        r, c = divmod(int(np.argmax(zeros)), n)
        mask[r, c] = 2
        if (mask[r] == 1).any():
            rc[r] = True
            cc[int(np.argmax(mask[r] == 1))] = False
            continue
        path = np.zeros((n, n), dtype=np.int64)
        path[r, c] = 1
        col = c
        for _ in range(n):
            if not (mask[:, col] == 1).any():
                break
            row = int(np.argmax(mask[:, col] == 1))
            pcol = int(np.argmax(mask[row] == 2))
            path[row, col] = 1
            path[row, pcol] = 1
            col = pcol
        mask = mask - path
        mask[mask == 2] = 0
        rc[:] = False
        cc = (mask == 1).any(axis=0)
    return mask


def qap_round(P, Omega, l, D, W):
    """utils/qap_helpers.py:625-653: for each of the first l Nystrom columns
    (indices past the last column read the last, as JAX's clamped gather
    does) and its negative, the assignment minimising max(c) - c of the
    column as an l x l matrix, priced by trace(W Pi D Pi^T); the smaller
    minimum (the positive one unless the negative is strictly smaller).
    Returns (best, values of the 2 l candidates)."""
    from .common import reconstruct_from_sketch
    E, _ = reconstruct_from_sketch(Omega, P)
    r = E.shape[1]
    D = np.asarray(D, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    vals = []
    for sign in (1.0, -1.0):
        for i in range(l):
            cm = sign * E[1:, min(i, r - 1)].reshape(l, l)
            cm = cm.max() - cm
            perm = munkres(l, cm)
            vals.append(float(np.trace(W @ perm @ D @ perm.T)))
    vals = np.array(vals)
    mp, mn = vals[:l].min(), vals[l:].min()
    return (mp if mp < mn else mn), vals
