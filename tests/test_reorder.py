"""Node renumbering of the host-side problem (usbs_amd.reorder): the permuted
SDP is the same SDP (CPU checks; the GPU solve is compared in
tests/test_gpu_usbs.py)."""
import numpy as np

from usbs_amd import maxcut_helpers, problems
from usbs_amd.reorder import permute_state, rcm_permutation, unpermute_rows


def _state(n=60, seed=5):
    W = problems.random_geometric(n, 0.25, seed)
    Om = problems.gaussian_matrix(n, 4, 1)
    return maxcut_helpers.initialize_state(W, 4, Omega=Om), W


def test_rcm_narrows_a_random_numbering():
    st, W = _state(400, 2)
    perm = rcm_permutation(st.C, st.A_indices, 400)
    assert np.array_equal(np.sort(perm), np.arange(400))
    i, j = np.asarray(st.C.indices).T
    pos = np.empty(400, dtype=np.int64)
    pos[perm] = np.arange(400)
    assert np.abs(pos[i] - pos[j]).max() < 0.5 * np.abs(i - j).max()


def test_permuted_state_is_the_same_sdp():
    """C' = P C P^T entry by entry, A*(y) permutes the same way, Omega and q0
    rows follow their nodes, and rows come back with unpermute_rows."""
    st, _ = _state()
    n = st.C.shape[0]
    rng = np.random.default_rng(0)
    perm = rng.permutation(n)
    q0 = problems.unit_start_vector(n, 3)
    st2, q02, p = permute_state(st, q0, perm=perm)
    assert np.array_equal(p, perm)
    Cd = st.C.to_scipy().toarray()
    C2 = st2.C.to_scipy().toarray()
    np.testing.assert_array_equal(C2, Cd[np.ix_(perm, perm)])
    y = rng.standard_normal(st.b.size)
    # A*(y) as a dense matrix from the triples, before and after
    def adj(Ai, Ad):
        M = np.zeros((n, n))
        np.add.at(M, (Ai[:, 1], Ai[:, 2]), Ad * y[Ai[:, 0]])
        return M
    Ai = np.asarray(st.A_indices).reshape(-1, 3)
    np.testing.assert_array_equal(adj(np.asarray(st2.A_indices), st2.A_data),
                                  adj(Ai, st.A_data)[np.ix_(perm, perm)])
    np.testing.assert_array_equal(st2.Omega, np.asarray(st.Omega)[perm])
    np.testing.assert_array_equal(q02, q0[perm])
    np.testing.assert_array_equal(unpermute_rows(st2.Omega, perm), st.Omega)
    # the spectrum of C - A*(y) does not depend on the numbering
    ev1 = np.linalg.eigvalsh(Cd - adj(Ai, st.A_data))
    ev2 = np.linalg.eigvalsh(C2 - adj(np.asarray(st2.A_indices), st2.A_data))
    np.testing.assert_allclose(ev2, ev1, rtol=0, atol=1e-12 * np.abs(ev1).max())
