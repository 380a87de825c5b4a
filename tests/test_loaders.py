"""Data loaders and the log line (scripts/warm_start_maxcut_usbs.py:82-99,
utils/qap_helpers.py:26-154, solver/usbs.py:514-530): files written here in
each format round-trip to the matrices they were written from; TSPLIB
distances are checked against their closed forms (identical points, points
on the equator, the pseudo-Euclidean ATT rule); the log line carries the
reference's fields in its order and parses back."""
import math
import re

import numpy as np
import pytest
import scipy.io
import scipy.sparse

from usbs_amd import loaders, problems
from usbs_amd.qap_helpers import build_objective_matrix


def test_qaplib_round_trip(tmp_path):
    D, F = problems.qap_instance(6, 1)
    for first, second, swap in ((D, F, False), (F, D, True)):
        path = tmp_path / "inst.dat"
        body = "\n".join(" ".join("%d" % v for v in row) for row in np.vstack([first, second]))
        path.write_text("  6\n\n" + body + "\n")
        l, D2, W2, C = loaders.load_and_process_qap(str(path))
        assert l == 6 and np.array_equal(D2, D) and np.array_equal(W2, F)   # W is the sparser one
        C_ref = build_objective_matrix(D, F)
        assert np.array_equal(C.indices, C_ref.indices) and np.array_equal(C.data, C_ref.data)
    l, D3, W3, _ = loaders.load_and_process_qap(str(path), num_drop=2)
    assert l == 4 and np.array_equal(D3, D[:4, :4]) and np.array_equal(W3, F[:4, :4])


def _tsp_file(path, n, header, section, rows):
    text = "NAME : t\nTYPE : TSP\nDIMENSION : %d\n%s\n%s\n%s\nEOF\n" % (
        n, header, section, "\n".join(rows))
    path.write_text(text)


@pytest.mark.parametrize("fmt", ["FULL_MATRIX", "UPPER_ROW", "LOWER_COL", "UPPER_DIAG_ROW",
                                 "LOWER_DIAG_COL", "LOWER_DIAG_ROW", "UPPER_DIAG_COL"])
def test_tsplib_explicit_weights(tmp_path, fmt):
    n = 5
    rng = np.random.default_rng(2)
    D = rng.integers(1, 50, (n, n)).astype(float)
    D = np.triu(D, 1)
    D = D + D.T
    if fmt == "FULL_MATRIX":
        vals = D.ravel()
    elif fmt in ("UPPER_ROW", "LOWER_COL"):
        vals = D[np.triu_indices(n, 1)]
    elif fmt in ("UPPER_DIAG_ROW", "LOWER_DIAG_COL"):
        vals = D[np.triu_indices(n, 0)]
    else:
        vals = D[np.tril_indices(n, 0)]
    path = tmp_path / "t.tsp"
    _tsp_file(path, n, "EDGE_WEIGHT_TYPE : EXPLICIT\nEDGE_WEIGHT_FORMAT : %s" % fmt,
              "EDGE_WEIGHT_SECTION", [" ".join("%d" % v for v in vals)])
    l, D2, W, C = loaders.load_and_process_tsp(str(path))
    assert l == n and np.array_equal(D2, D)
    # the canonical tour, weight 1/2 each way
    assert np.allclose(W.sum(0), 1.0) and all(W[i, (i + 1) % n] == 0.5 for i in range(n))


def test_tsplib_coordinates(tmp_path):
    pts = [(0.0, 0.0), (3.0, 4.0), (10.0, 0.5), (3.0, 4.0)]
    rows = ["%d %g %g" % (i + 1, x, y) for i, (x, y) in enumerate(pts)]
    path = tmp_path / "c.tsp"
    _tsp_file(path, 4, "EDGE_WEIGHT_TYPE : EUC_2D", "NODE_COORD_SECTION", rows)
    _, D, _, _ = loaders.load_and_process_tsp(str(path))
    assert D[0, 1] == 5.0 and D[1, 3] == 0.0 and D[0, 2] == 10.0
    _tsp_file(path, 4, "EDGE_WEIGHT_TYPE : CEIL_2D", "NODE_COORD_SECTION", rows)
    _, D, _, _ = loaders.load_and_process_tsp(str(path))
    assert D[0, 2] == 11.0                                   # ceil(sqrt(100.25))
    _tsp_file(path, 4, "EDGE_WEIGHT_TYPE : ATT", "NODE_COORD_SECTION", rows)
    _, D, _, _ = loaders.load_and_process_tsp(str(path))
    assert D[0, 1] == math.ceil(math.sqrt(25 / 10.0))         # pseudo-Euclidean
    # GEO: points on the equator d whole degrees apart are RRR d pi/180 apart
    geo = ["1 0.0 10.0", "2 0.0 13.0", "3 0.0 10.0"]
    _tsp_file(path, 3, "EDGE_WEIGHT_TYPE : GEO", "NODE_COORD_SECTION", geo)
    _, D, _, _ = loaders.load_and_process_tsp(str(path))
    assert D[0, 1] == int(6378.388 * 3.0 * 3.141592 / 180.0 + 1.0)
    assert D[0, 2] == 1.0                                    # identical points: int(0 + 1)


def test_maxcut_mat_round_trip(tmp_path):
    W = problems.erdos_renyi(40, 0.2, seed=3, signed=True)
    gset = tmp_path / "Gset_G0.mat"
    scipy.io.savemat(str(gset), {"Problem": {"title": "g", "A": scipy.sparse.csc_matrix(W)}})
    assert abs(loaders.load_maxcut_mat(str(gset)) - W).max() == 0
    dimacs = tmp_path / "DIMACS10_x.mat"
    scipy.io.savemat(str(dimacs), {"Problem": {"title": "d", "name": "x",
                                               "A": scipy.sparse.csc_matrix(W)}})
    assert abs(loaders.load_maxcut_mat(str(dimacs)) - W).max() == 0
    with pytest.raises(ValueError):
        loaders.load_maxcut_mat(str(tmp_path / "other.mat"))


def test_log_line_fields():
    rec = dict(t=7.0, time=1.5, pen_dual_obj=10.25, cand_pen_dual_obj=11.5, lb_spec_est=-3.0,
               primal_obj=9.125, dual_obj=8.0, obj_ub=9.5, obj_gap=1e-05, infeas_gap=0.003,
               max_infeas=0.0125, callback_val=float("nan"))
    line = loaders.format_log_line(rec, start_epoch=1000.0)
    names = re.findall(r"(\w+): ", line)
    assert names == ["t", "curr_time", "pen_dual_obj", "cand_pen_dual_obj", "lb_spec_est",
                     "primal_obj", "dual_obj", "primal_obj_ub", "obj_gap", "infeas_gap",
                     "max_infeas", "callback_val"]
    m = re.search(r"t:\s(\d+) - curr_time:\s([\.0-9]+).*obj_gap:\s([-e\.0-9]+).*callback_val:\s(.+)$",
                  line)
    assert m.group(1) == "7" and float(m.group(2)) == 1001.5 and float(m.group(3)) == 1e-05
    assert m.group(4) == "None"
