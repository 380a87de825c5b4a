"""Benchmark data loaders and the solver's log line, mirroring the reference.

* ``load_maxcut_mat``: Gset and DIMACS10 graphs from MATLAB files, as
  scripts/warm_start_maxcut_usbs.py:82-99 loads them (the adjacency is field
  1 of the ``Problem`` struct for Gset, field 2 for DIMACS10).  MATLAB v7.3
  files are HDF5, which the reference reads with ``mat73``; that package is
  not part of this environment, so such files raise.
* ``load_and_process_qap``: QAPLIB ``.dat`` (utils/qap_helpers.py:26-45).
* ``load_and_process_tsp``: TSPLIB ``.tsp`` as a QAP with the canonical tour
  as flow (utils/qap_helpers.py:48-154), in the formats the reference reads.
* ``format_log_line``: the line usbs() prints at every condition check
  (solver/usbs.py:514-530), from a record of ``usbs(..., log=[])``.

All of this is host-side setup, as in the reference.
"""
import math

import numpy as np
import scipy.io
import scipy.sparse
from scipy.spatial.distance import pdist, squareform

from .qap_helpers import build_objective_matrix


def load_maxcut_mat(path, kind=None):
    """The weighted adjacency (scipy CSC) of a Gset or DIMACS10 graph.  kind
    ("Gset" / "DIMACS") defaults to the one named in the path, as the
    reference decides."""
    if kind is None:
        kind = "Gset" if "Gset" in path else "DIMACS" if "DIMACS" in path else None
    if kind not in ("Gset", "DIMACS"):
        raise ValueError("Unknown path type")
    try:
        problem = scipy.io.loadmat(path)
    except NotImplementedError as e:  # MATLAB v7.3 (HDF5)
        raise NotImplementedError("MATLAB v7.3 files need an HDF5 reader (the reference uses "
                                  "mat73), which this environment lacks") from e
    C = problem["Problem"][0][0][1 if kind == "Gset" else 2]
    return scipy.sparse.csc_matrix(C)


def load_and_process_qap(fname, num_drop=0):
    """(l, D, W, C) from a QAPLIB file: n, then two n x n matrices; the sparser
    one is the flow W (by the reference's convention); the last num_drop
    facilities and locations dropped; C = build_objective_matrix(D, W)."""
    with open(fname, "r") as f:
        tok = f.read().strip().split()
    n = int(tok[0])
    M = np.array([float(v) for v in tok[1:]]).reshape(2 * n, n)
    D, W = M[:n], M[n:]
    if np.count_nonzero(D) < np.count_nonzero(W):
        D, W = W, D
    n_out = n - num_drop
    D, W = D[:n_out, :n_out], W[:n_out, :n_out]
    return n_out, D, W, build_objective_matrix(D, W)


def _geo(p1, p2):
    # TSPLIB's GEO distance with the constants the reference uses
    PI, RRR = 3.141592, 6378.388

    def rad(x):
        deg = np.round(x)
        return PI * (deg + 5.0 * (x - deg) / 3.0) / 180.0

    lat1, long1, lat2, long2 = rad(p1[0]), rad(p1[1]), rad(p2[0]), rad(p2[1])
    q1, q2, q3 = np.cos(long1 - long2), np.cos(lat1 - lat2), np.cos(lat1 + lat2)
    return int(RRR * np.arccos(0.5 * (q2 * (1.0 + q1) - q3 * (1.0 - q1))) + 1.0)


def _att(p1, p2):
    xd, yd = p1[0] - p2[0], p1[1] - p2[1]
    return np.ceil(np.sqrt((xd ** 2 + yd ** 2) / 10.0))


def load_and_process_tsp(fname, num_drop=0):
    """(l, D, W, C) from a TSPLIB file: D from NODE_COORD_SECTION (EUC_2D,
    CEIL_2D, ATT, GEO) or EDGE_WEIGHT_SECTION (FULL_MATRIX, UPPER_ROW /
    LOWER_COL, UPPER_DIAG_ROW / LOWER_DIAG_COL, LOWER_DIAG_ROW /
    UPPER_DIAG_COL); W the symmetric canonical tour with weights 1/2."""
    spec = {}
    D, coords = None, None
    with open(fname, "r") as f:
        lines = iter(f)
        for line in lines:
            parts = [p.strip() for p in line.split(":")]
            if len(parts) == 2:
                spec[parts[0]] = parts[1]
                continue
            if len(parts) != 1:
                raise ValueError("Something went wrong when reading file.")
            n = int(spec["DIMENSION"])
            if parts[0] == "NODE_COORD_SECTION":
                coords = []
                for row in lines:
                    row = row.strip()
                    if row == "EOF":
                        break
                    _, x, y = row.split()
                    coords.append([float(x), float(y)])
                if len(coords) != n:
                    raise ValueError("NODE_COORD_SECTION has %d nodes, DIMENSION %d" % (len(coords), n))
                break
            if parts[0] == "EDGE_WEIGHT_SECTION":
                vals = []
                for row in lines:
                    row = row.strip()
                    if row in ("EOF", "DISPLAY_DATA_SECTION"):
                        break
                    vals += [int(v) for v in row.split()]
                vals = np.array(vals, dtype=np.float64)
                fmt = spec["EDGE_WEIGHT_FORMAT"]
                D = np.zeros((n, n))
                if fmt == "FULL_MATRIX":
                    D[:] = vals.reshape(n, n)
                elif fmt in ("UPPER_ROW", "LOWER_COL"):
                    D[np.triu_indices(n, 1)] = vals
                    D = D + D.T
                elif fmt in ("LOWER_ROW", "UPPER_COL"):
                    raise NotImplementedError("edge weight format type not implemented")
                elif fmt in ("UPPER_DIAG_ROW", "LOWER_DIAG_COL"):
                    D[np.triu_indices(n, 0)] = vals
                    D = D + D.T
                elif fmt in ("LOWER_DIAG_ROW", "UPPER_DIAG_COL"):
                    D[np.tril_indices(n, 0)] = vals
                    D = D + D.T
                else:
                    raise ValueError("Unsupported EDGE_WEIGHT_FORMAT.")
                break
    if D is None:
        if coords is None:
            raise ValueError("no NODE_COORD_SECTION or EDGE_WEIGHT_SECTION")
        coords = np.array(coords)
        kind = spec["EDGE_WEIGHT_TYPE"]
        if kind == "GEO":
            D = squareform(pdist(coords, _geo))
        elif kind == "ATT":
            D = squareform(pdist(coords, _att))
        elif kind == "EUC_2D":
            D = np.round(squareform(pdist(coords, "euclidean")))
        elif kind == "CEIL_2D":
            D = np.ceil(squareform(pdist(coords, "euclidean")))
        else:
            raise ValueError("Unsupported EDGE_WEIGHT_TYPE.")
    n_out = n - num_drop
    D = D[:n_out, :n_out]
    W = np.zeros((n_out, n_out))
    for i in range(n_out):
        W[i, (i + 1) % n_out] += 0.5
        W[(i + 1) % n_out, i] += 0.5
    return n_out, D, W, build_objective_matrix(D, W)


def _num(v):
    if v is None or (isinstance(v, float) and math.isnan(v)):
        return "None"
    return repr(float(v))


def format_log_line(rec, start_epoch=0.0):
    """The reference's condition-check line (solver/usbs.py:514-530) for a
    record of usbs(..., log=[]): its curr_time is wall-clock time, here
    start_epoch (time.time() when the solve began) plus the record's seconds
    since the start; a record without a rounding callback prints None."""
    return ("t: %d - curr_time: %s - pen_dual_obj: %s - cand_pen_dual_obj: %s - lb_spec_est: %s"
            " - primal_obj: %s - dual_obj: %s - primal_obj_ub: %s - obj_gap: %s - infeas_gap: %s"
            " - max_infeas: %s - callback_val: %s"
            % (int(rec["t"]), _num(start_epoch + rec["time"]), _num(rec["pen_dual_obj"]),
               _num(rec["cand_pen_dual_obj"]), _num(rec["lb_spec_est"]), _num(rec["primal_obj"]),
               _num(rec["dual_obj"]), _num(rec["obj_ub"]), _num(rec["obj_gap"]),
               _num(rec["infeas_gap"]), _num(rec["max_infeas"]), _num(rec["callback_val"])))
