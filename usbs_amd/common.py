"""SDPState and its scaling, mirroring utils/common.py:11-96.

Host-side bookkeeping only (no arithmetic of the hot path): the state tuple
with the reference's 15 fields and the SCALE_* conversions.  C is held as a
``usbs_amd.sparse.BCOO`` (or any scipy sparse matrix).
"""
from collections import namedtuple

import numpy as np

from .sparse import BCOO

SDPState = namedtuple("SDPState", ["C", "A_indices", "A_data", "b", "b_ineq_mask", "X", "P",
                                   "Omega", "y", "z", "tr_X", "primal_obj", "SCALE_C", "SCALE_X",
                                   "SCALE_A"])


def _scale_C(C, s):
    if isinstance(C, BCOO):
        return BCOO((C.data * s, C.indices), C.shape)
    return (C * s).tocsr()


def scale_sdp_state(s):
    """utils/common.py:29-61."""
    A_indices = np.asarray(s.A_indices)
    SA = np.asarray(s.SCALE_A, dtype=np.float64)
    return s._replace(
        C=_scale_C(s.C, s.SCALE_C),
        A_data=np.asarray(s.A_data, dtype=np.float64) * SA[A_indices[:, 0]],
        b=s.b * s.SCALE_X * SA,
        X=None if s.X is None else s.X * s.SCALE_X,
        P=None if s.P is None else s.P * s.SCALE_X,
        z=s.z * SA * s.SCALE_X,
        tr_X=s.tr_X * s.SCALE_X,
        primal_obj=s.primal_obj * s.SCALE_X * s.SCALE_C)


def unscale_sdp_state(s):
    """utils/common.py:64-96."""
    A_indices = np.asarray(s.A_indices)
    SA = np.asarray(s.SCALE_A, dtype=np.float64)
    return s._replace(
        C=_scale_C(s.C, 1.0 / s.SCALE_C),
        A_data=np.asarray(s.A_data, dtype=np.float64) / SA[A_indices[:, 0]],
        b=(s.b / SA) / s.SCALE_X,
        X=None if s.X is None else s.X / s.SCALE_X,
        P=None if s.P is None else s.P / s.SCALE_X,
        z=(s.z / SA) / s.SCALE_X,
        tr_X=s.tr_X / s.SCALE_X,
        primal_obj=s.primal_obj / (s.SCALE_X * s.SCALE_C))
