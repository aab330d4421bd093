"""Reference-element tables on the host (for input generation only).

Gauss points and shape functions as FEM/mathlib.cpp (:1112-1215 tables,
:1285 tet5, :1661-2302 shape functions) and the point order of
CElement::SetGaussPoint (FEM/fem_ele.cpp:813).  The device kernels receive
the same tables from the library (pattern.cpp build_gauss_table).
"""
from __future__ import annotations

import numpy as np

from .mesh import HEX8, LINE2, QUAD4, TET4

_PKT = {1: [0.0], 2: [0.577350269189626, -0.577350269189626],
        3: [0.774596669241483, 0.0, -0.774596669241483],
        4: [0.861136311594053, 0.339981043584856, -0.339981043584856, -0.861136311594053]}
_TET5 = [(0.25, 0.25, 0.25), (0.166666666666667,) * 3,
         (0.5, 0.166666666666667, 0.166666666666667),
         (0.166666666666667, 0.5, 0.166666666666667),
         (0.166666666666667, 0.166666666666667, 0.5)]


def unit_points(etype: int, ng: int = 3) -> np.ndarray:
    p = _PKT[ng]
    if etype == LINE2:
        return np.array([[a, 0, 0] for a in p], float)
    if etype == QUAD4:
        return np.array([[p[g // ng], p[g % ng], 0] for g in range(ng * ng)], float)
    if etype == HEX8:
        pts = []
        for g in range(ng ** 3):
            r, s = g // (ng * ng), g % (ng * ng)
            t, s = s % ng, s // ng
            pts.append([p[r], p[s], p[t]])
        return np.array(pts, float)
    if etype == TET4:
        return np.array(_TET5, float)
    raise ValueError(etype)


def shape_values(etype: int, u: np.ndarray) -> np.ndarray:
    r, s, t = u[:, 0], u[:, 1], u[:, 2]
    if etype == LINE2:
        return 0.5 * np.stack([1 - r, 1 + r], 1)
    if etype == QUAD4:
        return 0.25 * np.stack([(1 + r) * (1 + s), (1 - r) * (1 + s), (1 - r) * (1 - s),
                                (1 + r) * (1 - s)], 1)
    if etype == HEX8:
        sr = np.array([1, -1, -1, 1, 1, -1, -1, 1])
        ss = np.array([1, 1, -1, -1, 1, 1, -1, -1])
        st = np.array([1, 1, 1, 1, -1, -1, -1, -1])
        return 0.125 * ((1 + sr * r[:, None]) * (1 + ss * s[:, None]) * (1 + st * t[:, None]))
    if etype == TET4:
        return np.stack([1 - r - s - t, r, s, t], 1)
    raise ValueError(etype)


def shape_tables(etype: int, ng: int = 3):
    u = unit_points(etype, ng)
    return shape_values(etype, u), u
