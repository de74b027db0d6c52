"""numpy restatement of the reintegration scheduler's host arithmetic --
TEST INFRASTRUCTURE ONLY, imported by tests/ as the checker.

  mat_to_dof   MattoDof, src/utils/matrixutil.cpp:38-64
  dof_to_mat   getT (gettrans, getRz, getRy, getRx), src/utils/matrixutil.h:20-63
  diff         pandia_integration::diff, src/core/integrate.cpp:211-216
  select       updateReintegrationBuffer, src/core/integrate.cpp:223-234
"""
from __future__ import annotations

import numpy as np


def mat_to_dof(T):
    T = np.asarray(T, np.float64).reshape(4, 4)
    return np.array([np.arctan2(T[2, 1], T[2, 2]), -np.arcsin(T[2, 0]), np.arctan2(T[1, 0], T[0, 0]),
                     T[0, 3], T[1, 3], T[2, 3]])


def dof_to_mat(x):
    x = np.asarray(x, np.float64)
    c, s = np.cos, np.sin
    tr = np.array([[1, 0, 0, x[3]], [0, 1, 0, x[4]], [0, 0, 1, x[5]], [0, 0, 0, 1.0]])
    rz = np.array([[c(x[2]), -s(x[2]), 0, 0], [s(x[2]), c(x[2]), 0, 0], [0, 0, 1, 0], [0, 0, 0, 1.0]])
    ry = np.array([[c(x[1]), 0, s(x[1]), 0], [0, 1, 0, 0], [-s(x[1]), 0, c(x[1]), 0], [0, 0, 0, 1.0]])
    rx = np.array([[1, 0, 0, 0], [0, c(x[0]), -s(x[0]), 0], [0, s(x[0]), c(x[0]), 0], [0, 0, 0, 1.0]])
    return ((tr @ rz) @ ry) @ rx


def diff(world, integrated):
    d = np.asarray(world, np.float64) - np.asarray(integrated, np.float64)
    return float(np.linalg.norm(np.concatenate([d[:3], 2 * d[3:]])))


def select(worlds, integrateds, treint):
    """(new list order, reintegration buffer length)."""
    d = [diff(a, b) for a, b in zip(worlds, integrateds)]
    order = sorted(range(len(d)), key=lambda k: -d[k])  # stable: ties keep list order
    m = 0
    while m < len(order) and not d[order[m]] < treint:
        m += 1
    return order, m
