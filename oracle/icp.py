"""numpy face of the ICP part of the oracle (icp_oracle.c) and the
reference's ICP loop -- TEST INFRASTRUCTURE ONLY, imported by tests/ as the
checker.

kokkosICP(d1, d2, tx, ty, intr, ...) restates kokkosICP.cpp:23-191 on the
CPU: the per-pixel kernels and the sums come from icp_oracle.c (reference
order, RMSE summed in float), the rotation from numpy's SVD of the centred
cross-covariance as the reference takes it from Eigen's JacobiSVD
(R = V U^T, kokkosICP.cpp:173-177).
"""
from __future__ import annotations

import numpy as np

from . import lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def depth_to_points(depth, intr):
    d = _c(depth, np.uint16)
    h, w = d.shape
    out = np.zeros((h, w, 3), np.float32)
    lib().orc_depth_to_points(d.ctypes.data, h, w, *(float(v) for v in intr), out.ctypes.data)
    return out


def normals(pcd):
    p = _c(pcd, np.float32)
    h, w = p.shape[:2]
    out = np.zeros_like(p)
    lib().orc_normals(p.ctypes.data, h, w, out.ctypes.data)
    return out


def associate(x, y, yn, Txy, intr, icp_type=0, n_search=5, max_dist=0.075):
    x, y = _c(x, np.float32), _c(y, np.float32)
    yn = _c(yn, np.float32) if yn is not None else np.zeros_like(y)
    h, w = x.shape[:2]
    T = _c(np.asarray(Txy, np.float32).reshape(16), np.float32)
    out = np.zeros_like(x)
    lib().orc_icp_associate(x.ctypes.data, y.ctypes.data, yn.ctypes.data, T.ctypes.data, h, w,
                            *(float(v) for v in intr), int(icp_type), int(n_search), float(max_dist),
                            out.ctypes.data)
    return out


def sums(x, xn, Txy):
    """[points, correspondences, centre x (3), centre x_nearest (3), centred
    H (9, row-major), RMSE]."""
    x, xn = _c(x, np.float32), _c(xn, np.float32)
    h, w = x.shape[:2]
    T = _c(np.asarray(Txy, np.float64).reshape(16), np.float64)
    out = np.zeros(18, np.float64)
    lib().orc_icp_sums(x.ctypes.data, xn.ctypes.data, T.ctypes.data, h, w, out.ctypes.data)
    return out


def kokkosICP(d1, d2, tx, ty, intr, icp_type=0, n_it=10, n_search=5, max_dist=0.075):
    """(T 4x4, lowest RMSE, best fitness, per-iteration records)."""
    x = depth_to_points(d1, intr)
    y = depth_to_points(d2, intr)
    yn = normals(y)
    T = np.linalg.inv(np.asarray(ty, np.float64)) @ np.asarray(tx, np.float64)
    lowest, best_fit, best = np.float32(1e6), 0.0, np.eye(4)
    records = []
    for _ in range(n_it):
        xn = associate(x, y, yn, T.astype(np.float32), intr, icp_type, n_search, max_dist)
        s = sums(x, xn, T)
        if s[1] == 0:
            break
        rmse, fit = np.float32(s[17]), s[1] / s[0]
        best_fit = max(best_fit, fit)
        U, _, Vt = np.linalg.svd(s[8:17].reshape(3, 3))
        R = Vt.T @ U.T
        T = np.eye(4)
        T[:3, :3] = R
        T[:3, 3] = s[5:8] - R @ s[2:5]
        records.append({"rmse": float(rmse), "fitness": fit, "T": T.copy()})
        if rmse < lowest:
            lowest, best = rmse, T.copy()
    return best, float(lowest), float(best_fit), records
