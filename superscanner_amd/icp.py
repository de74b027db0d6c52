"""Projective ICP between two depth frames and the point / normal maps it is
built on, on the GPU through the C ABI (include/tsdf.h):

  CreatePcdFromDepthView(depth, intr)    imageutil.cpp:89-111
  CalculateNormalsFromPCDView(pcd)       imageutil.cpp:116-157
  associate(x, y, y_normals, Txy, ...)   kokkosICP.cpp:69-115 ("Closest Point Kernel")
  moments(x, x_nearest, Txy)             the sums of kokkosICP.cpp:117-171
  kokkosICP(d1, d2, tx, ty, intr, ...)   kokkosICP.cpp:23-191 (the depth-view overload)

Maps are device float32 tensors [H, W, 3] (row-major; the reference's Views
are LayoutLeft Kvec3f).  Intrinsics are (fx, fy, cx, cy) in float; the
reference's ci = H - cy, cj = cx convention is applied inside.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import TsdfIcpParams, TsdfView, check

PointToPoint = _lib.ICP_POINT_TO_POINT
PointToPlane = _lib.ICP_POINT_TO_PLANE


def _stream(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _view(h, w, intr):
    fx, fy, cx, cy = (float(v) for v in intr)
    return TsdfView(int(h), int(w), fx, fy, cx, cy)


def params(icp_type=PointToPoint, n_it=None, n_search=None, max_dist=None) -> TsdfIcpParams:
    """The reference's constants (n_it 10, n_search 5, max_dist 0.075 m) unless given."""
    p = TsdfIcpParams()
    _lib.load().tsdf_icp_default_params(C.byref(p))
    p.type = int(icp_type)
    if n_it is not None:
        p.n_iterations = int(n_it)
    if n_search is not None:
        p.n_search = int(n_search)
    if max_dist is not None:
        p.max_dist = float(max_dist)
    return p


def _dev_u16(depth):
    if isinstance(depth, torch.Tensor):
        if depth.dtype not in (torch.int16, torch.uint16) or not depth.is_cuda:
            raise ValueError("depth must be a device uint16 (or int16 view) tensor [H, W]")
        return depth.contiguous()
    a = np.ascontiguousarray(depth, dtype=np.uint16)
    return torch.from_numpy(a.view(np.int16)).cuda()


def _dev_pts(p, h=None, w=None):
    if not (isinstance(p, torch.Tensor) and p.is_cuda and p.dtype == torch.float32 and p.dim() == 3
            and p.shape[2] == 3):
        raise ValueError("point / normal maps are device float32 tensors [H, W, 3]")
    if h is not None and tuple(p.shape[:2]) != (h, w):
        raise ValueError(f"map of shape {tuple(p.shape[:2])}, expected {(h, w)}")
    return p.contiguous()


def CreatePcdFromDepthView(depth, intr, stream=None) -> torch.Tensor:
    """imageutil.cpp:89-111: ((j - cx) / fx d, (i - (H - cy)) / fy d, d), d = depth / 1000."""
    d = _dev_u16(depth)
    h, w = d.shape
    out = torch.empty((h, w, 3), dtype=torch.float32, device=d.device)
    v = _view(h, w, intr)
    check(_lib.load().tsdf_depth_to_points(C.c_void_p(d.data_ptr()), C.byref(v), C.c_void_p(out.data_ptr()),
                                           _stream(stream)), "tsdf_depth_to_points")
    return out


def CalculateNormalsFromPCDView(pcd, stream=None) -> torch.Tensor:
    """imageutil.cpp:116-157: (0, 0, 1) on the border and next to z == 0."""
    p = _dev_pts(pcd)
    h, w = p.shape[:2]
    out = torch.empty_like(p)
    check(_lib.load().tsdf_points_to_normals(C.c_void_p(p.data_ptr()), h, w, C.c_void_p(out.data_ptr()),
                                             _stream(stream)), "tsdf_points_to_normals")
    return out


def associate(x, y, y_normals, Txy, intr, icp_type=PointToPoint, n_search=5, max_dist=0.075,
              stream=None) -> torch.Tensor:
    """The closest point of y for every point of x under the float transform
    Txy (kokkosICP.cpp:69-115); (0, 0, 0) where none."""
    x = _dev_pts(x)
    h, w = x.shape[:2]
    y = _dev_pts(y, h, w)
    yn = _dev_pts(y_normals, h, w) if y_normals is not None else None
    T = np.ascontiguousarray(np.asarray(Txy, np.float32).reshape(16))
    p = params(icp_type, n_search=n_search, max_dist=max_dist)
    v = _view(h, w, intr)
    out = torch.empty_like(x)
    check(_lib.load().tsdf_icp_associate(C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()),
                                         C.c_void_p(yn.data_ptr()) if yn is not None else None,
                                         T.ctypes.data_as(C.POINTER(C.c_float)), C.byref(v), C.byref(p),
                                         C.c_void_p(out.data_ptr()), _stream(stream)), "tsdf_icp_associate")
    return out


def moments(x, x_nearest, Txy, stream=None) -> np.ndarray:
    """[points of x, correspondences, sum x (3), sum x_nearest (3),
    sum x x_nearest^T (9, row-major), sum |x_nearest - Txy x|^2] in double."""
    x = _dev_pts(x)
    h, w = x.shape[:2]
    xn = _dev_pts(x_nearest, h, w)
    T = np.ascontiguousarray(np.asarray(Txy, np.float64).reshape(16))
    out = np.zeros(_lib.ICP_MOMENTS, np.float64)
    dp = C.POINTER(C.c_double)
    check(_lib.load().tsdf_icp_moments(C.c_void_p(x.data_ptr()), C.c_void_p(xn.data_ptr()), T.ctypes.data_as(dp),
                                       h, w, out.ctypes.data_as(dp), _stream(stream)), "tsdf_icp_moments")
    return out


def moments_noalloc(x, x_nearest, Txy, stream=None) -> np.ndarray:
    """The sums of ``moments`` (bit-identical) through the entry point that
    allocates no device memory (tsdf_icp_moments_noalloc)."""
    x = _dev_pts(x)
    h, w = x.shape[:2]
    xn = _dev_pts(x_nearest, h, w)
    T = np.ascontiguousarray(np.asarray(Txy, np.float64).reshape(16))
    out = np.zeros(_lib.ICP_MOMENTS, np.float64)
    dp = C.POINTER(C.c_double)
    check(_lib.load().tsdf_icp_moments_noalloc(C.c_void_p(x.data_ptr()), C.c_void_p(xn.data_ptr()),
                                               T.ctypes.data_as(dp), h, w, out.ctypes.data_as(dp),
                                               _stream(stream)), "tsdf_icp_moments_noalloc")
    return out


def kokkosICP(d1, d2, tx, ty, intr, ICPType=PointToPoint, n_it=10, n_search=5, max_dist=0.075, stream=None):
    """kokkosICP(space, d1, d2, tx, ty, ICPType, &rsme), kokkosICP.cpp:23-191:
    the transform taking the points of d1 (pose tx) onto those of d2 (pose
    ty), as (T 4x4 float64, rmse, fitness)."""
    a, b = _dev_u16(d1), _dev_u16(d2)
    if a.shape != b.shape:
        raise ValueError("depth images of different sizes")  # the reference only prints (:30-33)
    h, w = a.shape
    v = _view(h, w, intr)
    p = params(ICPType, n_it=n_it, n_search=n_search, max_dist=max_dist)
    dp = C.POINTER(C.c_double)
    TX = np.ascontiguousarray(np.asarray(tx, np.float64).reshape(16))
    TY = np.ascontiguousarray(np.asarray(ty, np.float64).reshape(16))
    T = np.zeros(16, np.float64)
    rmse, fit = C.c_float(0), C.c_float(0)
    check(_lib.load().tsdf_icp(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), TX.ctypes.data_as(dp),
                               TY.ctypes.data_as(dp), C.byref(v), C.byref(p), T.ctypes.data_as(dp),
                               C.byref(rmse), C.byref(fit), _stream(stream)), "tsdf_icp")
    return T.reshape(4, 4), float(rmse.value), float(fit.value)


def release() -> None:
    """Frees the workspace kokkosICP keeps on the current device between
    calls (tsdf_icp_release); the next call allocates it again."""
    check(_lib.load().tsdf_icp_release(), "tsdf_icp_release")
