"""Differentiable depth rendering of the TSDF, its autograd binding and
Gauss-Newton depth tracking.

No reference counterpart: the reference raycast (GPUVoxelGrid.cpp:561-715)
is not differentiable and the reference tracks with a point-to-plane ICP
linearisation (kokkosICP.cpp:69-115).  Here the depth of each pixel is the
first zero crossing of the trilinear TSDF along its ray, and its exact
derivative with respect to the 6-DoF pose and the voxel sdf values comes
from the implicit function theorem at the root (libtsdf_hip.so
tsdf_render / tsdf_render_backward / tsdf_track_normal_eq).

Poses are camera->world 4x4.  Pose composition and the 6x6 normal-equation
solve run in fp64 on the host (tsdf_se3_gauss_newton_step); the render runs
in fp32 on the GPU.  Perturbations are left-multiplicative,
T' = exp(xi^) T with xi = (v, w): translation first, rotation second.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import TsdfView, check


def _hat(w):
    return np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])


def se3_exp(xi) -> np.ndarray:
    """exp of xi = (v, w) as a 4x4 float64 matrix."""
    xi = np.asarray(xi, np.float64)
    v, w = xi[:3], xi[3:]
    th2 = float(w @ w)
    th = np.sqrt(th2)
    if th < 1e-5:
        A, B, Cc = 1 - th2 / 6, 0.5 - th2 / 24, 1 / 6 - th2 / 120
    else:
        A, B, Cc = np.sin(th) / th, (1 - np.cos(th)) / th2, (th - np.sin(th)) / (th2 * th)
    K = _hat(w)
    K2 = K @ K
    T = np.eye(4)
    T[:3, :3] = np.eye(3) + A * K + B * K2
    T[:3, 3] = (np.eye(3) + B * K + Cc * K2) @ v
    return T


def so3_left_jacobian(w) -> np.ndarray:
    w = np.asarray(w, np.float64)
    th2 = float(w @ w)
    th = np.sqrt(th2)
    K = _hat(w)
    if th < 1e-5:
        return np.eye(3) + 0.5 * K + K @ K / 6
    return np.eye(3) + (1 - np.cos(th)) / th2 * K + (th - np.sin(th)) / (th2 * th) * K @ K


def se3_left_jacobian(xi) -> np.ndarray:
    """6x6 left Jacobian of SE(3): exp((xi + e)^) = exp((J e)^) exp(xi^) to
    first order in e (Barfoot, State Estimation for Robotics, 7.1.5)."""
    xi = np.asarray(xi, np.float64)
    rho, phi = xi[:3], xi[3:]
    th2 = float(phi @ phi)
    th = np.sqrt(th2)
    J = so3_left_jacobian(phi)
    P, R = _hat(phi), _hat(rho)
    if th < 1e-4:
        Q = 0.5 * R + (P @ R + R @ P + P @ R @ P) / 6.0 - (P @ P @ R + R @ P @ P - 3 * P @ R @ P) / 24.0
    else:
        s, c = np.sin(th), np.cos(th)
        Q = (0.5 * R + (th - s) / th ** 3 * (P @ R + R @ P + P @ R @ P)
             + (th2 + 2 * c - 2) / (2 * th ** 4) * (P @ P @ R + R @ P @ P - 3 * P @ R @ P)
             + (2 * th - 3 * s + th * c) / (2 * th ** 5) * (P @ R @ P @ P + P @ P @ R @ P))
    out = np.zeros((6, 6))
    out[:3, :3] = J
    out[:3, 3:] = Q
    out[3:, 3:] = J
    return out


def view_of(h, w, intr) -> TsdfView:
    fx, fy, cx, cy = intr
    return TsdfView(int(h), int(w), float(fx), float(fy), float(cx), float(cy))


def _poses_f32(T):
    T = np.asarray(T, np.float64).reshape(-1, 4, 4)
    return np.ascontiguousarray(T.astype(np.float32).reshape(-1))


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def render(grid, T, h, w, intr, sdf=None, normals=False):
    """Depth [n, h, w] (and unit normals [n, h, w, 3]) of views T [n, 4, 4]."""
    T = np.asarray(T, np.float64).reshape(-1, 4, 4)
    n = T.shape[0]
    dev = grid.device
    depth = torch.empty((n, h, w), dtype=torch.float32, device=dev)
    nrm = torch.empty((n, h, w, 3), dtype=torch.float32, device=dev) if normals else None
    E = _poses_f32(T)
    cam = view_of(h, w, intr)
    check(grid._L.tsdf_render(grid._h, n, _fptr(E), C.byref(cam), _ptr(sdf), _ptr(depth), _ptr(nrm), _stream()),
          "tsdf_render")
    return (depth, nrm) if normals else depth


def render_backward(grid, T, h, w, intr, depth, grad_depth, sdf=None, want_sdf=True):
    """(grad_pose [n, 6] = dL/dxi of the left perturbation at T, grad_sdf
    [n_bricks, 512] or None)."""
    T = np.asarray(T, np.float64).reshape(-1, 4, 4)
    n = T.shape[0]
    dev = grid.device
    gp = torch.empty((n, 6), dtype=torch.float32, device=dev)
    gs = None
    if want_sdf:
        rows = sdf.shape[0] if sdf is not None else grid.size()
        gs = torch.zeros((max(rows, 1), 512), dtype=torch.float32, device=dev)
    E = _poses_f32(T)
    cam = view_of(h, w, intr)
    check(grid._L.tsdf_render_backward(grid._h, n, _fptr(E), C.byref(cam), _ptr(sdf), _ptr(depth.contiguous()),
                                       _ptr(grad_depth.contiguous()), _ptr(gp), _ptr(gs), _stream()),
          "tsdf_render_backward")
    return gp, gs


class RenderDepth(torch.autograd.Function):
    """depth = render(exp(xi^) T_base, sdf): differentiable in xi [n, 6]
    (float64, host) and sdf [n_bricks, 512] (float32, device; the stored
    values of the map, e.g. from GPUVoxelGrid.sdf_tensor())."""

    @staticmethod
    def forward(ctx, xi, sdf, grid, T_base, h, w, intr):
        xi_np = xi.detach().cpu().numpy().astype(np.float64).reshape(-1, 6)
        Tb = np.asarray(T_base, np.float64).reshape(-1, 4, 4)
        T = np.stack([se3_exp(x) @ Tk for x, Tk in zip(xi_np, Tb)])
        depth = render(grid, T, h, w, intr, sdf=sdf.detach().contiguous())
        ctx.grid, ctx.T, ctx.hw, ctx.intr, ctx.xi = grid, T, (h, w), intr, xi_np
        ctx.save_for_backward(sdf, depth)
        ctx.xi_device, ctx.xi_dtype = xi.device, xi.dtype
        return depth

    @staticmethod
    def backward(ctx, grad_depth):
        sdf, depth = ctx.saved_tensors
        h, w = ctx.hw
        gp, gs = render_backward(ctx.grid, ctx.T, h, w, ctx.intr, depth, grad_depth, sdf=sdf.detach().contiguous(),
                                 want_sdf=ctx.needs_input_grad[1])
        g_left = gp.double().cpu().numpy()
        # chain rule through T = exp(xi^) T_base: dL/dxi = J_l(xi)^T dL/d(left perturbation)
        gxi = np.stack([se3_left_jacobian(x).T @ g for x, g in zip(ctx.xi, g_left)])
        gxi = torch.from_numpy(gxi).to(device=ctx.xi_device, dtype=ctx.xi_dtype)
        return gxi, gs, None, None, None, None, None


def gauss_newton_step(normal_eq, T):
    """One fp64 Gauss-Newton update per view from normal_eq [n, 29] (host
    float64): returns (T_new [n, 4, 4], delta [n, 6])."""
    L = _lib.load()
    ne = np.ascontiguousarray(normal_eq, np.float64).reshape(-1, 29)
    T = np.ascontiguousarray(T, np.float64).reshape(-1, 4, 4)
    out = np.empty_like(T)
    delta = np.zeros((T.shape[0], 6))
    dp = C.POINTER(C.c_double)
    for k in range(T.shape[0]):
        H = np.ascontiguousarray(ne[k, :21])
        b = np.ascontiguousarray(ne[k, 21:27])
        Tin = np.ascontiguousarray(T[k].reshape(16))
        Tout = np.empty(16)
        d = np.empty(6)
        st = L.tsdf_se3_gauss_newton_step(H.ctypes.data_as(dp), b.ctypes.data_as(dp), Tin.ctypes.data_as(dp),
                                          Tout.ctypes.data_as(dp), d.ctypes.data_as(dp))
        if st != 0:  # degenerate view: keep its pose
            out[k] = T[k]
            continue
        out[k] = Tout.reshape(4, 4)
        delta[k] = d
    return out, delta


def track(grid, depth_obs, T_init, intr, iters=20, max_residual=0.1):
    """Gauss-Newton depth tracking of n views against observed depth
    depth_obs [n, h, w] (device float32, metres, 0 = none) from T_init
    [n, 4, 4].  Returns (T [n, 4, 4] float64, history: per iteration the
    per-view RMS residual and inlier count)."""
    n, h, w = depth_obs.shape
    T = np.asarray(T_init, np.float64).reshape(n, 4, 4).copy()
    ne = torch.empty((n, 29), dtype=torch.float64, device=grid.device)
    cam = view_of(h, w, intr)
    dobs = depth_obs.contiguous()
    hist = []
    for _ in range(iters):
        E = _poses_f32(T)
        check(grid._L.tsdf_track_normal_eq(grid._h, n, _fptr(E), C.byref(cam), _ptr(dobs), float(max_residual),
                                           _ptr(ne), None, _stream()), "tsdf_track_normal_eq")
        host = ne.cpu().numpy()
        cnt = np.maximum(host[:, 28], 1.0)
        hist.append({"rms": np.sqrt(host[:, 27] / cnt), "count": host[:, 28].copy()})
        T, _ = gauss_newton_step(host, T)
    return T, hist
