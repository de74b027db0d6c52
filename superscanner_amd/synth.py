"""Seeded synthetic scenes, trajectories and depth frames.

The workloads (sizes, scene content, trajectories) are the five benchmark
configurations of this path:

  plain_small   640x480 x 30 frames, plane z=0 + sphere r=0.3 m + box 0.4 m,
                circular trajectory r=1.5 m, N(0, 1 mm) noise, fx=fy=525,
                cx=319.5, cy=239.5, vl 5 mm, trunc 4 voxels, 2^20 buckets
  room_large    848x480 x 600 frames, seeded random room (6 boxes, 2 spheres,
                floor/walls), random walk 5 mm / 0.5 deg per step, vl 4 mm,
                trunc 5 voxels, 2^22 buckets, max weight 128
  reintegrate   room_large scene, 64 keyframes, poses perturbed by
                N(0, 1 cm / 0.5 deg), per-keyframe valid fraction U(0.3, 1)
  diff_raycast  16 views at 320x240 over the plain_small TSDF
  fp16_large    1280x720 x 2000 frames of a 12 m x 8 m multi-room
                (scene_multiroom / traj_multiroom), vl 2 mm, trunc 4 voxels,
                2^24 buckets, fp16 storage

Depth is rendered by libtsdf_synth.so (double-precision ray casting with the
reference camera convention ci = H - cy, cj = cx).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtsdf_synth.so")

PLANE, SPHERE, BOX, ROOM = 0, 1, 2, 3


class Prim(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("a", C.c_double * 6)]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build with `make -C {_HERE}`")
        L = C.CDLL(LIB_PATH)
        L.syn_render_depth.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int, C.c_int,
                                       C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_uint64, C.c_double, C.c_int, C.c_void_p]
        L.syn_render_depth.restype = C.c_int
        _lib = L
    return _lib


def prims_array(prims):
    arr = (Prim * max(1, len(prims)))()
    for k, (kind, vals) in enumerate(prims):
        arr[k].kind = kind
        for m, v in enumerate(vals):
            arr[k].a[m] = float(v)
    return arr, len(prims)


def render_depth(prims, pose, H, W, intr, noise_mm=1.0, max_range_m=10.0, seed=0, keep=1.0,
                 tile=16):
    """uint16 depth in mm of `prims` seen from `pose` (4x4 camera->world)."""
    arr, n = prims_array(prims)
    P = np.ascontiguousarray(np.asarray(pose, np.float64).reshape(16))
    out = np.zeros((H, W), np.uint16)
    fx, fy, cx, cy = intr
    st = _load().syn_render_depth(C.cast(arr, C.c_void_p), n, P.ctypes.data_as(C.POINTER(C.c_double)),
                                  H, W, fx, fy, cx, cy, noise_mm, max_range_m,
                                  C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF), keep, tile, out.ctypes.data)
    if st != 0:
        raise ValueError("syn_render_depth rejected its arguments")
    return out


# --------------------------------------------------------------------------
# poses
# --------------------------------------------------------------------------

def look_at(eye, target, up=(0.0, 0.0, 1.0)):
    """Camera->world pose with camera z toward `target`, x = z cross up,
    y = z cross x (image rows grow downward)."""
    eye = np.asarray(eye, np.float64)
    z = np.asarray(target, np.float64) - eye
    z /= np.linalg.norm(z)
    x = np.cross(z, np.asarray(up, np.float64))
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    T = np.eye(4)
    T[:3, 0], T[:3, 1], T[:3, 2], T[:3, 3] = x, y, z, eye
    return T


def axis_angle(axis, ang):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * (K @ K)


def pose_f32(T):
    """fp64 pose -> the fp32 pose handed to the TSDF (the reference casts
    with .cast<float>(), integrate.cpp:98)."""
    return np.asarray(T, np.float64).astype(np.float32)


# --------------------------------------------------------------------------
# scenes
# --------------------------------------------------------------------------

PLAIN_SMALL = dict(H=480, W=640, intr=(525.0, 525.0, 319.5, 239.5), voxel_length=0.005,
                   truncation=0.02, n_buckets=1 << 20, frames=30, noise_mm=1.0, max_range_m=4.0,
                   max_weight=0.0)
ROOM_LARGE = dict(H=480, W=848, intr=(430.0, 430.0, 423.5, 239.5), voxel_length=0.004,
                  truncation=0.02, n_buckets=1 << 22, frames=600, noise_mm=1.0, max_range_m=6.0,
                  max_weight=128.0)
FP16_LARGE = dict(H=720, W=1280, intr=(640.0, 640.0, 639.5, 359.5), voxel_length=0.002,
                  truncation=0.008, n_buckets=1 << 24, frames=2000, noise_mm=1.0, max_range_m=5.0,
                  max_weight=0.0)


def scene_plain_small():
    """Plane z=0, sphere r=0.3 m, axis-aligned box of side 0.4 m."""
    return [(PLANE, (0, 0, 1, 0)),
            (SPHERE, (0.35, 0.25, 0.3, 0.3)),
            (BOX, (-0.55, -0.45, 0.0, -0.15, -0.05, 0.4))]


def traj_plain_small(n=30, radius=1.5, height=1.0, target=(0.0, 0.0, 0.2)):
    """Circular trajectory of radius 1.5 m around the origin, looking at it."""
    return [look_at((radius * math.cos(2 * math.pi * k / n), radius * math.sin(2 * math.pi * k / n),
                     height), target) for k in range(n)]


def scene_room(seed=1, size=(5.0, 4.0, 2.6)):
    """Seeded random room: floor/ceiling/walls + 6 boxes + 2 spheres."""
    rng = np.random.default_rng(seed)
    sx, sy, sz = size
    prims = [(ROOM, (-sx / 2, -sy / 2, 0.0, sx / 2, sy / 2, sz))]
    for _ in range(6):
        w, d, h = rng.uniform(0.2, 0.8, 3)
        cx, cy = rng.uniform(-sx / 2 + 0.5, sx / 2 - 0.5), rng.uniform(-sy / 2 + 0.5, sy / 2 - 0.5)
        if abs(cx) < 0.8 and abs(cy) < 0.8:  # keep the camera's start free
            cx += math.copysign(1.0, cx if cx != 0 else 1.0)
        prims.append((BOX, (cx - w / 2, cy - d / 2, 0.0, cx + w / 2, cy + d / 2, h)))
    for _ in range(2):
        r = rng.uniform(0.15, 0.4)
        cx, cy = rng.uniform(-sx / 2 + 0.6, sx / 2 - 0.6), rng.uniform(-sy / 2 + 0.6, sy / 2 - 0.6)
        prims.append((SPHERE, (cx, cy, rng.uniform(r, 1.8), r)))
    return prims


def traj_random_walk(n=600, seed=2, step_m=0.005, step_deg=0.5, start=(0.0, 0.0, 1.4),
                     bounds=((-1.6, 1.6), (-1.2, 1.2), (1.0, 1.8))):
    """Random walk: each step moves 5 mm along a heading that turns by at most
    step_deg (reflected at `bounds`, so the camera stays inside the room), and
    rotates the camera by step_deg about a random axis biased toward the
    world vertical; the first pose looks along +x, slightly down."""
    rng = np.random.default_rng(seed)
    T = look_at(start, (start[0] + 1.0, start[1], start[2] - 0.3))
    heading = np.array([1.0, 0.3, 0.0])
    heading /= np.linalg.norm(heading)
    poses = []
    for _ in range(n):
        poses.append(T.copy())
        turn = axis_angle((0, 0, 1), math.radians(rng.uniform(-step_deg, step_deg)))
        heading = turn @ heading
        nxt = T[:3, 3] + step_m * heading
        for a in range(3):
            if not bounds[a][0] <= nxt[a] <= bounds[a][1]:
                heading[a] = -heading[a]
        T[:3, 3] += step_m * heading
        ax = rng.normal(size=3) * np.array([0.3, 0.3, 1.0])
        T[:3, :3] = axis_angle(ax, math.radians(step_deg)) @ T[:3, :3]
    return poses


# multi-room floor plan: 3 x 2 rooms of 4 m x 4 m in a 12 m x 8 m x 2.6 m
# shell, 0.1 m interior walls with 1 m doors
MULTIROOM = dict(size=(12.0, 8.0, 2.6), rooms=(3, 2), room=4.0, wall=0.1, door=1.0)


def scene_multiroom(seed=5):
    """Seeded multi-room scene: outer shell, interior walls with doors, and
    in every room 3 boxes and 1 sphere placed in the room's corners (the
    trajectory runs along the rooms' centre lines)."""
    rng = np.random.default_rng(seed)
    sx, sy, sz = MULTIROOM["size"]
    t, door, rs = MULTIROOM["wall"], MULTIROOM["door"], MULTIROOM["room"]
    prims = [(ROOM, (-sx / 2, -sy / 2, 0.0, sx / 2, sy / 2, sz))]
    # walls x = -2, 2 (along y), one door per room they separate
    for wx in (-2.0, 2.0):
        y0 = -sy / 2
        for dy in (-2.0, 2.0):
            prims.append((BOX, (wx - t / 2, y0, 0.0, wx + t / 2, dy - door / 2, sz)))
            y0 = dy + door / 2
        prims.append((BOX, (wx - t / 2, y0, 0.0, wx + t / 2, sy / 2, sz)))
    # wall y = 0 (along x), one door per room
    x0 = -sx / 2
    for dx in (-4.0, 0.0, 4.0):
        prims.append((BOX, (x0, -t / 2, 0.0, dx - door / 2, t / 2, sz)))
        x0 = dx + door / 2
    prims.append((BOX, (x0, -t / 2, 0.0, sx / 2, t / 2, sz)))
    for cx in (-4.0, 0.0, 4.0):
        for cy in (-2.0, 2.0):
            corners = [(1, 1), (1, -1), (-1, 1), (-1, -1)]
            rng.shuffle(corners)
            for k, (qx, qy) in enumerate(corners):
                px = cx + qx * rng.uniform(1.1, 1.4)
                py = cy + qy * rng.uniform(1.1, 1.4)
                if k < 3:
                    w, d, h = rng.uniform(0.25, 0.5, 2).tolist() + [rng.uniform(0.3, 1.2)]
                    prims.append((BOX, (px - w / 2, py - d / 2, 0.0, px + w / 2, py + d / 2, h)))
                else:
                    r = rng.uniform(0.15, 0.3)
                    prims.append((SPHERE, (px, py, rng.uniform(r, 1.5), r)))
    del rs
    return prims


def traj_multiroom(n=2000, seed=6, height=1.4):
    """Loop through all six rooms via their doors (room centres and door
    centres as waypoints, about 24 m), looking along the direction of travel
    with a slow yaw sweep of +-35 deg and a pitch of -12 deg +- 5 deg."""
    rng = np.random.default_rng(seed)
    way = [(-4, -2), (-2, -2), (0, -2), (2, -2), (4, -2), (4, 0), (4, 2), (2, 2), (0, 2), (-2, 2),
           (-4, 2), (-4, 0), (-4, -2)]
    way = np.array(way, np.float64)
    seg = np.linalg.norm(np.diff(way, axis=0), axis=1)
    cum = np.concatenate([[0.0], np.cumsum(seg)])
    L = cum[-1]
    ph = rng.uniform(0, 2 * math.pi, 2)
    poses = []
    for k in range(n):
        s = L * k / n
        i = min(int(np.searchsorted(cum, s, side="right")) - 1, len(seg) - 1)
        u = (s - cum[i]) / seg[i]
        pos = way[i] + u * (way[i + 1] - way[i])
        # heading: tangent blended toward the next segment near its end
        tan = (way[i + 1] - way[i]) / seg[i]
        if i + 1 < len(seg) and u > 0.7:
            nxt = (way[i + 2] - way[i + 1]) / seg[i + 1]
            a = (u - 0.7) / 0.3
            tan = (1 - a) * tan + a * nxt
            tan /= np.linalg.norm(tan)
        yaw = math.atan2(tan[1], tan[0]) + math.radians(35.0) * math.sin(2 * math.pi * k / 400.0 + ph[0])
        pitch = math.radians(-12.0 + 5.0 * math.sin(2 * math.pi * k / 300.0 + ph[1]))
        eye = np.array([pos[0], pos[1], height])
        fwd = np.array([math.cos(yaw) * math.cos(pitch), math.sin(yaw) * math.cos(pitch), math.sin(pitch)])
        poses.append(look_at(eye, eye + fwd))
    return poses


def perturb(T, rng, sigma_t=0.01, sigma_deg=0.5):
    """Pose-graph style update: N(0, 1 cm) translation, N(0, 0.5 deg) rotation."""
    P = np.array(T, np.float64)
    ax = rng.normal(size=3)
    P[:3, :3] = axis_angle(ax, math.radians(rng.normal(0.0, sigma_deg))) @ P[:3, :3]
    P[:3, 3] += rng.normal(0.0, sigma_t, 3)
    return P


def frames(prims, poses, cfg, seed=0, keep=None, H=None, W=None):
    """Depth stack [n, H, W] uint16 for the given poses."""
    H = H or cfg["H"]
    W = W or cfg["W"]
    out = np.zeros((len(poses), H, W), np.uint16)
    for k, T in enumerate(poses):
        out[k] = render_depth(prims, T, H, W, cfg["intr"], cfg["noise_mm"], cfg["max_range_m"],
                              seed=seed * 1000003 + k, keep=1.0 if keep is None else keep[k])
    return out


def scaled_intr(cfg, H, W):
    """Intrinsics rescaled to an H x W image of the same field of view."""
    fx, fy, cx, cy = cfg["intr"]
    sx, sy = W / cfg["W"], H / cfg["H"]
    return (fx * sx, fy * sy, (cx + 0.5) * sx - 0.5, (cy + 0.5) * sy - 0.5)
