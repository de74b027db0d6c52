"""The raycast march the gfx950 kernel runs (tsdf_march.hpp), compiled for
the host, reproduces the oracle's raycast bit for bit -- points, depth, UB
flags and the number of voxel reads and hash probes -- on the plain_small and
room_large scenes (tests/cpp/march_host.cpp)."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from superscanner_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def march_lib(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    so = tmp_path_factory.mktemp("march") / "libmarch_host.so"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                    "-fhip-fp32-correctly-rounded-divide-sqrt", "-o", str(so),
                    os.path.join(ROOT, "tests", "cpp", "march_host.cpp")], check=True, capture_output=True)
    L = C.CDLL(str(so))
    vp, f, i = C.c_void_p, C.c_float, C.c_int
    L.march_host.argtypes = [i, i, f, f, f, f, f, f, f, i, vp, vp, vp, vp, C.c_uint32, i, vp, vp, vp, vp,
                             C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_int)]
    L.march_host.restype = C.c_int
    L.march_tile_reuse.argtypes = [i, i, f, f, f, f, f, f, f, i, vp, vp, vp, vp, C.c_uint32, i, vp, vp]
    L.march_tile_reuse.restype = C.c_int
    return L


def run_host(L, o, cfg, intr, E, H, W, use_sb):
    ex = o.export()
    n = len(ex["keys"])
    nb = 1 << max(10, (2 * n).bit_length())
    pts = np.zeros((H, W, 3), np.float32)
    dep = np.zeros((H, W), np.float32)
    ub = np.zeros((H, W), np.uint8)
    ns, npr, cert = C.c_uint64(), C.c_uint64(), C.c_int()
    E = np.ascontiguousarray(E, np.float32).reshape(16)
    f = C.c_float
    st = L.march_host(H, W, f(intr[0]), f(intr[1]), f(intr[2]), f(intr[3]), f(cfg["voxel_length"]),
                      f(cfg["truncation"]), f(10.0), n, ex["keys"].ctypes.data, ex["coords"].ctypes.data,
                      ex["sdf"].ctypes.data, ex["w"].ctypes.data, C.c_uint32(nb), int(use_sb), E.ctypes.data,
                      pts.ctypes.data, dep.ctypes.data, ub.ctypes.data, C.byref(ns), C.byref(npr), C.byref(cert))
    assert st == 0
    assert cert.value == int(use_sb > 0)  # these scenes are certified alias-free, no brick near the camera
    return pts, dep, ub, ns.value, npr.value


def check(L, o, cfg, intr, E, H, W):
    r = o.raycast(E)
    for use_sb in (0, 1, 2):  # no skip, superblock distance field, tile depth bounds
        compare(run_host(L, o, cfg, intr, E, H, W, use_sb), r)
    return r


def compare(res, r):
    pts, dep, ub, ns, npr = res
    assert (ub == r["ub"]).all(), int((ub != r["ub"]).sum())
    ok = r["ub"] == 0
    assert (pts[ok].view(np.uint32) == r["points"][ok].view(np.uint32)).all()
    assert (dep[ok].view(np.uint32) == r["depth"][ok].view(np.uint32)).all()
    assert ns == r["counts"]["samples"] and npr == r["counts"]["probes"], (ns, npr, r["counts"])


def test_march_plain_small(march_lib):
    H, W = 240, 320
    cfg = synth.PLAIN_SMALL
    intr = synth.scaled_intr(cfg, H, W)
    poses = synth.traj_plain_small(n=12)
    depth = synth.frames(synth.scene_plain_small(), poses, dict(cfg, intr=intr), seed=3, H=H, W=W)
    o = oracle.OracleGrid(cfg["voxel_length"], cfg["truncation"], H, W)
    o.set_intrinsics(*intr)
    for k, T in enumerate(poses):
        o.integrate(depth[k], synth.pose_f32(T))
    for k in (0, 5):
        r = check(march_lib, o, cfg, intr, synth.pose_f32(poses[k]), H, W)
        assert r["counts"]["n_hit"] > 0.5 * H * W


def test_march_room(march_lib):
    H, W = 240, 424
    cfg = synth.ROOM_LARGE
    intr = synth.scaled_intr(cfg, H, W)
    poses = synth.traj_random_walk(n=20, seed=2)
    depth = synth.frames(synth.scene_room(seed=1), poses, dict(cfg, intr=intr), seed=4, H=H, W=W)
    o = oracle.OracleGrid(cfg["voxel_length"], cfg["truncation"], H, W, max_weight=128)
    o.set_intrinsics(*intr)
    for k in range(20):
        o.integrate(depth[k], synth.pose_f32(poses[k]))
        if k == 12:
            o.integrate(depth[2], synth.pose_f32(poses[2]), -1)
    check(march_lib, o, cfg, intr, synth.pose_f32(poses[19]), H, W)


def test_tile_reuse_of_lookups_and_voxels(march_lib):
    """What a per-workgroup LDS cache could serve on the room workload
    (848 x 480, a map of 40 frames, tile depth bounds on): the raycast replayed
    tile by tile as the kernel groups its rays.  The figures are printed; the
    assertions hold the shape of the result: a tile's loaded lookups repeat
    (an ideal key cache would serve > 90 % of them), its voxel reads mostly do
    not (< 40 %), and with every lane of a wavefront waiting for the slowest,
    fewer than half of the wavefront lookup steps would be served whole."""
    H, W = 480, 848
    cfg = synth.ROOM_LARGE
    intr = cfg["intr"]
    poses = synth.traj_random_walk(n=40, seed=2)
    depth = synth.frames(synth.scene_room(seed=1), poses, dict(cfg), seed=4)
    oracle.set_threads(8)
    try:
        o = oracle.OracleGrid(cfg["voxel_length"], cfg["truncation"], H, W, max_weight=128, max_bricks=1 << 20)
        o.set_intrinsics(*intr)
        for k in range(40):
            o.integrate(depth[k], synth.pose_f32(poses[k]))
    finally:
        oracle.set_threads(1)
    ex = o.export()
    n = len(ex["keys"])
    nb = 1 << max(10, (2 * n).bit_length())
    E = np.ascontiguousarray(synth.pose_f32(poses[39]), np.float32).reshape(16)
    st = np.zeros(10, np.uint64)
    f = C.c_float
    assert march_lib.march_tile_reuse(H, W, f(intr[0]), f(intr[1]), f(intr[2]), f(intr[3]), f(cfg["voxel_length"]),
                                      f(cfg["truncation"]), f(10.0), n, ex["keys"].ctypes.data,
                                      ex["coords"].ctypes.data, ex["sdf"].ctypes.data, ex["w"].ctypes.data,
                                      C.c_uint32(nb), 2, E.ctypes.data, st.ctypes.data) == 0
    # the traced march is the march: same sample and probe counts
    _, _, _, ns, npr = run_host(march_lib, o, cfg, intr, E, H, W, 2)
    assert (int(st[8]), int(st[9])) == (ns, npr)
    lookups, keys, vreads, voxels, lane_hits, steps, whole, kmax = (int(x) for x in st[:8])
    ideal_key, ideal_vox = 1 - keys / lookups, 1 - voxels / vreads
    print(f"\nloaded lookups {lookups}, distinct (tile, key) {keys}: ideal key-cache hit {ideal_key:.3f}; "
          f"voxel reads {vreads}, distinct (tile, voxel) {voxels}: ideal voxel reuse {ideal_vox:.3f}; "
          f"lockstep lane hits {lane_hits / lookups:.3f}, wavefront steps served whole {whole / steps:.3f}; "
          f"most keys in one tile {kmax}")
    assert ideal_key > 0.9 and ideal_vox < 0.4
    assert whole / steps < 0.5 and kmax < 1024
