"""cleanMemory and mesh extraction on the GPU, through the C ABI.

  * mesh: the GPU triangle soup equals the numpy marching-tetrahedra oracle
    (oracle/mesh.py, pinned on closed forms by test_mesh_oracle.py) run on
    the GPU's own exported bricks, as an exact multiset of float32 triangles,
    for fp32 and fp16 storage and for maps with deintegrated (w = 0, w = -1)
    voxels; a short output buffer reports the full count and holds a subset.
  * clean: after the same integrate / deintegrate sequence the GPU and the
    oracle (orc_clean, GPUVoxelGrid.cpp:185-196, 785-805) erase the same
    bricks, and the grids stay bit-identical through further integration and
    the raycast (with and without empty-superblock skipping); a cleaned map
    equals an uncleaned one on every brick that holds a non-zero weight.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from gpu_helpers import (assert_bits_equal, assert_grids_equal, assert_raycast_equal, dev_u8, dev_u16, make_pair,
                         sorted_export_gpu)
from oracle import mesh
from superscanner_amd import _lib, synth
from superscanner_amd.voxel_grid import GPUVoxelGrid

pytestmark = pytest.mark.gpu


def small_frames(n=10, H=240, W=320):
    cfg = synth.PLAIN_SMALL
    intr = synth.scaled_intr(cfg, H, W)
    poses = synth.traj_plain_small(n=n)
    depth = synth.frames(synth.scene_plain_small(), poses, dict(cfg, intr=intr), seed=1, H=H, W=W)
    return cfg, intr, [synth.pose_f32(T) for T in poses], depth


def room_frames(n, H=240, W=424, seed=3):
    cfg = synth.ROOM_LARGE
    intr = synth.scaled_intr(cfg, H, W)
    poses = synth.traj_random_walk(n=n, seed=seed)
    depth = synth.frames(synth.scene_room(seed=1), poses, dict(cfg, intr=intr), seed=seed, H=H, W=W)
    return cfg, intr, [synth.pose_f32(T) for T in poses], depth


def oracle_mesh_of(g):
    ex = {k: v.cpu().numpy() for k, v in g.export_bricks().items()}
    return mesh.extract_mesh(ex["coords"], ex["sdf"], ex["w"], g.vl)


def mesh_with_keys(g):
    m = g.extract_mesh(method="tetrahedra")
    n = m["triangles"].shape[0]
    keys = torch.empty((max(n, 1), 3), dtype=torch.int64, device=g.device)
    tris = torch.empty((max(n, 1), 3, 3), dtype=torch.float32, device=g.device)
    cnt = C.c_uint64(0)
    assert g._L.tsdf_extract_mesh(g._h, C.c_void_p(tris.data_ptr()), None, C.c_void_p(keys.data_ptr()), n,
                                  C.byref(cnt), None) == 0 and cnt.value == n
    return m, tris[:n].cpu().numpy(), keys[:n].cpu().numpy()


def key_map(tris, keys):
    return dict(zip(keys.reshape(-1).tolist(), map(tuple, tris.reshape(-1, 3).view(np.uint32).tolist())))


def assert_mesh_equal(g):
    got, tk, kg = mesh_with_keys(g)
    ex = {k: v.cpu().numpy() for k, v in g.export_bricks().items()}
    exp, ko = mesh.extract_mesh(ex["coords"], ex["sdf"], ex["w"], g.vl, keys=True)
    tg = got["triangles"].cpu().numpy()
    assert tg.shape == exp.shape, (tg.shape, exp.shape)
    assert np.array_equal(mesh.canonical(tg), mesh.canonical(exp)), "triangle multisets differ"
    # the same vertex keys name the same positions
    assert key_map(tk, kg) == key_map(exp, ko), "vertex keys differ"
    # the welded mesh indexes the same soup, one vertex per key
    assert_bits_equal(got["vertices"][got["faces"]], tg, "indexed mesh")
    assert got["vertices"].shape[0] == len(np.unique(ko))
    return tg


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_mesh_equals_oracle(gpu, storage):
    cfg, intr, poses, depth = small_frames(n=4 if storage == "f32" else 2)
    g = GPUVoxelGrid(1 << 26, cfg["truncation"], cfg["voxel_length"], 240, 320, n_buckets=1 << 18,
                     n_bricks=1 << 17, storage=storage)
    g.setIntrinsic(*intr)
    for k in range(len(poses)):
        g.integrate(None, dev_u16(depth[k]), poses[k])
    t = assert_mesh_equal(g)
    assert len(t) > 5000
    g.close()


def test_mesh_colours_equal_oracle(gpu):
    cfg, intr, poses, depth = small_frames(n=6)
    g, _ = make_pair(cfg["voxel_length"], cfg["truncation"], 240, 320, intr, with_color=True)
    rng = np.random.default_rng(11)
    for k in range(6):
        rgb = rng.integers(0, 256, size=(240, 320, 3), dtype=np.uint8)
        g.integrate(dev_u8(rgb), dev_u16(depth[k]), poses[k])
    got = g.extract_mesh(method="tetrahedra")
    ex = {k: v.cpu().numpy() for k, v in g.export_bricks().items()}
    t, c = mesh.extract_mesh(ex["coords"], ex["sdf"], ex["w"], g.vl, ex["rgb"])
    tg, cg = got["triangles"].cpu().numpy(), got["triangle_colors"].cpu().numpy()
    assert len(t) > 3000 and tg.shape == t.shape
    assert np.array_equal(mesh.canonical(tg, cg), mesh.canonical(t, c)), "coloured triangle multisets differ"
    assert got["colors"].shape == got["vertices"].shape
    # unit normals, except at vertices whose triangles all have zero area (a
    # vertex on a voxel centre whose sdf is exactly 0)
    nrm = np.linalg.norm(got["normals"], axis=1)
    assert np.abs(nrm[nrm > 0] - 1).max() < 1e-5 and (nrm > 0).mean() > 0.999
    g.close()


def test_mesh_with_deintegrated_voxels(gpu):
    cfg, intr, poses, depth = room_frames(12)
    g, _ = make_pair(cfg["voxel_length"], cfg["truncation"], 240, 424, intr)
    for k in range(12):
        g.integrate(None, dev_u16(depth[k]), poses[k])
    for k in (2, 5, 2):  # the second deintegration of frame 2 leaves w = -1 behind
        g.deIntegrate(None, dev_u16(depth[k]), poses[k])
    w = g.export_bricks()["w"]
    assert bool((w < 0).any()) and bool((w == 0).any())
    assert len(assert_mesh_equal(g)) > 1000
    g.close()


def test_mesh_short_buffer(gpu):
    cfg, intr, poses, depth = small_frames(n=2, H=120, W=160)
    g, _ = make_pair(cfg["voxel_length"], cfg["truncation"], 120, 160, intr)
    for k in range(2):
        g.integrate(None, dev_u16(depth[k]), poses[k])
    full = g.extract_mesh(method="tetrahedra")["triangles"].cpu().numpy()
    cap = 100
    buf = torch.full((cap + 1, 3, 3), 7.0, dtype=torch.float32, device=gpu)
    n = C.c_uint64(0)
    assert g._L.tsdf_extract_mesh(g._h, C.c_void_p(buf.data_ptr()), None, None, cap, C.byref(n), None) == 0
    assert n.value == len(full) > cap
    part = buf.cpu().numpy()
    assert (part[cap] == 7.0).all()  # nothing written past the cap
    fc = {tuple(r) for r in mesh.canonical(full).tolist()}
    assert all(tuple(r) in fc for r in mesh.canonical(part[:cap]).tolist())
    # a zero-capacity call only counts
    assert g._L.tsdf_extract_mesh(g._h, None, None, None, 0, C.byref(n), None) == 0
    assert n.value == len(full)
    # colours from a grid without colour are refused before any launch
    assert g._L.tsdf_extract_mesh(g._h, C.c_void_p(buf.data_ptr()), C.c_void_p(buf.data_ptr()), None, cap, C.byref(n),
                                  None) == _lib.TSDF_BAD_ARG
    g.close()


def test_mesh_of_empty_grid(gpu):
    cfg, intr, _, _ = small_frames(n=1)
    g, _ = make_pair(cfg["voxel_length"], cfg["truncation"], 240, 320, intr)
    m = g.extract_mesh(method="tetrahedra")
    assert m["triangles"].shape[0] == 0 and m["faces"].shape == (0, 3)
    g.close()


def deintegrated_pair(color=False):
    cfg, intr, poses, depth = small_frames()
    g, o = make_pair(cfg["voxel_length"], cfg["truncation"], 240, 320, intr, with_color=color)
    rng = np.random.default_rng(4)
    rgb = rng.integers(0, 256, size=(len(poses), 240, 320, 3), dtype=np.uint8) if color else None
    for k in range(len(poses)):
        g.integrate(dev_u8(rgb[k]) if color else None, dev_u16(depth[k]), poses[k])
        o.integrate(depth[k], poses[k], 1, rgb[k] if color else None)
    for k in range(5):
        g.deIntegrate(dev_u8(rgb[k]) if color else None, dev_u16(depth[k]), poses[k])
        o.integrate(depth[k], poses[k], -1, rgb[k] if color else None)
    return g, o, poses, depth, rgb


@pytest.mark.parametrize("color", [False, True])
def test_clean_equals_oracle(gpu, color):
    g, o, poses, depth, rgb = deintegrated_pair(color)
    before = g.size()
    removed = g.cleanMemory()
    assert removed == o.clean() and removed > 0
    assert g.size() == o.n_bricks() == before - removed
    assert_grids_equal(g, o, color=color)
    # the rebuilt table serves lookups and new allocations
    for k in (0, 2, 4):
        g.integrate(dev_u8(rgb[k]) if color else None, dev_u16(depth[k]), poses[k])
        o.integrate(depth[k], poses[k], 1, rgb[k] if color else None)
    assert g.size() == o.n_bricks()
    assert_grids_equal(g, o, color=color)
    for sb in (0, 1):
        g.set_tuning(raycast_skip=sb)
        assert_raycast_equal(g, o, poses[3], color=color)
    assert g.cleanMemory() == o.clean()
    assert_grids_equal(g, o, color=color)
    g.close()


def test_clean_keeps_every_weighted_brick(gpu):
    cfg, intr, poses, depth = small_frames()
    grids = []
    for _ in range(2):
        g = GPUVoxelGrid(1 << 26, cfg["truncation"], cfg["voxel_length"], 240, 320, n_buckets=1 << 18,
                         n_bricks=1 << 17, storage="f16")
        g.setIntrinsic(*intr)
        for k in range(len(poses)):
            g.integrate(None, dev_u16(depth[k]), poses[k])
        for k in range(0, 10, 2):
            g.deIntegrate(None, dev_u16(depth[k]), poses[k])
        grids.append(g)
    a, b = grids
    assert a.cleanMemory() > 0
    ea, eb = sorted_export_gpu(a), sorted_export_gpu(b)
    keep = (eb["w"] != 0).any(1)
    assert np.array_equal(ea["coords"], eb["coords"][keep])
    assert_bits_equal(ea["w"], eb["w"][keep], "weights")
    assert_bits_equal(ea["sdf"], eb["sdf"][keep], "sdf")
    for k in range(0, 10, 2):
        a.integrate(None, dev_u16(depth[k]), poses[k])
        b.integrate(None, dev_u16(depth[k]), poses[k])
    ea, eb = sorted_export_gpu(a), sorted_export_gpu(b)
    ia = {tuple(c): i for i, c in enumerate(ea["coords"].tolist())}
    rows = np.array([ia.get(tuple(c), -1) for c in eb["coords"].tolist()])
    # every brick of the uncleaned map is in the cleaned one unless it is empty
    assert ((rows >= 0) | ((eb["w"] == 0).all(1))).all()
    assert (rows >= 0).sum() == len(ea["coords"])
    assert_bits_equal(ea["w"][rows[rows >= 0]], eb["w"][rows >= 0], "weights after reintegration")
    assert_bits_equal(ea["sdf"][rows[rows >= 0]], eb["sdf"][rows >= 0], "sdf after reintegration")
    for g in grids:
        g.close()


def test_clean_everything_then_integrate(gpu):
    cfg, intr, poses, depth = small_frames(n=3)
    g, o = make_pair(cfg["voxel_length"], cfg["truncation"], 240, 320, intr)
    for k in range(3):
        g.integrate(None, dev_u16(depth[k]), poses[k])
        o.integrate(depth[k], poses[k], 1)
    for k in range(3):
        g.deIntegrate(None, dev_u16(depth[k]), poses[k])
        o.integrate(depth[k], poses[k], -1)
    n = g.size()
    assert g.cleanMemory() == n == o.clean()
    assert g.size() == 0 and g.cleanMemory() == 0
    assert g.extract_mesh(method="tetrahedra")["triangles"].shape[0] == 0
    g.integrate(None, dev_u16(depth[1]), poses[1])
    o.integrate(depth[1], poses[1], 1)
    assert_grids_equal(g, o)
    assert_raycast_equal(g, o, poses[1])
    g.close()
