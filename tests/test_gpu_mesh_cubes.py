"""Marching cubes on the GPU (tsdf_marching_cubes, the mesh of the reference's
marchingCubesHost, GPUVoxelGrid.cpp:200-251) against the numpy oracle
(oracle/mesh_cubes.py, pinned by test_mc_oracle.py) run on the GPU's own
exported bricks, through the C ABI:

  * the vertex set: the same lattice-edge keys, each at a bit-identical
    double position, with a bit-identical colour on a colour grid;
  * the triangles: the same multiset of vertex-key triples, orientation
    included;
  * the vertex normals: unit vectors within 1e-9 of the oracle's (the GPU
    adds the face normals in another order; vertices whose face normals
    nearly cancel are left out);

for fp32 and fp16 storage, colour, a map with deintegrated voxels (w = 0
drops cells, w = -1 does not), a room-sized map, short buffers (sizes
reported, nothing written), an empty grid and a colour request to a grid
without colour.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from gpu_helpers import dev_u8, dev_u16, make_pair
from oracle import mesh_cubes
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


def assert_f64_bits(got, exp, what):
    got, exp = np.ascontiguousarray(got, np.float64), np.ascontiguousarray(exp, np.float64)
    assert got.shape == exp.shape, f"{what}: shape {got.shape} vs {exp.shape}"
    bad = got.view(np.int64) != exp.view(np.int64)
    assert not bad.any(), f"{what}: {int(bad.sum())} of {bad.size} differ, first {got[bad][:3]} vs {exp[bad][:3]}"


def assert_cubes_equal(g, color=False):
    got = g.extract_mesh(method="cubes")
    ex = {k: v.cpu().numpy() for k, v in g.export_bricks().items()}
    exp = mesh_cubes.marching_cubes(ex["coords"], ex["sdf"], ex["w"], g.vl, ex["rgb"] if color else None)
    o = np.argsort(got["vertex_keys"])
    assert np.array_equal(got["vertex_keys"][o], exp["vertex_keys"]), "vertex sets differ"
    assert_f64_bits(got["vertices"][o], exp["vertices"], "vertex positions")
    if color:
        assert_f64_bits(got["colors"][o], exp["colors"], "vertex colours")
    else:
        assert got["colors"] is None
    kg = mesh_cubes.canonical_faces(got["vertex_keys"][got["faces"]])
    ke = mesh_cubes.canonical_faces(exp["vertex_keys"][exp["faces"]])
    assert kg.shape == ke.shape and np.array_equal(kg, ke), "triangle multisets differ"
    # normals: the same sums in another order; compare where the sum is not a
    # near-cancellation of its terms
    v, f = exp["vertices"], exp["faces"]
    fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    mag = np.zeros(len(v))
    raw = np.zeros_like(v)
    for q in range(3):
        np.add.at(mag, f[:, q], np.linalg.norm(fn, axis=1))
        np.add.at(raw, f[:, q], fn)
    well = np.linalg.norm(raw, axis=1) > 1e-3 * mag
    assert well.mean() > 0.99
    assert np.abs(got["normals"][o][well] - exp["normals"][well]).max() < 1e-9
    zero = np.linalg.norm(exp["normals"], axis=1) == 0
    assert np.array_equal(np.linalg.norm(got["normals"][o], axis=1) == 0, zero)
    return got


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_cubes_equal_oracle(gpu, storage):
    cfg, intr, poses, depth = small_frames(n=4 if storage == "f32" else 2)
    g = GPUVoxelGrid(1 << 26, cfg["truncation"], cfg["voxel_length"], 240, 320, n_buckets=1 << 18,
                     n_bricks=1 << 17, storage=storage)
    g.setIntrinsic(*intr)
    for k in range(len(poses)):
        g.integrate(None, dev_u16(depth[k]), poses[k])
    m = assert_cubes_equal(g)
    assert len(m["faces"]) > 2000
    g.close()


def test_cubes_colours_equal_oracle(gpu):
    cfg, intr, poses, depth = small_frames(n=6)
    g, _ = make_pair(cfg["voxel_length"], cfg["truncation"], 240, 320, intr, with_color=True)
    rng = np.random.default_rng(11)
    for k in range(6):
        rgb = rng.integers(0, 256, size=(240, 320, 3), dtype=np.uint8)
        g.integrate(dev_u8(rgb), dev_u16(depth[k]), poses[k])
    m = assert_cubes_equal(g, color=True)
    assert len(m["faces"]) > 1000 and 0.0 <= m["colors"].min() and m["colors"].max() <= 1.0
    g.close()


def test_cubes_with_deintegrated_voxels(gpu):
    cfg, intr, poses, depth = room_frames(12)
    g, _ = make_pair(cfg["voxel_length"], cfg["truncation"], 240, 424, intr)
    for k in range(12):
        g.integrate(None, dev_u16(depth[k]), poses[k])
    for k in (2, 5, 2):  # the second deintegration of frame 2 leaves w = -1 behind
        g.deIntegrate(None, dev_u16(depth[k]), poses[k])
    w = g.export_bricks()["w"]
    assert bool((w < 0).any()) and bool((w == 0).any())
    assert len(assert_cubes_equal(g)["faces"]) > 1000
    g.close()


def test_cubes_room_map(gpu):
    cfg, intr, poses, depth = room_frames(40, seed=7)
    g, _ = make_pair(cfg["voxel_length"], cfg["truncation"], 240, 424, intr, n_bricks=1 << 18)
    for k in range(40):
        g.integrate(None, dev_u16(depth[k]), poses[k])
    m = assert_cubes_equal(g)
    print(f"room map: {g.size()} bricks, {len(m['vertices'])} vertices, {len(m['faces'])} triangles", flush=True)
    assert len(m["faces"]) > 50000
    g.close()


def call(g, verts, cols, nrm, faces, keys, vcap, tcap):
    nv, nt = C.c_uint64(0), C.c_uint64(0)
    p = (lambda t: C.c_void_p(t.data_ptr()) if t is not None else None)
    rc = g._L.tsdf_marching_cubes(g._h, p(verts), p(cols), p(nrm), p(faces), p(keys), vcap, tcap, C.byref(nv),
                                  C.byref(nt), None)
    return rc, nv.value, nt.value


def test_cubes_short_buffers_and_refusals(gpu):
    cfg, intr, poses, depth = small_frames(n=2, H=120, W=160)
    g, _ = make_pair(cfg["voxel_length"], cfg["truncation"], 120, 160, intr)
    for k in range(2):
        g.integrate(None, dev_u16(depth[k]), poses[k])
    full = g.extract_mesh(method="cubes")
    V, T = len(full["vertices"]), len(full["faces"])
    assert V > 100 and T > 100
    verts = torch.full((V, 3), 7.0, dtype=torch.float64, device=gpu)
    faces = torch.full((T, 3), -7, dtype=torch.int32, device=gpu)
    # sizes only: zero capacities, then one short capacity each
    assert call(g, None, None, None, None, None, 0, 0) == (0, V, T)
    assert call(g, verts, None, None, faces, None, V - 1, T) == (0, V, T)
    assert call(g, verts, None, None, faces, None, V, T - 1) == (0, V, T)
    assert bool((verts == 7.0).all()) and bool((faces == -7).all())  # nothing written
    # exact capacities: the same mesh
    keys = torch.empty((V,), dtype=torch.int64, device=gpu)
    assert call(g, verts, None, None, faces, keys, V, T) == (0, V, T)
    k, f = keys.cpu().numpy(), faces.long().cpu().numpy()
    assert f.min() >= 0 and f.max() < V
    assert np.array_equal(mesh_cubes.canonical_faces(k[f]),
                          mesh_cubes.canonical_faces(full["vertex_keys"][full["faces"]]))
    # colours from a grid without colour, null buffers with a capacity
    assert call(g, verts, verts, None, faces, None, V, T)[0] == _lib.TSDF_BAD_ARG
    assert call(g, None, None, None, faces, None, V, T)[0] == _lib.TSDF_BAD_ARG
    assert call(g, verts, None, None, None, None, V, T)[0] == _lib.TSDF_BAD_ARG
    g.close()


def test_cubes_of_empty_grid(gpu):
    cfg, intr, _, _ = small_frames(n=1)
    g, _ = make_pair(cfg["voxel_length"], cfg["truncation"], 240, 320, intr)
    m = g.marchingCubesHost()
    assert m["faces"].shape == (0, 3) and m["vertices"].shape == (0, 3) and g.mesh is m
    g.close()
