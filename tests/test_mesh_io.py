"""Mesh and brick formats on the host side:

  * the Open3D voxel layout of the brick export: marchingCubesHost
    (GPUVoxelGrid.cpp:226-241) walks (x, y, z) with z fastest, writes Open3D
    index x + 8y + 64z from storage voxel number vcount = 64x + 8y + z, and
    Open3D's UniformTSDFVolume::IndexOf(X, Y, Z) is 64X + 8Y + Z; replaying
    that loop shows the volume holds storage voxel (x, y, z) at
    IndexOf(x, y, z), which storage_to_open3d reproduces;
  * the PLY writer scales colours by color_scale (1 for the marching-cubes
    mesh, whose colours are Open3D's 0..1, 255 for the tetrahedra) and reads
    back what it wrote.
"""
import numpy as np
import torch

from superscanner_amd import mesh_io
from superscanner_amd.voxel_grid import storage_to_open3d


def storage_value(b, x, y, z):
    return 1000 * b + x + 10 * y + 100 * z


def test_reference_copy_loop_keeps_each_voxel_at_its_own_coordinates():
    storage = np.array([storage_value(0, v & 7, (v >> 3) & 7, v >> 6) for v in range(512)])
    o3d = np.zeros(512, np.int64)
    vcount = 0
    for x in range(8):
        for y in range(8):
            for z in range(8):
                o3d[x + 8 * y + 64 * z] = storage[vcount]
                vcount += 1
    for x in range(8):
        for y in range(8):
            for z in range(8):
                assert o3d[64 * x + 8 * y + z] == storage_value(0, x, y, z)


def test_storage_to_open3d_numpy_and_torch():
    n = 3
    v = np.arange(512)
    x, y, z = v & 7, (v >> 3) & 7, v >> 6
    a = np.stack([storage_value(b, x, y, z) for b in range(n)]).astype(np.float32)
    rgb = np.stack([a, a + 0.25, a + 0.5], -1)
    for arr, col in ((a, rgb), (torch.from_numpy(a), torch.from_numpy(rgb))):
        o = np.asarray(storage_to_open3d(arr))
        oc = np.asarray(storage_to_open3d(col))
        assert o.shape == (n, 512) and oc.shape == (n, 512, 3)
        for b in range(n):
            for (xx, yy, zz) in ((0, 0, 0), (7, 0, 0), (0, 7, 0), (0, 0, 7), (3, 5, 6), (7, 7, 7)):
                k = 64 * xx + 8 * yy + zz
                assert o[b, k] == storage_value(b, xx, yy, zz)
                assert np.array_equal(oc[b, k], [o[b, k], o[b, k] + 0.25, o[b, k] + 0.5])


def test_ply_colour_scale(tmp_path):
    verts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
    faces = np.array([[0, 1, 2]])
    for scale, col in ((1.0, [[0.0, 0.5, 1.0]] * 3), (255.0, [[0.0, 127.5, 255.0]] * 3)):
        m = {"vertices": verts, "faces": faces, "normals": np.tile([0.0, 0.0, 1.0], (3, 1)), "colors": np.array(col),
             "color_scale": scale}
        p = tmp_path / f"m{int(scale)}.ply"
        mesh_io.write_ply(p, m)
        r = mesh_io.read_ply(p)
        assert np.array_equal(r["colors"][0], [0, 128, 255])
        assert np.array_equal(r["faces"], faces)
