"""Pins the numpy marching-tetrahedra oracle (oracle/mesh.py) on fields whose
meshes are known in closed form:

  * one cell with a single inside corner: 6 triangles on the midpoints of the
    corner's 7 edges, total area 3 * (vl / 2)^2, all facing away from it;
  * a plane z = z0 over a block of bricks: every vertex on the plane, every
    normal +z, area = (meshed span in x) * (meshed span in y), with the span
    ending at the last voxel centre of the last brick;
  * unobserved voxels (w = 0 or the deintegration remainder w = -1) remove
    exactly the cells that contain them;
  * a sphere: a closed, consistently oriented surface (every directed edge
    once, its reverse once), Euler characteristic 2, vertices on the sphere
    to within the chord error, positive enclosed volume near 4/3 pi r^3.

Each planted mistake of oracle.mesh.set_mutation fails one of these checks.
"""
import numpy as np
import pytest

from oracle import mesh

VL = 0.01


def grid_field(fn, lo, hi, vl=VL):
    """Bricks lo..hi (inclusive, per axis) with sdf = fn(voxel centres), w = 1."""
    rng = [np.arange(lo[a], hi[a] + 1) for a in range(3)]
    coords = np.stack(np.meshgrid(*rng, indexing="ij"), -1).reshape(-1, 3)
    v = np.arange(512)
    vx, vy, vz = v & 7, (v >> 3) & 7, v >> 6
    px = mesh._voxel_mid(vx[None, :], coords[:, 0:1], vl).astype(np.float64)
    py = mesh._voxel_mid(vy[None, :], coords[:, 1:2], vl).astype(np.float64)
    pz = mesh._voxel_mid(vz[None, :], coords[:, 2:3], vl).astype(np.float64)
    sdf = fn(px, py, pz).astype(np.float32)
    w = np.ones_like(sdf)
    return coords, sdf, w


def normals(t):
    t = t.astype(np.float64)
    return np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])


def area(t):
    return 0.5 * np.linalg.norm(normals(t), axis=1).sum()


def closed_and_oriented(t):
    """(every directed edge once and its reverse once, Euler characteristic)"""
    verts, inv = np.unique(t.reshape(-1, 3).view(np.uint32), axis=0, return_inverse=True)
    f = inv.reshape(-1, 3)
    d = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    n = len(verts)
    key = d[:, 0].astype(np.int64) * n + d[:, 1]
    rkey = d[:, 1].astype(np.int64) * n + d[:, 0]
    uk, cnt = np.unique(key, return_counts=True)
    ok = bool((cnt == 1).all()) and np.isin(rkey, uk).all()
    n_edges = len(np.unique(np.sort(d, 1), axis=0))
    return ok, len(verts) - n_edges + len(f)


@pytest.fixture(autouse=True)
def _no_mutation():
    mesh.set_mutation(0)
    yield
    mesh.set_mutation(0)


def one_cell():
    """Cell (2, 3, 4) of brick (1, -2, 0): corner 0 inside, the rest outside."""
    coords = np.array([[1, -2, 0]])
    sdf = np.full((1, 512), 1.0, np.float32)
    w = np.zeros((1, 512), np.float32)
    for c in range(8):
        v = (2 + (c & 1)) + 8 * (3 + ((c >> 1) & 1)) + 64 * (4 + ((c >> 2) & 1))
        w[0, v] = 1.0
        if c == 0:
            sdf[0, v] = -1.0
    return coords, sdf, w


def test_single_cell_known_answer():
    coords, sdf, w = one_cell()
    t = mesh.extract_mesh(coords, sdf, w, VL)
    assert t.shape == (6, 3, 3)
    p0 = np.array([mesh._voxel_mid(2, 1, VL), mesh._voxel_mid(3, -2, VL), mesh._voxel_mid(4, 0, VL)], np.float64)
    # the vertices are the midpoints of the 7 edges from corner 0
    mids = {tuple(np.round((p0 + 0.5 * VL * np.array([c & 1, c >> 1 & 1, c >> 2 & 1])) / VL * 2).astype(int))
            for c in range(1, 8)}
    got = {tuple(np.round(v / VL * 2).astype(int)) for v in t.reshape(-1, 3).astype(np.float64)}
    assert got == mids
    assert abs(area(t) - 3 * (VL / 2) ** 2) < 1e-4 * VL ** 2
    # facing away from the inside corner
    n = normals(t)
    centre = t.astype(np.float64).mean(1)
    assert ((n * (centre - p0)).sum(1) > 0).all()


def test_single_cell_inverted():
    coords, sdf, w = one_cell()
    t = mesh.extract_mesh(coords, -sdf, w, VL)  # now only corner 0 is outside
    assert t.shape == (6, 3, 3)
    p0 = np.array([mesh._voxel_mid(2, 1, VL), mesh._voxel_mid(3, -2, VL), mesh._voxel_mid(4, 0, VL)], np.float64)
    n = normals(t)
    centre = t.astype(np.float64).mean(1)
    assert ((n * (centre - p0)).sum(1) < 0).all()  # facing corner 0, the free side


Z0 = 0.0537


def plane(lo=(0, 0, 0), hi=(2, 2, 1)):
    return grid_field(lambda x, y, z: z - Z0, lo, hi)


def check_plane(t, nx, ny):
    z = t[:, :, 2].astype(np.float64)
    assert np.abs(z - Z0).max() < 1e-6, np.abs(z - Z0).max()
    assert (normals(t)[:, 2] > 0).all()
    assert abs(area(t) - nx * ny * VL * VL) < 1e-5 * nx * ny * VL * VL


def test_plane_closed_form():
    t = mesh.extract_mesh(*plane(), VL)
    check_plane(t, 23, 23)  # 3 bricks = 24 voxel centres = 23 cells per axis


def test_plane_single_brick():
    t = mesh.extract_mesh(*plane((0, 0, 0), (0, 0, 0)), VL)
    check_plane(t, 7, 7)  # no neighbour bricks: 7 cells per axis


def test_plane_unobserved_voxels_remove_their_cells():
    coords, sdf, w = plane()
    full = area(mesh.extract_mesh(coords, sdf, w, VL))
    # voxel (x, y, z) = (5, 9, 4) is centred at z = 0.045, just below the plane:
    # it is a corner of 4 of the cells that the plane crosses
    b = int(np.flatnonzero((coords == [0, 1, 0]).all(1))[0])
    for bad in (0.0, -1.0):
        w2 = w.copy()
        w2[b, 5 + 8 * 1 + 64 * 4] = bad
        t = mesh.extract_mesh(coords, sdf, w2, VL)
        assert abs(full - area(t) - 4 * VL * VL) < 1e-4 * VL * VL


SPHERE_C = np.array([0.1013, 0.0987, 0.1021])
SPHERE_R = 0.06


def sphere():
    return grid_field(lambda x, y, z: np.sqrt((x - SPHERE_C[0]) ** 2 + (y - SPHERE_C[1]) ** 2 +
                                              (z - SPHERE_C[2]) ** 2) - SPHERE_R, (0, 0, 0), (2, 2, 2))


def signed_volume(t):
    t = t.astype(np.float64)
    return (t[:, 0] * np.cross(t[:, 1], t[:, 2])).sum() / 6.0


def test_sphere_closed_oriented_and_on_surface():
    t = mesh.extract_mesh(*sphere(), VL)
    assert len(t) > 1000
    ok, chi = closed_and_oriented(t)
    assert ok and chi == 2
    r = np.linalg.norm(t.reshape(-1, 3).astype(np.float64) - SPHERE_C, axis=1)
    # linear interpolation of the distance along an edge of length <= sqrt(3) vl
    # lands inside the sphere by at most (sqrt(3) vl)^2 / (8 r) ... plus rounding
    assert (r - SPHERE_R).max() < 1e-6 and (SPHERE_R - r).max() < 3 * VL * VL / (8 * SPHERE_R) + 1e-6
    v = signed_volume(t)
    exact = 4.0 / 3.0 * np.pi * SPHERE_R ** 3
    assert 0.97 * exact < v < exact


def test_empty_and_unobserved():
    assert mesh.extract_mesh(np.zeros((0, 3)), np.zeros((0, 512)), np.zeros((0, 512)), VL).shape == (0, 3, 3)
    coords, sdf, w = sphere()
    assert mesh.extract_mesh(coords, sdf, np.zeros_like(w), VL).shape == (0, 3, 3)


def test_canonical_is_rotation_and_order_invariant():
    t = mesh.extract_mesh(*sphere(), VL)
    rng = np.random.default_rng(0)
    rot = rng.integers(0, 3, len(t))
    idx = (rot[:, None] + np.arange(3)) % 3
    t2 = np.take_along_axis(t, idx[:, :, None], 1)[rng.permutation(len(t))]
    assert np.array_equal(mesh.canonical(t), mesh.canonical(t2))
    # a flipped triangle is a different triangle
    t3 = t.copy()
    t3[0, [1, 2]] = t3[0, [2, 1]]
    assert not np.array_equal(mesh.canonical(t), mesh.canonical(t3))


def test_colour_of_a_linear_field():
    coords, sdf, w = sphere()
    v = np.arange(512)
    px = mesh._voxel_mid((v & 7)[None, :], coords[:, 0:1], VL).astype(np.float64)
    pz = mesh._voxel_mid((v >> 6)[None, :], coords[:, 2:3], VL).astype(np.float64)
    rgb = np.stack([100 + 500 * px, 50 + 300 * pz, np.full_like(px, 7.0)], -1).astype(np.float32)
    t, c = mesh.extract_mesh(coords, sdf, w, VL, rgb)
    assert np.array_equal(t, mesh.extract_mesh(coords, sdf, w, VL))
    # a colour linear in space is linear along every edge: the vertex colour
    # is the field at the vertex
    tv = t.astype(np.float64)
    assert np.abs(c[..., 0] - (100 + 500 * tv[..., 0])).max() < 1e-3
    assert np.abs(c[..., 1] - (50 + 300 * tv[..., 2])).max() < 1e-3
    assert (c[..., 2] == 7.0).all()
    # the colour rides with its vertex through the rotation of canonical()
    a = mesh.canonical(t, c)
    rot = np.roll(np.arange(3), 1)
    assert np.array_equal(a, mesh.canonical(t[:, rot], c[:, rot]))
    assert not np.array_equal(a, mesh.canonical(t[:, rot], c))


def test_ply_round_trip(tmp_path):
    from superscanner_amd import mesh_io
    t = mesh.extract_mesh(*sphere(), VL)
    verts, inv = np.unique(t.reshape(-1, 3).view(np.uint32), axis=0, return_inverse=True)
    m = {"vertices": verts.view(np.float32).reshape(-1, 3), "faces": inv.reshape(-1, 3),
         "normals": np.tile(np.float32([0, 0, 1]), (len(verts), 1)),
         "colors": np.tile(np.float32([255.4, 0.6, 300.0]), (len(verts), 1))}
    mesh_io.write_ply(tmp_path / "m.ply", m)
    r = mesh_io.read_ply(tmp_path / "m.ply")
    assert np.array_equal(r["vertices"], m["vertices"]) and np.array_equal(r["faces"], m["faces"])
    assert np.array_equal(r["normals"], m["normals"])
    assert (r["colors"] == [255, 1, 255]).all()
    assert np.array_equal(r["vertices"][r["faces"]], t)  # the indexed mesh is the soup
    # saveCPUMesh: out-of-range channels zeroed, y and z flipped
    mesh_io.save_cpu_mesh(tmp_path / "s.ply", m)
    r = mesh_io.read_ply(tmp_path / "s.ply")
    assert (r["colors"] == [0, 1, 0]).all()  # 255.4 and 300 are out of range
    assert np.array_equal(r["vertices"], m["vertices"] * np.float32([1, -1, -1]))
    assert np.array_equal(r["normals"], np.tile(np.float32([0, 0, -1]), (len(verts), 1)))


def decode_key(k):
    k = np.asarray(k, np.int64)
    m19 = (1 << 19) - 1
    g = np.stack([(k >> 43) & m19, (k >> 24) & m19, (k >> 5) & m19], -1) - (1 << 18)
    c = k & 31
    d = np.stack([c % 3, (c // 3) % 3, c // 9], -1) - 1
    return g, d


def test_vertex_keys_name_their_edges():
    coords, sdf, w = sphere()
    t, keys = mesh.extract_mesh(coords, sdf, w, VL, keys=True)
    assert np.array_equal(t, mesh.extract_mesh(coords, sdf, w, VL))
    flat_k = keys.reshape(-1)
    flat_p = t.reshape(-1, 3).view(np.uint32)
    uk, first, inv = np.unique(flat_k, return_index=True, return_inverse=True)
    # a key names one position, and distinct keys distinct positions here
    assert np.array_equal(flat_p, flat_p[first][inv])
    assert len(np.unique(flat_p, axis=0)) == len(uk)
    # the vertex lies on the segment between the two voxel centres the key names
    g, d = decode_key(flat_k)
    pa = (g + 0.5) * VL
    pb = (g + d + 0.5) * VL
    v = t.reshape(-1, 3).astype(np.float64)
    u = ((v - pa) * (pb - pa)).sum(1) / ((pb - pa) ** 2).sum(1)
    assert (u >= -1e-6).all() and (u <= 1 + 1e-6).all()
    assert np.abs(pa + u[:, None] * (pb - pa) - v).max() < 1e-6
    assert (np.abs(d).sum(1) > 0).all()
    # welded by key, the sphere is closed and consistently oriented
    f = inv.reshape(-1, 3)
    de = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    n = len(uk)
    kk = de[:, 0].astype(np.int64) * n + de[:, 1]
    ukk, cnt = np.unique(kk, return_counts=True)
    assert (cnt == 1).all() and np.isin(de[:, 1].astype(np.int64) * n + de[:, 0], ukk).all()


@pytest.mark.parametrize("m", [1, 2, 3])
def test_planted_mistakes_are_caught(m):
    mesh.set_mutation(m)
    caught = []
    try:
        test_plane_closed_form()
    except AssertionError:
        caught.append("plane")
    try:
        test_sphere_closed_oriented_and_on_surface()
    except AssertionError:
        caught.append("sphere")
    try:
        test_single_cell_known_answer()
    except AssertionError:
        caught.append("cell")
    assert caught, f"mutation {m} passed every closed-form check"
