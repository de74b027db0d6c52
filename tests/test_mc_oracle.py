"""Pins the numpy marching-cubes oracle (oracle/mesh_cubes.py), the restatement
of Open3D's ScalableTSDFVolume::ExtractTriangleMesh that the reference's
marchingCubesHost (GPUVoxelGrid.cpp:200-251) calls.

Open3D is not in this image, so the case table and the arithmetic are pinned
by what the mathematics fixes:

  * table structure, row by row: the edges a row uses are exactly the edges
    whose ends differ in sign; the triangles of a cell form a surface whose
    open border runs along the cube faces, pairing the sign-change edges of
    every face; on every ambiguous face (two diagonal corners inside) the
    pairing cuts off the inside corners, the one rule that makes neighbouring
    cells agree; a row without an ambiguous face is its complement row
    mirrored; every row has as many triangles as its border polygons
    (k vertices -> k - 2 triangles) that this face rule predicts;
  * the 15 base configurations on one cell with corner values +-1: triangle
    counts, vertices at edge midpoints, area in closed form where the
    polygons are planar, every triangle facing the positive corners;
  * an oblique plane over a block of bricks: vertices on the plane, normals
    equal to the plane's, area equal to the plane's section of the box of
    voxel centres (clipped independently);
  * a sphere and a random field: closed, consistently oriented surfaces
    (every directed edge once, its reverse once); the sphere with Euler
    characteristic 2, its volume near 4/3 pi r^3, normals outward;
  * w == 0 drops exactly the cells that contain the voxel, w = -1 drops none;
  * vertex colours of a linear colour field equal the field / 255.

Each planted mistake of mesh_cubes.set_mutation fails at least one check,
and the HIP source's tables equal the oracle's.
"""
import itertools
import os
import re

import numpy as np
import pytest

from oracle import mesh_cubes as mc

VL = 0.01
VL32 = np.float64(np.float32(VL))  # the grid's voxel length
C = mc.CORNER
EDGES = list(zip(mc.LO.tolist(), mc.HI.tolist()))
FACES = [frozenset(k for k in range(8) if C[k][a] == v) for a in range(3) for v in (0, 1)]
HIP_SRC = os.path.join(os.path.dirname(__file__), "..", "superscanner_amd", "csrc", "tsdf_mc.hpp")


@pytest.fixture(autouse=True)
def _no_mutation():
    mc.set_mutation(0)
    yield
    mc.set_mutation(0)


def inside(case):
    return [(case >> k) & 1 for k in range(8)]


def sign_edges(case):
    ins = inside(case)
    return {e for e, (a, b) in enumerate(EDGES) if ins[a] != ins[b]}


def face_of(e1, e2):
    """The cube face both edges lie on, or None."""
    s = [f for f in FACES if set(EDGES[e1]) <= f and set(EDGES[e2]) <= f]
    return s[0] if s else None


def rule_pairs(case):
    """The face segments the face rule predicts: on each face, its two
    sign-change edges, or, with four (an ambiguous face), the pairs that
    share an inside corner."""
    ins = inside(case)
    pairs = set()
    for f in FACES:
        fe = [e for e in sign_edges(case) if set(EDGES[e]) <= f]
        if len(fe) == 2:
            pairs.add(frozenset(fe))
        elif len(fe) == 4:
            for i in range(4):
                for j in range(i + 1, 4):
                    com = set(EDGES[fe[i]]) & set(EDGES[fe[j]])
                    if len(com) == 1 and ins[com.pop()]:
                        pairs.add(frozenset((fe[i], fe[j])))
    return pairs


def tris_of(case):
    r = mc.TRI[case]
    return [tuple(r[i:i + 3]) for i in range(0, len(r), 3)]


def segments(case):
    seg = {}
    for t in tris_of(case):
        for k in range(3):
            a, b = t[k], t[(k + 1) % 3]
            seg[(a, b)] = seg.get((a, b), 0) + 1
    return seg


# ---------------------------------------------------------------------------
# the table
# ---------------------------------------------------------------------------

def test_rows_use_exactly_the_sign_change_edges():
    for c in range(256):
        r = mc.TRI[c]
        assert len(r) % 3 == 0, c
        assert set(r) == sign_edges(c), c
        for t in tris_of(c):
            assert len(set(t)) == 3, (c, t)


def test_cells_are_manifold_with_borders_on_the_faces():
    for c in range(256):
        seg = segments(c)
        border = set()
        for (a, b), k in seg.items():
            if face_of(a, b) is not None:
                assert k + seg.get((b, a), 0) == 1, (c, a, b)  # a border segment, once
                border.add(frozenset((a, b)))
            else:
                assert k == 1 and seg.get((b, a), 0) == 1, (c, a, b)  # interior: once each way
        # the border is the face rule's pairing (ambiguous faces cut off inside corners)
        assert border == rule_pairs(c), c


def test_triangle_counts_follow_the_border_polygons():
    for c in range(256):
        pairs = rule_pairs(c)
        adj = {e: [] for e in sign_edges(c)}
        for p in pairs:
            a, b = tuple(p)
            adj[a].append(b)
            adj[b].append(a)
        assert all(len(v) == 2 for v in adj.values()), c
        seen, expect = set(), 0
        for e0 in adj:
            if e0 in seen:
                continue
            k, e, prev = 0, e0, None
            while True:
                seen.add(e)
                k += 1
                nxt = adj[e][0] if adj[e][0] != prev else adj[e][1]
                prev, e = e, nxt
                if e == e0:
                    break
            expect += k - 2
        assert len(tris_of(c)) == expect, c


def has_ambiguous_face(c):
    return any(len([e for e in sign_edges(c) if set(EDGES[e]) <= f]) == 4 for f in FACES)


def test_unambiguous_rows_mirror_their_complements():
    def canon(t):
        i = t.index(min(t))
        return t[i:] + t[:i]

    n = 0
    for c in range(256):
        if has_ambiguous_face(c):
            continue
        a = sorted(canon(t) for t in tris_of(c))
        b = sorted(canon(tuple(reversed(t))) for t in tris_of(255 - c))
        assert a == b, c
        n += 1
    assert n == 136


def test_hip_tables_equal_the_oracles():
    src = open(HIP_SRC).read()
    body = src[src.index("kMcTri[256][16] = {"):]
    body = body[:body.index("};")]
    rows = re.findall(r"\{([-\d,\s]+)\}", body)
    assert len(rows) == 256
    for c, r in enumerate(rows):
        v = [int(x) for x in r.split(",")]
        assert len(v) == 16
        assert (v[:v.index(-1)] if -1 in v else v) == mc.TRI[c], c
    ntri = re.search(r"kMcNtri\[256\] = \{([\d,\s]+)\}", src).group(1)
    assert [int(x) for x in ntri.split(",")] == [len(t) // 3 for t in mc.TRI]

    def arr(name):
        return [int(x) for x in re.search(name + r"\[\d+\] = \{([\d,\s]+)\}", src).group(1).split(",")]

    assert arr("kMcCorner") == [int(a) + 2 * int(b) + 4 * int(c) for a, b, c in C]
    assert arr("kMcEdgeLo") == mc.LO.tolist()
    assert arr("kMcEdgeHi") == mc.HI.tolist()
    assert arr("kMcEdgeAxis") == mc.AXIS.tolist()


# ---------------------------------------------------------------------------
# fields
# ---------------------------------------------------------------------------

CELL_BRICK = (1, -2, 0)
CELL = (2, 3, 4)


def one_cell(case):
    """Cell CELL of brick CELL_BRICK: corner k -1 when inside, +1 outside;
    the other voxels unobserved."""
    coords = np.array([CELL_BRICK])
    sdf = np.ones((1, 512), np.float32)
    w = np.zeros((1, 512), np.float32)
    for k in range(8):
        x, y, z = (CELL[a] + int(C[k][a]) for a in range(3))
        v = x + 8 * y + 64 * z
        w[0, v] = 1.0
        sdf[0, v] = -1.0 if (case >> k) & 1 else 1.0
    return coords, sdf, w


def corner_pos(k):
    g = np.array(CELL_BRICK) * 8 + np.array(CELL) + C[k]
    return (g + 0.5) * VL32


def tri_area(v, f):
    a = v[f[:, 1]] - v[f[:, 0]]
    b = v[f[:, 2]] - v[f[:, 0]]
    return 0.5 * np.linalg.norm(np.cross(a, b), axis=1)


# the 15 base configurations (inside corners), triangle counts, and the area
# in units of vl^2 where the polygons are planar
BASE = [
    ((), 0, 0.0),
    ((0,), 1, np.sqrt(3) / 8),
    ((0, 1), 2, np.sqrt(2) / 2),
    ((0, 2), 2, 2 * np.sqrt(3) / 8),
    ((0, 6), 2, 2 * np.sqrt(3) / 8),
    ((1, 2, 3), 3, None),
    ((0, 1, 6), 3, np.sqrt(2) / 2 + np.sqrt(3) / 8),
    ((1, 3, 6), 3, 3 * np.sqrt(3) / 8),
    ((0, 1, 2, 3), 2, 1.0),
    ((0, 1, 3, 4), 4, None),
    ((0, 1, 6, 7), 4, 2 * np.sqrt(2) / 2),
    ((0, 1, 2, 6), 4, None),
    ((1, 2, 3, 4), 4, None),
    ((0, 2, 5, 7), 4, 4 * np.sqrt(3) / 8),
    ((1, 2, 3, 7), 4, None),
]


def check_cell(case, n_expect, area_vl2):
    m = mc.marching_cubes(*one_cell(case), VL)
    v, f = m["vertices"], m["faces"]
    if len(f) != n_expect:
        return False
    if n_expect == 0:
        return True
    # vertices at the midpoints of sign-change edges
    mids = [0.5 * (corner_pos(a) + corner_pos(b)) for e, (a, b) in enumerate(EDGES) if e in sign_edges(case)]
    for p in v:
        if min(np.abs(p - q).max() for q in mids) > 1e-12:
            return False
    if area_vl2 is not None and abs(tri_area(v, f).sum() - area_vl2 * VL32 * VL32) > 1e-12:
        return False
    # every triangle faces the positive corners
    ins = inside(case)
    for t in f:
        p = v[t]
        n = np.cross(p[1] - p[0], p[2] - p[0])
        flux = 0.0
        for q in p:
            e = min(range(12), key=lambda e: np.abs(0.5 * (corner_pos(EDGES[e][0]) + corner_pos(EDGES[e][1])) - q).max())
            a, b = EDGES[e]
            d = corner_pos(b) - corner_pos(a) if ins[a] else corner_pos(a) - corner_pos(b)
            flux += n @ d
        if flux <= 0:
            return False
    return True


@pytest.mark.parametrize("k", range(15))
def test_base_configurations(k):
    corners, n, area = BASE[k]
    case = sum(1 << c for c in corners)
    assert check_cell(case, n, area)
    # the complement: the same surface with the other orientation, or (with an
    # ambiguous face) the connected variant of the face rule
    assert len(mc.marching_cubes(*one_cell(255 - case), VL)["faces"]) == len(tris_of(255 - case))


def block_field(fn, nb=2, w=1.0, rgb_fn=None):
    """nb^3 bricks from (-1, 0, -1) with sdf = fn(x, y, z) at voxel centres."""
    rng = [np.arange(-1, -1 + nb), np.arange(0, nb), np.arange(-1, -1 + nb)]
    coords = np.stack(np.meshgrid(*rng, indexing="ij"), -1).reshape(-1, 3)
    v = np.arange(512)
    loc = np.stack([v & 7, (v >> 3) & 7, v >> 6], 1)
    g = coords[:, None, :] * 8 + loc[None]
    p = (g + 0.5) * VL32
    sdf = fn(p[..., 0], p[..., 1], p[..., 2]).astype(np.float32)
    ww = np.full(sdf.shape, w, np.float32)
    rgb = rgb_fn(p).astype(np.float32) if rgb_fn is not None else None
    return coords, sdf, ww, rgb, p


PLANE_N = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])
PLANE_X0 = np.array([-0.004, 0.081, 0.003])


def plane_fn(x, y, z):
    return PLANE_N[0] * (x - PLANE_X0[0]) + PLANE_N[1] * (y - PLANE_X0[1]) + PLANE_N[2] * (z - PLANE_X0[2])


def section_area(n, x0, lo, hi):
    """Area of the plane n.(x - x0) = 0 inside the box [lo, hi] (convex
    polygon from the plane's crossings of the box's 12 edges)."""
    corners = np.array([[lo[0] if i & 1 == 0 else hi[0], lo[1] if i & 2 == 0 else hi[1],
                         lo[2] if i & 4 == 0 else hi[2]] for i in range(8)])
    d = (corners - x0) @ n
    pts = []
    for i in range(8):
        for j in range(i + 1, 8):
            if bin(i ^ j).count("1") == 1 and d[i] * d[j] < 0:
                t = d[i] / (d[i] - d[j])
                pts.append(corners[i] + t * (corners[j] - corners[i]))
    pts = np.array(pts)
    c = pts.mean(0)
    u = np.cross(n, [1.0, 0, 0])
    u /= np.linalg.norm(u)
    w = np.cross(n, u)
    ang = np.arctan2((pts - c) @ w, (pts - c) @ u)
    pts = pts[np.argsort(ang)]
    s = np.zeros(3)
    for i in range(len(pts)):
        s += np.cross(pts[i] - c, pts[(i + 1) % len(pts)] - c)
    return 0.5 * abs(s @ n)


def check_plane():
    coords, sdf, w, _, p = block_field(plane_fn)
    m = mc.marching_cubes(coords, sdf, w, VL)
    v, f, vn = m["vertices"], m["faces"], m["normals"]
    if len(f) < 100:
        return False
    if np.abs((v - PLANE_X0) @ PLANE_N).max() > 1e-8:
        return False
    if np.abs(vn - PLANE_N).max() > 1e-6:
        return False
    lo, hi = p.reshape(-1, 3).min(0), p.reshape(-1, 3).max(0)
    return abs(tri_area(v, f).sum() / section_area(PLANE_N, PLANE_X0, lo, hi) - 1) < 1e-5


def test_plane():
    assert check_plane()


def closed_and_oriented(f, nv):
    d = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    key = d[:, 0] * nv + d[:, 1]
    rkey = d[:, 1] * nv + d[:, 0]
    uk, cnt = np.unique(key, return_counts=True)
    ok = bool((cnt == 1).all()) and bool(np.isin(rkey, uk).all())
    n_edges = len(np.unique(np.sort(d, 1), axis=0))
    return ok, nv - n_edges + len(f)


SPHERE_C = np.array([0.0431, 0.1197, 0.0412])
SPHERE_R = 0.092


def sphere_field():
    return block_field(lambda x, y, z: np.sqrt((x - SPHERE_C[0]) ** 2 + (y - SPHERE_C[1]) ** 2 +
                                               (z - SPHERE_C[2]) ** 2) - SPHERE_R, nb=3)


def check_sphere():
    coords, sdf, w, _, _ = sphere_field()
    m = mc.marching_cubes(coords, sdf, w, VL)
    v, f, vn = m["vertices"], m["faces"], m["normals"]
    ok, chi = closed_and_oriented(f, len(v))
    if not ok or chi != 2:
        return False
    vol = np.einsum("ij,ij->i", v[f[:, 0]], np.cross(v[f[:, 1]], v[f[:, 2]])).sum() / 6
    if abs(vol / (4 / 3 * np.pi * SPHERE_R ** 3) - 1) > 0.015:
        return False
    if np.abs(np.linalg.norm(v - SPHERE_C, axis=1) - SPHERE_R).max() > 0.05 * VL:
        return False
    return bool((np.einsum("ij,ij->i", vn, v - SPHERE_C) > 0).all())


def test_sphere():
    assert check_sphere()


def test_random_field_is_closed():
    rng = np.random.default_rng(5)
    coords, sdf, w, _, p = block_field(lambda x, y, z: np.zeros_like(x), nb=3)
    g = np.round(p / VL - 0.5).astype(np.int64)
    lo, hi = g.reshape(-1, 3).min(0), g.reshape(-1, 3).max(0)
    border = ((g == lo) | (g == hi)).any(-1)
    sdf = np.where(border, 1.0, rng.uniform(-1, 1, sdf.shape)).astype(np.float32)
    m = mc.marching_cubes(coords, sdf, w, VL)
    ok, _ = closed_and_oriented(m["faces"], len(m["vertices"]))
    assert ok
    # the field reaches nearly every case of the table
    S, W, _ = mc.padded(coords, sdf, w)
    f = np.stack([S[:, dz:dz + 8, dy:dy + 8, dx:dx + 8].reshape(-1) for dx, dy, dz in C], 1)
    assert len(np.unique(((f < 0) << np.arange(8)).sum(1))) >= 250


def test_shared_vertices_are_bit_identical():
    coords, sdf, w, _, _ = sphere_field()
    k, p = mc.marching_cubes(coords, sdf, w, VL)["instances"]
    o = np.argsort(k, kind="stable")
    k, p = k[o], p[o]
    same = k[1:] == k[:-1]
    assert same.sum() > len(k) // 2
    assert np.array_equal(p[1:][same].view(np.int64), p[:-1][same].view(np.int64))


def tri_cells(keys):
    """Low corner of the cell each triangle (a vertex-key triple) lies in:
    the one cell that holds all three of its edges."""
    G = np.stack([((keys >> s) & 0x7FFFF) - mc.KEY_BIAS for s in (43, 24, 5)], -1)  # [T, 3, 3]
    A = keys & 31
    out = np.zeros((len(keys), 3), np.int64)
    found = np.zeros(len(keys), np.int64)
    for o in itertools.product((0, 1), repeat=3):
        c0 = G[:, 0] - np.array(o)
        ok = np.ones(len(keys), bool)
        for k in range(3):
            d = G[:, k] - c0
            ok &= ((d >= 0) & (d <= 1)).all(1) & (np.take_along_axis(d, A[:, k:k + 1], 1)[:, 0] == 0)
        out[ok] = c0[ok]
        found += ok
    return out if (found == 1).all() else None


def keyed(m):
    return m["vertex_keys"][m["faces"]]


def check_observation():
    coords, sdf, w, _, p = block_field(plane_fn)
    full = keyed(mc.marching_cubes(coords, sdf, w, VL))
    # a voxel next to the plane
    d = np.abs(plane_fn(p[..., 0], p[..., 1], p[..., 2]))
    b, v = np.unravel_index(np.argmin(d), d.shape)
    w2 = w.copy()
    w2[b, v] = -1.0  # a deintegration remainder: still observed
    if not np.array_equal(mc.canonical_faces(keyed(mc.marching_cubes(coords, sdf, w2, VL))), mc.canonical_faces(full)):
        return False
    w0 = w.copy()
    w0[b, v] = 0.0
    got = mc.canonical_faces(keyed(mc.marching_cubes(coords, sdf, w0, VL)))
    # expected: every triangle but those of the 8 cells around the voxel
    gv = coords[b] * 8 + np.array([v & 7, (v >> 3) & 7, v >> 6])
    cell = tri_cells(full)
    if cell is None:
        return False
    drop = np.all((cell >= gv - 1) & (cell <= gv), axis=1)
    return 0 < drop.sum() and np.array_equal(got, mc.canonical_faces(full[~drop]))


def test_unobserved_voxels_drop_their_cells():
    assert check_observation()


COL_A = np.array([40.0, 200.0, 120.0])
COL_B = np.array([[900.0, -300.0, 100.0], [200.0, 500.0, -700.0], [-400.0, 100.0, 800.0]])


def check_colours():
    coords, sdf, w, rgb, _ = block_field(plane_fn, rgb_fn=lambda p: COL_A + (p - PLANE_X0) @ COL_B.T)
    m = mc.marching_cubes(coords, sdf, w, VL, rgb)
    v, col = m["vertices"], m["colors"]
    expect = (COL_A + (v - PLANE_X0) @ COL_B.T) / 255.0
    return np.abs(col - expect).max() < 1e-6


def test_colours_are_the_field_over_255():
    assert check_colours()


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 6])
def test_planted_mistakes_are_caught(m):
    mc.set_mutation(m)
    checks = [check_plane, check_sphere, check_observation, check_colours] + \
             [lambda k=k: check_cell(sum(1 << c for c in BASE[k][0]), BASE[k][1], BASE[k][2]) for k in range(15)]
    assert not all(c() for c in checks)
