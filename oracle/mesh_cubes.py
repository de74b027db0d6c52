"""numpy restatement of the reference's mesh extraction -- TEST
INFRASTRUCTURE ONLY, imported by tests/ as the checker.

marchingCubesHost (GPUVoxelGrid.cpp:200-251) copies the bricks into an
Open3D 0.12 ScalableTSDFVolume (volume_unit_resolution 8, RGB8 colour; the
copy loop at :226-241 writes storage voxel (z, y, x) at Open3D voxel
(z, y, x), its two transpositions cancelling) and calls
ScalableTSDFVolume::ExtractTriangleMesh (:248) and ComputeVertexNormals
(:249).  Open3D is a third-party dependency absent from the reference tree
and from this image, so its published algorithm is restated here:

  * corners k = 0..7 of cell (x, y, z) at offsets (0,0,0) (1,0,0) (1,1,0)
    (0,1,0) (0,0,1) (1,0,1) (1,1,1) (0,1,1); a corner in the next unit
    along an axis reads that unit, a missing unit reads w = 0;
  * any corner with w == 0 drops the cell; bit k of the case is sdf_k < 0;
    cases 0 and 255 emit nothing;
  * edge e from corner LO[e] to HI[e] along AXIS[e]; its vertex at
    (vl/2 + vl G) + e_a f0 vl / (f0 + f1) in double, G the global index of
    the low voxel, f0 = |sdf_lo|, f1 = |sdf_hi|; colour
    (f1 c0 + f0 c1) / (f0 + f1) with c = rgb / 255;
  * triangle t of row TRI[case] emitted as (e[3t], e[3t+2], e[3t+1]);
  * vertex normals: sum over the vertex's triangles of (v1 - v0) x (v2 - v0)
    in triangle order, then divided by the norm (a zero sum stays zero).

TRI is the published Lorensen-Cline case table in Bourke's numbering
(Open3D's tri_table).  No copy of it exists in this image to check against;
tests/test_mc_oracle.py pins it by its structure (exact edge sets, manifold
cells, one rule on every ambiguous face, complementary rows) and by closed
forms.
"""
from __future__ import annotations

import numpy as np

CORNER = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], np.int64)
LO = np.array([0, 1, 3, 0, 4, 5, 7, 4, 0, 1, 2, 3])
HI = np.array([1, 2, 2, 3, 5, 6, 6, 7, 4, 5, 6, 7])
AXIS = np.array([0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2])

_TRI_TEXT = """
-1
0 8 3
0 1 9
1 8 3 9 8 1
1 2 10
0 8 3 1 2 10
9 2 10 0 2 9
2 8 3 2 10 8 10 9 8
3 11 2
0 11 2 8 11 0
1 9 0 2 3 11
1 11 2 1 9 11 9 8 11
3 10 1 11 10 3
0 10 1 0 8 10 8 11 10
3 9 0 3 11 9 11 10 9
9 8 10 10 8 11
4 7 8
4 3 0 7 3 4
0 1 9 8 4 7
4 1 9 4 7 1 7 3 1
1 2 10 8 4 7
3 4 7 3 0 4 1 2 10
9 2 10 9 0 2 8 4 7
2 10 9 2 9 7 2 7 3 7 9 4
8 4 7 3 11 2
11 4 7 11 2 4 2 0 4
9 0 1 8 4 7 2 3 11
4 7 11 9 4 11 9 11 2 9 2 1
3 10 1 3 11 10 7 8 4
1 11 10 1 4 11 1 0 4 7 11 4
4 7 8 9 0 11 9 11 10 11 0 3
4 7 11 4 11 9 9 11 10
9 5 4
9 5 4 0 8 3
0 5 4 1 5 0
8 5 4 8 3 5 3 1 5
1 2 10 9 5 4
3 0 8 1 2 10 4 9 5
5 2 10 5 4 2 4 0 2
2 10 5 3 2 5 3 5 4 3 4 8
9 5 4 2 3 11
0 11 2 0 8 11 4 9 5
0 5 4 0 1 5 2 3 11
2 1 5 2 5 8 2 8 11 4 8 5
10 3 11 10 1 3 9 5 4
4 9 5 0 8 1 8 10 1 8 11 10
5 4 0 5 0 11 5 11 10 11 0 3
5 4 8 5 8 10 10 8 11
9 7 8 5 7 9
9 3 0 9 5 3 5 7 3
0 7 8 0 1 7 1 5 7
1 5 3 3 5 7
9 7 8 9 5 7 10 1 2
10 1 2 9 5 0 5 3 0 5 7 3
8 0 2 8 2 5 8 5 7 10 5 2
2 10 5 2 5 3 3 5 7
7 9 5 7 8 9 3 11 2
9 5 7 9 7 2 9 2 0 2 7 11
2 3 11 0 1 8 1 7 8 1 5 7
11 2 1 11 1 7 7 1 5
9 5 8 8 5 7 10 1 3 10 3 11
5 7 0 5 0 9 7 11 0 1 0 10 11 10 0
11 10 0 11 0 3 10 5 0 8 0 7 5 7 0
11 10 5 7 11 5
10 6 5
0 8 3 5 10 6
9 0 1 5 10 6
1 8 3 1 9 8 5 10 6
1 6 5 2 6 1
1 6 5 1 2 6 3 0 8
9 6 5 9 0 6 0 2 6
5 9 8 5 8 2 5 2 6 3 2 8
2 3 11 10 6 5
11 0 8 11 2 0 10 6 5
0 1 9 2 3 11 5 10 6
5 10 6 1 9 2 9 11 2 9 8 11
6 3 11 6 5 3 5 1 3
0 8 11 0 11 5 0 5 1 5 11 6
3 11 6 0 3 6 0 6 5 0 5 9
6 5 9 6 9 11 11 9 8
5 10 6 4 7 8
4 3 0 4 7 3 6 5 10
1 9 0 5 10 6 8 4 7
10 6 5 1 9 7 1 7 3 7 9 4
6 1 2 6 5 1 4 7 8
1 2 5 5 2 6 3 0 4 3 4 7
8 4 7 9 0 5 0 6 5 0 2 6
7 3 9 7 9 4 3 2 9 5 9 6 2 6 9
3 11 2 7 8 4 10 6 5
5 10 6 4 7 2 4 2 0 2 7 11
0 1 9 4 7 8 2 3 11 5 10 6
9 2 1 9 11 2 9 4 11 7 11 4 5 10 6
8 4 7 3 11 5 3 5 1 5 11 6
5 1 11 5 11 6 1 0 11 7 11 4 0 4 11
0 5 9 0 6 5 0 3 6 11 6 3 8 4 7
6 5 9 6 9 11 4 7 9 7 11 9
10 4 9 6 4 10
4 10 6 4 9 10 0 8 3
10 0 1 10 6 0 6 4 0
8 3 1 8 1 6 8 6 4 6 1 10
1 4 9 1 2 4 2 6 4
3 0 8 1 2 9 2 4 9 2 6 4
0 2 4 4 2 6
8 3 2 8 2 4 4 2 6
10 4 9 10 6 4 11 2 3
0 8 2 2 8 11 4 9 10 4 10 6
3 11 2 0 1 6 0 6 4 6 1 10
6 4 1 6 1 10 4 8 1 2 1 11 8 11 1
9 6 4 9 3 6 9 1 3 11 6 3
8 11 1 8 1 0 11 6 1 9 1 4 6 4 1
3 11 6 3 6 0 0 6 4
6 4 8 11 6 8
7 10 6 7 8 10 8 9 10
0 7 3 0 10 7 0 9 10 6 7 10
10 6 7 1 10 7 1 7 8 1 8 0
10 6 7 10 7 1 1 7 3
1 2 6 1 6 8 1 8 9 8 6 7
2 6 9 2 9 1 6 7 9 0 9 3 7 3 9
7 8 0 7 0 6 6 0 2
7 3 2 6 7 2
2 3 11 10 6 8 10 8 9 8 6 7
2 0 7 2 7 11 0 9 7 6 7 10 9 10 7
1 8 0 1 7 8 1 10 7 6 7 10 2 3 11
11 2 1 11 1 7 10 6 1 6 7 1
8 9 6 8 6 7 9 1 6 11 6 3 1 3 6
0 9 1 11 6 7
7 8 0 7 0 6 3 11 0 11 6 0
7 11 6
7 6 11
3 0 8 11 7 6
0 1 9 11 7 6
8 1 9 8 3 1 11 7 6
10 1 2 6 11 7
1 2 10 3 0 8 6 11 7
2 9 0 2 10 9 6 11 7
6 11 7 2 10 3 10 8 3 10 9 8
7 2 3 6 2 7
7 0 8 7 6 0 6 2 0
2 7 6 2 3 7 0 1 9
1 6 2 1 8 6 1 9 8 8 7 6
10 7 6 10 1 7 1 3 7
10 7 6 1 7 10 1 8 7 1 0 8
0 3 7 0 7 10 0 10 9 6 10 7
7 6 10 7 10 8 8 10 9
6 8 4 11 8 6
3 6 11 3 0 6 0 4 6
8 6 11 8 4 6 9 0 1
9 4 6 9 6 3 9 3 1 11 3 6
6 8 4 6 11 8 2 10 1
1 2 10 3 0 11 0 6 11 0 4 6
4 11 8 4 6 11 0 2 9 2 10 9
10 9 3 10 3 2 9 4 3 11 3 6 4 6 3
8 2 3 8 4 2 4 6 2
0 4 2 4 6 2
1 9 0 2 3 4 2 4 6 4 3 8
1 9 4 1 4 2 2 4 6
8 1 3 8 6 1 8 4 6 6 10 1
10 1 0 10 0 6 6 0 4
4 6 3 4 3 8 6 10 3 0 3 9 10 9 3
10 9 4 6 10 4
4 9 5 7 6 11
0 8 3 4 9 5 11 7 6
5 0 1 5 4 0 7 6 11
11 7 6 8 3 4 3 5 4 3 1 5
9 5 4 10 1 2 7 6 11
6 11 7 1 2 10 0 8 3 4 9 5
7 6 11 5 4 10 4 2 10 4 0 2
3 4 8 3 5 4 3 2 5 10 5 2 11 7 6
7 2 3 7 6 2 5 4 9
9 5 4 0 8 6 0 6 2 6 8 7
3 6 2 3 7 6 1 5 0 5 4 0
6 2 8 6 8 7 2 1 8 4 8 5 1 5 8
9 5 4 10 1 6 1 7 6 1 3 7
1 6 10 1 7 6 1 0 7 8 7 0 9 5 4
4 0 10 4 10 5 0 3 10 6 10 7 3 7 10
7 6 10 7 10 8 5 4 10 4 8 10
6 9 5 6 11 9 11 8 9
3 6 11 0 6 3 0 5 6 0 9 5
0 11 8 0 5 11 0 1 5 5 6 11
6 11 3 6 3 5 5 3 1
1 2 10 9 5 11 9 11 8 11 5 6
0 11 3 0 6 11 0 9 6 5 6 9 1 2 10
11 8 5 11 5 6 8 0 5 10 5 2 0 2 5
6 11 3 6 3 5 2 10 3 10 5 3
5 8 9 5 2 8 5 6 2 3 8 2
9 5 6 9 6 0 0 6 2
1 5 8 1 8 0 5 6 8 3 8 2 6 2 8
1 5 6 2 1 6
1 3 6 1 6 10 3 8 6 5 6 9 8 9 6
10 1 0 10 0 6 9 5 0 5 6 0
0 3 8 5 6 10
10 5 6
11 5 10 7 5 11
11 5 10 11 7 5 8 3 0
5 11 7 5 10 11 1 9 0
10 7 5 10 11 7 9 8 1 8 3 1
11 1 2 11 7 1 7 5 1
0 8 3 1 2 7 1 7 5 7 2 11
9 7 5 9 2 7 9 0 2 2 11 7
7 5 2 7 2 11 5 9 2 3 2 8 9 8 2
2 5 10 2 3 5 3 7 5
8 2 0 8 5 2 8 7 5 10 2 5
9 0 1 5 10 3 5 3 7 3 10 2
9 8 2 9 2 1 8 7 2 10 2 5 7 5 2
1 3 5 3 7 5
0 8 7 0 7 1 1 7 5
9 0 3 9 3 5 5 3 7
9 8 7 5 9 7
5 8 4 5 10 8 10 11 8
5 0 4 5 11 0 5 10 11 11 3 0
0 1 9 8 4 10 8 10 11 10 4 5
10 11 4 10 4 5 11 3 4 9 4 1 3 1 4
2 5 1 2 8 5 2 11 8 4 5 8
0 4 11 0 11 3 4 5 11 2 11 1 5 1 11
0 2 5 0 5 9 2 11 5 4 5 8 11 8 5
9 4 5 2 11 3
2 5 10 3 5 2 3 4 5 3 8 4
5 10 2 5 2 4 4 2 0
3 10 2 3 5 10 3 8 5 4 5 8 0 1 9
5 10 2 5 2 4 1 9 2 9 4 2
8 4 5 8 5 3 3 5 1
0 4 5 1 0 5
8 4 5 8 5 3 9 0 5 0 3 5
9 4 5
4 11 7 4 9 11 9 10 11
0 8 3 4 9 7 9 11 7 9 10 11
1 10 11 1 11 4 1 4 0 7 4 11
3 1 4 3 4 8 1 10 4 7 4 11 10 11 4
4 11 7 9 11 4 9 2 11 9 1 2
9 7 4 9 11 7 9 1 11 2 11 1 0 8 3
11 7 4 11 4 2 2 4 0
11 7 4 11 4 2 8 3 4 3 2 4
2 9 10 2 7 9 2 3 7 7 4 9
9 10 7 9 7 4 10 2 7 8 7 0 2 0 7
3 7 10 3 10 2 7 4 10 1 10 0 4 0 10
1 10 2 8 7 4
4 9 1 4 1 7 7 1 3
4 9 1 4 1 7 0 8 1 8 7 1
4 0 3 7 4 3
4 8 7
9 10 8 10 11 8
3 0 9 3 9 11 11 9 10
0 1 10 0 10 8 8 10 11
3 1 10 11 3 10
1 2 11 1 11 9 9 11 8
3 0 9 3 9 11 1 2 9 2 11 9
0 2 11 8 0 11
3 2 11
2 3 8 2 8 10 10 8 9
9 10 2 0 9 2
2 3 8 2 8 10 0 1 8 1 10 8
1 10 2
1 3 8 9 1 8
0 9 1
0 3 8
-1
"""
TRI = [[int(v) for v in line.split()] if line.strip() != "-1" else [] for line in _TRI_TEXT.strip("\n").split("\n")]
assert len(TRI) == 256

KEY_BIAS = 1 << 18

# Planted mistakes that the closed-form tests must catch (0 = none):
#   1 emits the table's triangle order (faces into the occupied side),
#   2 interpolates from the high end of the edge,
#   3 takes w <= 0 as unobserved (drops deintegration remainders w < 0),
#   4 leaves the colours in 0..255,
#   5 numbers the corners x + 2y + 4z instead of the table's order,
#   6 drops the last triangle of every row with more than one.
_mutation = 0


def set_mutation(m: int) -> None:
    global _mutation
    _mutation = int(m)


def edge_key(g, axis):
    """Key of lattice edge (global voxel g [.., 3], axis): g biased by 2^18 per
    axis in bits 43-61 / 24-42 / 5-23, the axis in bits 0-4."""
    g = np.asarray(g, np.int64) + KEY_BIAS
    return (g[..., 0] << 43) | (g[..., 1] << 24) | (g[..., 2] << 5) | np.asarray(axis, np.int64)


def padded(coords, sdf, w, rgb=None):
    """Each brick with one voxel plane of its +x / +y / +z neighbours:
    (S, W, C) indexed [brick, z, y, x] over 0..8; a missing neighbour reads
    sdf 0, w 0."""
    coords = np.asarray(coords, np.int64).reshape(-1, 3)
    n = coords.shape[0]
    s4 = np.asarray(sdf, np.float32).reshape(n, 8, 8, 8)
    w4 = np.asarray(w, np.float32).reshape(n, 8, 8, 8)
    c4 = np.asarray(rgb, np.float32).reshape(n, 8, 8, 8, 3) if rgb is not None else None
    index = {tuple(c): k for k, c in enumerate(coords.tolist())}
    S = np.zeros((n, 9, 9, 9), np.float32)
    W = np.zeros((n, 9, 9, 9), np.float32)
    Cp = np.zeros((n, 9, 9, 9, 3), np.float32) if c4 is not None else None
    for q in range(8):
        dx, dy, dz = q & 1, (q >> 1) & 1, (q >> 2) & 1
        ids = np.array([index.get((c[0] + dx, c[1] + dy, c[2] + dz), -1) for c in coords.tolist()], np.int64)
        has = ids >= 0
        dst = tuple(slice(8, 9) if e else slice(0, 8) for e in (dz, dy, dx))
        src = tuple(slice(0, 1) if e else slice(0, 8) for e in (dz, dy, dx))
        S[(has,) + dst] = s4[ids[has]][(slice(None),) + src]
        W[(has,) + dst] = w4[ids[has]][(slice(None),) + src]
        if Cp is not None:
            Cp[(has,) + dst] = c4[ids[has]][(slice(None),) + src]
    return S, W, Cp


def marching_cubes(coords, sdf, w, vl, rgb=None):
    """Mesh of exported bricks (coords [n, 3], sdf / w [n, 512] voxel
    x + 8y + 64z, rgb [n, 512, 3] or None) -> dict: vertex_keys [V] int64
    (sorted), vertices [V, 3] float64, colors [V, 3] float64 or None,
    faces [T, 3] int64 into the vertices, normals [V, 3] float64, and
    instances: (key [m], position [m, 3]) of every (triangle, corner)
    before welding."""
    coords = np.asarray(coords, np.int64).reshape(-1, 3)
    n = coords.shape[0]
    vl64 = np.float64(np.float32(vl))
    half = vl64 * 0.5
    empty = {"vertex_keys": np.zeros(0, np.int64), "vertices": np.zeros((0, 3)), "faces": np.zeros((0, 3), np.int64),
             "normals": np.zeros((0, 3)), "colors": np.zeros((0, 3)) if rgb is not None else None,
             "instances": (np.zeros(0, np.int64), np.zeros((0, 3)))}
    if n == 0:
        return empty
    S, W, Cp = padded(coords, sdf, w, rgb)
    corner = CORNER
    if _mutation == 5:
        corner = np.array([(k & 1, (k >> 1) & 1, (k >> 2) & 1) for k in range(8)], np.int64)
    f = np.zeros((n * 512, 8), np.float32)
    obs = np.ones(n * 512, bool)
    col = np.zeros((n * 512, 8, 3), np.float64) if Cp is not None else None
    for k in range(8):
        dx, dy, dz = corner[k]
        f[:, k] = S[:, dz:dz + 8, dy:dy + 8, dx:dx + 8].reshape(-1)
        wk = W[:, dz:dz + 8, dy:dy + 8, dx:dx + 8].reshape(-1)
        obs &= (wk > 0) if _mutation == 3 else (wk != 0)
        if col is not None:
            ck = Cp[:, dz:dz + 8, dy:dy + 8, dx:dx + 8].reshape(-1, 3).astype(np.float64)
            col[:, k] = ck if _mutation == 4 else ck / 255.0
    case = ((f < 0).astype(np.int64) << np.arange(8)).sum(1)
    case[~obs] = 0
    cells = np.flatnonzero((case != 0) & (case != 255))
    if cells.size == 0:
        return empty
    b, cell = cells >> 9, cells & 511
    g0 = coords[b] * 8 + np.stack([cell & 7, (cell >> 3) & 7, cell >> 6], 1)
    fc, cc, cs = f[cells].astype(np.float64), case[cells], col[cells] if col is not None else None
    keys, pos, cols = [], [], []
    for c in np.unique(cc):
        row = TRI[int(c)]
        nt = len(row) // 3
        if _mutation == 6 and nt > 1:
            nt -= 1
        sel = cc == c
        for t in range(nt):
            order = (row[3 * t], row[3 * t + 1], row[3 * t + 2]) if _mutation == 1 else \
                    (row[3 * t], row[3 * t + 2], row[3 * t + 1])
            tk, tp, tc = [], [], []
            for e in order:
                lo, hi, a = LO[e], HI[e], AXIS[e]
                g = g0[sel] + corner[lo]
                f0, f1 = np.abs(fc[sel, lo]), np.abs(fc[sel, hi])
                p = half + vl64 * g.astype(np.float64)
                frac = (f1 if _mutation == 2 else f0) * vl64 / (f0 + f1)
                p[:, a] = p[:, a] + frac
                tk.append(edge_key(g, a))
                tp.append(p)
                if cs is not None:
                    c0, c1 = cs[sel, lo], cs[sel, hi]
                    tc.append((f1[:, None] * c0 + f0[:, None] * c1) / (f0 + f1)[:, None])
            keys.append(np.stack(tk, 1))
            pos.append(np.stack(tp, 1))
            if cs is not None:
                cols.append(np.stack(tc, 1))
    tkeys = np.concatenate(keys, 0)
    tpos = np.concatenate(pos, 0)
    uk, first, inv = np.unique(tkeys.reshape(-1), return_index=True, return_inverse=True)
    verts = tpos.reshape(-1, 3)[first]
    faces = inv.reshape(-1, 3).astype(np.int64)
    vcol = np.concatenate(cols, 0).reshape(-1, 3)[first] if cs is not None else None
    return {"vertex_keys": uk, "vertices": verts, "faces": faces, "normals": vertex_normals(verts, faces),
            "colors": vcol if rgb is not None else None, "instances": (tkeys.reshape(-1), tpos.reshape(-1, 3))}


def vertex_normals(verts, faces):
    """ComputeVertexNormals: unnormalised face normals summed per vertex in
    face order, then divided by their norm (a zero sum stays zero)."""
    v = np.asarray(verts, np.float64)
    f = np.asarray(faces, np.int64)
    a = v[f[:, 1]] - v[f[:, 0]]
    b = v[f[:, 2]] - v[f[:, 0]]
    fn = np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1], a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                   a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], 1)
    vn = np.zeros_like(v)
    for q in range(3):
        np.add.at(vn, f[:, q], fn)
    z = vn[:, 0] * vn[:, 0] + vn[:, 1] * vn[:, 1] + vn[:, 2] * vn[:, 2]
    r = np.sqrt(z)
    nz = z > 0
    vn[nz] = vn[nz] / r[nz, None]
    return vn


def canonical_faces(keys):
    """Triangles as vertex-key triples [T, 3], each rotated to start at its
    smallest key (orientation kept), rows sorted: equal arrays are equal
    triangle multisets."""
    k = np.asarray(keys, np.int64).reshape(-1, 3)
    r = np.argmin(k, 1)
    idx = (r[:, None] + np.arange(3)[None, :]) % 3
    k = np.take_along_axis(k, idx, 1)
    o = np.lexsort((k[:, 2], k[:, 1], k[:, 0]))
    return k[o]
