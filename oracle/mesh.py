"""numpy restatement of the GPU mesh extraction (marching tetrahedra over the
bricks, superscanner_amd/csrc/tsdf_maintain.hpp) -- TEST INFRASTRUCTURE ONLY,
imported by tests/ as the checker.

No reference counterpart: the reference meshes through Open3D's
ScalableTSDFVolume::ExtractTriangleMesh (GPUVoxelGrid.cpp:200-251).  The
definition restated here, in float32 with the kernel's operation order:
  * cell (x, y, z) of brick b has corners c = (c & 1, c >> 1 & 1, c >> 2 & 1)
    at voxels (x, y, z) + c (in b or a neighbouring brick), positions from
    getVoxelMidPoint (GPUVoxelGrid.h:275-278: float block origin, double
    offset, rounded once);
  * a cell is meshed when all 8 voxels exist with weight > 0;
  * 6 tetrahedra (0,1,3,7) (0,3,2,7) (0,2,6,7) (0,6,4,7) (0,4,5,7) (0,5,1,7);
    corners with sdf < 0 are inside; one corner apart -> 1 triangle across its
    3 edges, two and two -> 2 triangles (e00, e01, e11), (e00, e11, e10);
  * edge point pa + t (pb - pa), t = sa / (sa - sb), a the inside corner;
  * each triangle is flipped when its normal points away from the outside
    corners (d = mean of outside corners - mean of inside corners);
  * vertex colours ca + t (cb - ca) on the same edge and t.
"""
from __future__ import annotations

import numpy as np

TETS = ((0, 1, 3, 7), (0, 3, 2, 7), (0, 2, 6, 7), (0, 6, 4, 7), (0, 4, 5, 7), (0, 5, 1, 7))

# Planted mistakes that the closed-form tests must catch (0 = none):
#   1 interpolates from the outside corner with the inside corner's t,
#   2 faces the triangles into the occupied side,
#   3 drops the last tetrahedron of every cell.
_mutation = 0


def set_mutation(m: int) -> None:
    global _mutation
    _mutation = int(m)


def _voxel_mid(v, vb, vl):
    vbl = np.float32(8.0) * np.float32(vl)
    base = (np.asarray(vb).astype(np.float32) * vbl).astype(np.float64)
    return (base + (np.asarray(v).astype(np.float64) + 0.5) * np.float64(np.float32(vl))).astype(np.float32)


def _edge(p, s, a, b):
    t = s[:, a] / (s[:, a] - s[:, b])
    if _mutation == 1:
        return p[:, b] + t[:, None] * (p[:, a] - p[:, b])
    return p[:, a] + t[:, None] * (p[:, b] - p[:, a])


def _edge_col(c, s, a, b):
    t = s[:, a] / (s[:, a] - s[:, b])
    return c[:, a] + t[:, None] * (c[:, b] - c[:, a])


def vertex_key(g_in, d):
    """Key of the vertex on the edge from the inside voxel g_in [.., 3]
    (global voxel index) to g_in + d: the index biased by 2^18 per axis in
    bits 43-61 / 24-42 / 5-23 and (dx+1) + 3 (dy+1) + 9 (dz+1) in bits 0-4."""
    g = np.asarray(g_in, np.int64) + (1 << 18)
    d = np.asarray(d, np.int64)
    return (g[..., 0] << 43) | (g[..., 1] << 24) | (g[..., 2] << 5) | ((d[..., 0] + 1) + 3 * (d[..., 1] + 1)
                                                                        + 9 * (d[..., 2] + 1))


def extract_mesh(coords, sdf, w, vl, rgb=None, keys=False):
    """Triangle soup [m, 3, 3] float32 of exported bricks (coords [n, 3],
    sdf / w [n, 512], voxel x + 8y + 64z); with rgb [n, 512, 3] also the
    vertex colours [m, 3, 3], interpolated along the same edges; with keys
    the vertex keys [m, 3] (vertex_key).  Several outputs come as a tuple
    (soup, colours if rgb, keys if keys)."""
    coords = np.asarray(coords, np.int64).reshape(-1, 3)
    sdf = np.asarray(sdf, np.float32).reshape(-1, 512)
    w = np.asarray(w, np.float32).reshape(-1, 512)
    n = coords.shape[0]
    if n == 0:
        z = np.zeros((0, 3, 3), np.float32)
        outs = [z] + ([z.copy()] if rgb is not None else []) + ([np.zeros((0, 3), np.int64)] if keys else [])
        return tuple(outs) if len(outs) > 1 else z
    if rgb is not None:
        rgb = np.asarray(rgb, np.float32).reshape(-1, 512, 3)
    index = {tuple(c): k for k, c in enumerate(coords.tolist())}
    nbr = np.full((n, 8), -1, np.int64)
    for k, c in enumerate(coords.tolist()):
        for q in range(8):
            nbr[k, q] = index.get((c[0] + (q & 1), c[1] + ((q >> 1) & 1), c[2] + ((q >> 2) & 1)), -1)
    # each brick padded by one voxel plane from its +x / +y / +z neighbours,
    # indexed [brick, z, y, x] (0..8); a missing neighbour leaves weight 0
    S = np.zeros((n, 9, 9, 9), np.float32)
    Wt = np.zeros((n, 9, 9, 9), np.float32)
    Cp = np.zeros((n, 9, 9, 9, 3), np.float32) if rgb is not None else None
    s4, w4 = sdf.reshape(n, 8, 8, 8), w.reshape(n, 8, 8, 8)
    c4 = rgb.reshape(n, 8, 8, 8, 3) if rgb is not None else None
    for q in range(8):
        d = (q >> 2 & 1, q >> 1 & 1, q & 1)  # z, y, x
        dst = tuple(slice(8, 9) if e else slice(0, 8) for e in d)
        src = tuple(slice(0, 1) if e else slice(0, 8) for e in d)
        has = nbr[:, q] >= 0
        ids = nbr[has, q]
        S[(has,) + dst] = s4[ids][(slice(None),) + src]
        Wt[(has,) + dst] = w4[ids][(slice(None),) + src]
        if Cp is not None:
            Cp[(has,) + dst] = c4[ids][(slice(None),) + src]
    m = n * 512
    s = np.zeros((m, 8), np.float32)
    ok = np.ones(m, bool)
    for c in range(8):
        dz, dy, dx = c >> 2 & 1, c >> 1 & 1, c & 1
        s[:, c] = S[:, dz:dz + 8, dy:dy + 8, dx:dx + 8].reshape(-1)
        ok &= Wt[:, dz:dz + 8, dy:dy + 8, dx:dx + 8].reshape(-1) > 0
    # only observed cells with a sign change yield triangles
    ok &= (s < 0).any(1) & (s >= 0).any(1)
    cells = np.flatnonzero(ok)
    s = s[cells]
    B, cell = cells >> 9, cells & 511
    X, Y, Z = cell & 7, (cell >> 3) & 7, cell >> 6
    i9 = np.arange(9)
    px = _voxel_mid(i9[None, :] & 7, coords[:, 0:1] + (i9[None, :] >> 3), vl)
    py = _voxel_mid(i9[None, :] & 7, coords[:, 1:2] + (i9[None, :] >> 3), vl)
    pz = _voxel_mid(i9[None, :] & 7, coords[:, 2:3] + (i9[None, :] >> 3), vl)
    p = np.zeros((cells.size, 8, 3), np.float32)
    col = np.zeros((cells.size, 8, 3), np.float32) if rgb is not None else None
    for c in range(8):
        dz, dy, dx = c >> 2 & 1, c >> 1 & 1, c & 1
        p[:, c, 0] = px[B, X + dx]
        p[:, c, 1] = py[B, Y + dy]
        p[:, c, 2] = pz[B, Z + dz]
        if col is not None:
            col[:, c] = Cp[B, Z + dz, Y + dy, X + dx]
    out, out_col, out_key = [], [], []
    gbase = coords[B] * 8 + np.stack([X, Y, Z], 1)  # global index of corner 0 of each cell
    corner = np.array([[c & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)], np.int64)
    for tet in (TETS[:5] if _mutation == 3 else TETS):
        inside = s[:, list(tet)] < 0
        for case in range(1, 15):
            bits = [(case >> k) & 1 for k in range(4)]
            sel = np.all(inside == np.array(bits, bool), axis=1)
            if not sel.any():
                continue
            ss, pp = s[sel], p[sel]
            ins = [tet[k] for k in range(4) if bits[k]]
            outs = [tet[k] for k in range(4) if not bits[k]]
            if len(ins) == 1:
                edges = [[(ins[0], outs[0]), (ins[0], outs[1]), (ins[0], outs[2])]]
            elif len(outs) == 1:
                edges = [[(ins[0], outs[0]), (ins[1], outs[0]), (ins[2], outs[0])]]
            else:
                e00, e01, e11, e10 = (ins[0], outs[0]), (ins[0], outs[1]), (ins[1], outs[1]), (ins[1], outs[0])
                edges = [[e00, e01, e11], [e00, e11, e10]]
            tris = [np.stack([_edge(pp, ss, a, b) for a, b in e], 1) for e in edges]
            gb = gbase[sel]
            tkeys = [np.stack([vertex_key(gb + corner[a], corner[b] - corner[a]) for a, b in e], 1)
                     for e in edges] if keys else None
            cc = col[sel] if col is not None else None
            tcols = [np.stack([_edge_col(cc, ss, a, b) for a, b in e], 1) for e in edges] if cc is not None else None
            d = np.zeros((pp.shape[0], 3), np.float32)
            for k in ins:
                d = d - pp[:, k] / np.float32(len(ins))
            for k in outs:
                d = d + pp[:, k] / np.float32(len(outs))
            for j, t in enumerate(tris):
                a_ = t[:, 1] - t[:, 0]
                b_ = t[:, 2] - t[:, 0]
                nx = a_[:, 1] * b_[:, 2] - a_[:, 2] * b_[:, 1]
                ny = a_[:, 2] * b_[:, 0] - a_[:, 0] * b_[:, 2]
                nz = a_[:, 0] * b_[:, 1] - a_[:, 1] * b_[:, 0]
                flip = (nx * d[:, 0] + ny * d[:, 1] + nz * d[:, 2]) < 0
                if _mutation == 2:
                    flip = ~flip
                t = t.copy()
                t[flip, 1], t[flip, 2] = t[flip, 2].copy(), t[flip, 1].copy()
                out.append(t)
                if tcols is not None:
                    tc = tcols[j].copy()
                    tc[flip, 1], tc[flip, 2] = tc[flip, 2].copy(), tc[flip, 1].copy()
                    out_col.append(tc)
                if tkeys is not None:
                    tk = tkeys[j].copy()
                    tk[flip, 1], tk[flip, 2] = tk[flip, 2].copy(), tk[flip, 1].copy()
                    out_key.append(tk)
    tris = np.concatenate(out, 0) if out else np.zeros((0, 3, 3), np.float32)
    outs = [tris]
    if rgb is not None:
        outs.append(np.concatenate(out_col, 0) if out_col else np.zeros((0, 3, 3), np.float32))
    if keys:
        outs.append(np.concatenate(out_key, 0) if out_key else np.zeros((0, 3), np.int64))
    return tuple(outs) if len(outs) > 1 else tris


def _less(a, b):
    """Row-wise lexicographic a < b of integer arrays [m, k]."""
    lt = np.zeros(a.shape[0], bool)
    eq = np.ones(a.shape[0], bool)
    for j in range(a.shape[1]):
        lt |= eq & (a[:, j] < b[:, j])
        eq &= a[:, j] == b[:, j]
    return lt


def canonical(tris, cols=None):
    """Triangles as rows of float32 bit patterns (9 per triangle, 18 with
    vertex colours), each rotated so its smallest vertex comes first (the
    winding is kept), rows in a fixed order: two triangle soups are the same
    multiset exactly when these arrays are equal.  The order is by a 64-bit
    hash of the row; rows are compared in full, so a hash collision can only
    make equal soups compare unequal, never the reverse."""
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 3, 3)
    if cols is not None:
        t = np.concatenate([t, np.ascontiguousarray(cols, np.float32).reshape(-1, 3, 3)], 2)
    k = t.shape[2]
    t = np.ascontiguousarray(t).view(np.uint32).astype(np.int64)
    if t.shape[0] == 0:
        return np.zeros((0, 3 * k), np.int64)
    first = np.zeros(t.shape[0], np.int64)
    for q in (1, 2):
        cur = np.take_along_axis(t, first[:, None, None], 1)[:, 0]
        first = np.where(_less(t[:, q], cur), q, first)
    idx = (first[:, None] + np.arange(3)[None, :]) % 3
    r = np.take_along_axis(t, idx[:, :, None], 1).reshape(-1, 3 * k)
    h = np.zeros(r.shape[0], np.uint64)
    mul = np.uint64(0x9E3779B97F4A7C15)
    for j in range(r.shape[1]):
        h = (h ^ r[:, j].astype(np.uint64)) * mul
        h ^= h >> np.uint64(29)
    order = np.lexsort((r[:, 0], h))
    return r[order]
