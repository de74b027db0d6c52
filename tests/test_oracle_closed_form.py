"""Closed-form checks that pin the oracle's integrate / deintegrate / raycast.

A fronto-parallel plane at depth D seen by a camera whose optical axis is
normal to it gives every pixel the same camera depth, so the projective SDF
of the reference (GPUVoxelGrid.cpp:357: (E.p - v).Ez) is D - z_c(v) for each
voxel that projects into the image, the weight counts frames, a deintegrate
undoes an integrate (w back to 0, sdf to 0, :382-391), and the voxel-centre
raycast (:561-715) returns the centre of the first voxel at or behind the
plane.  The allocated block set is checked against a brute-force float64
enumeration of each pixel's +-trunc box.  Every check is also run against
oracle builds with a planted mistake, which must fail at least one of them.
"""
import math

import numpy as np
import pytest

import oracle

H, W = 60, 80
FX = FY = 70.0
CX, CY = 39.5, 29.5
VL = 0.01
TRUNC = 0.04
D_MM = 1013


def rot_y(deg):
    a = math.radians(deg)
    return np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])


def pose(R=np.eye(3), t=(0, 0, 0)):
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = t
    return T.astype(np.float32)


POSES = {"identity": pose(), "rotated": pose(rot_y(25.0) @ np.array(
    [[1, 0, 0], [0, math.cos(0.2), -math.sin(0.2)], [0, math.sin(0.2), math.cos(0.2)]]), (0.31, -0.17, 0.23))}


def make_grid(**kw):
    g = oracle.OracleGrid(VL, TRUNC, H, W, **kw)
    g.set_intrinsics(FX, FY, CX, CY)
    return g


def plane_depth():
    return np.full((H, W), D_MM, np.uint16)


def pixel_points_world(E):
    """float64 world points of all pixels at camera depth D (reference camera:
    ci = H - cy, cj = cx)."""
    D = D_MM / 1000.0
    ci, cj = H - CY, CX
    i, j = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    pc = np.stack([(j - cj) / FX * D, (i - ci) / FY * D, np.full(i.shape, D)], -1).reshape(-1, 3)
    E = E.astype(np.float64)
    return pc @ E[:3, :3].T + E[:3, 3]


def brute_force_blocks(E):
    """(certain, possible) block sets: each pixel's +-trunc box enumerated in
    float64.  A box face within 1e-4 blocks of a block face may round either
    way in float32, so those blocks are only 'possible'."""
    pw = pixel_points_world(E)
    vbl = 8 * VL
    blo, bhi = (pw - TRUNC) / vbl, (pw + TRUNC) / vbl
    eps = 1e-4
    lo_min, lo_max = np.floor(blo - eps).astype(int), np.floor(blo + eps).astype(int)
    hi_min, hi_max = np.floor(bhi - eps).astype(int), np.floor(bhi + eps).astype(int)

    def enum(lo, hi):
        s = set()
        for a, b in zip(lo, hi):
            for x in range(a[0], b[0] + 1):
                for y in range(a[1], b[1] + 1):
                    for z in range(a[2], b[2] + 1):
                        s.add((x, y, z))
        return s
    return enum(lo_max, hi_min), enum(lo_min, hi_max)


def assert_block_set(got, E):
    certain, possible = brute_force_blocks(E)
    assert certain <= got, f"{len(certain - got)} required blocks missing"
    assert got <= possible, f"{len(got - possible)} blocks outside every pixel box"
    assert len(possible) - len(certain) < 0.05 * len(certain)


def voxel_centres(coords):
    """float64 centres [n, 512, 3] of the voxels of the given blocks
    (voxel v = x + 8y + 64z)."""
    v = np.arange(512)
    off = np.stack([v % 8, (v // 8) % 8, v // 64], -1) + 0.5
    return (coords[:, None, :] * 8 + off[None]) * VL


def expected_plane(E, n_frames=1):
    """(mask of voxels that must be updated, expected sdf) for the plane."""
    def fn(coords):
        vm = voxel_centres(coords.astype(np.float64)).reshape(-1, 3)
        E64 = E.astype(np.float64)
        vc = (vm - E64[:3, 3]) @ E64[:3, :3]  # camera coordinates
        row = vc[:, 1] * FY / vc[:, 2] + (H - CY)
        col = vc[:, 0] * FX / vc[:, 2] + CX
        sdf = D_MM / 1000.0 - vc[:, 2]
        # (int) truncates toward zero, so rows/cols in (-1, 0) land on pixel 0
        # and pass the reference's bounds check (GPUVoxelGrid.cpp:348-350)
        inside = (row > -1) & (row < H) & (col > -1) & (col < W) & (vc[:, 2] > 0)
        # decisions away from pixel borders and the truncation edge are certain
        sure = (np.abs(row - np.round(row)) > 1e-3) & (np.abs(col - np.round(col)) > 1e-3) & \
               (np.abs(np.abs(sdf) - TRUNC) > 1e-5)
        upd = inside & (np.abs(sdf) < TRUNC)
        return upd.reshape(len(coords), 512), sdf.reshape(len(coords), 512), sure.reshape(len(coords), 512)
    return fn


def check_plane_integration(name):
    """Raises AssertionError when the oracle disagrees with the closed form."""
    E = POSES[name]
    g = make_grid()
    st, cnt = g.integrate(plane_depth(), E)
    assert st == 0
    ex = g.export()
    got = set(map(tuple, ex["coords"].tolist()))
    assert_block_set(got, E)
    upd, sdf, sure = expected_plane(E)(ex["coords"])
    assert ((ex["w"] == 1) == upd)[sure].all(), "updated voxel set"
    assert (ex["w"][~upd & sure] == 0).all()
    err = np.abs(ex["sdf"] - sdf)[upd & sure]
    assert err.max() < 3e-6, f"sdf error {err.max()}"
    assert cnt["n_touched"] == len(got) and cnt["n_new"] == len(got)
    # three identical frames: w counts frames, the running mean stays put
    for _ in range(2):
        g.integrate(plane_depth(), E)
    ex = g.export()
    assert ((ex["w"] == 3) == upd)[sure].all()
    assert np.abs(ex["sdf"] - sdf)[upd & sure].max() < 5e-6
    # deintegrate all three: back to the empty state (GPUVoxelGrid.cpp:382-391)
    for _ in range(3):
        g.deintegrate(plane_depth(), E)
    ex = g.export()
    assert (ex["w"] == 0).all() and (ex["sdf"] == 0).all(), "integrate o deintegrate"
    g.close()


def check_plane_raycast():
    """Voxel-centre march from the camera: the reference's crossing formula
    returns the centre of the first voxel at or behind the plane, so the
    depth lies in [D, D + vl) near the optical axis."""
    E = POSES["identity"]
    g = make_grid()
    for _ in range(2):
        g.integrate(plane_depth(), E)
    r = g.raycast(E)
    D = D_MM / 1000.0
    dep = r["depth"]
    hit = dep > 0
    centre = np.zeros_like(hit)
    centre[H // 2 - 10:H // 2 + 10, W // 2 - 10:W // 2 + 10] = True
    assert hit[centre].all(), "central rays must hit the plane"
    d = dep[centre & (r["ub"] == 0)]
    assert (d >= D - 1e-5).all() and (d < D + VL + 1e-5).all(), (d.min(), d.max())
    # the point lies on its pixel ray through the camera centre
    p = r["points"][hit & (r["ub"] == 0)].astype(np.float64)
    i, j = np.nonzero(hit & (r["ub"] == 0))
    ray = np.stack([(j - CX) / FX, (i - (H - CY)) / FY, np.ones_like(i, dtype=float)], -1)
    ray /= np.linalg.norm(ray, axis=1, keepdims=True)
    perp = np.linalg.norm(p - (p * ray).sum(1, keepdims=True) * ray, axis=1)
    assert perp.max() < 1e-5
    g.close()


def check_deintegrate_quirk():
    """Deintegrating a frame that was never integrated drives w to -1 (no w==0
    guard, GPUVoxelGrid.cpp:382-391): (s*0 - sdf)/(0 - 1) = sdf."""
    E = POSES["identity"]
    g = make_grid()
    g.deintegrate(plane_depth(), E)
    ex = g.export()
    upd, sdf, sure = expected_plane(E)(ex["coords"])
    assert ((ex["w"] == -1) == upd)[sure].all()
    assert np.abs(ex["sdf"] - sdf)[upd & sure].max() < 3e-6
    g.close()


CHECKS = [
    lambda: check_plane_integration("identity"),
    lambda: check_plane_integration("rotated"),
    check_plane_raycast,
    check_deintegrate_quirk,
]


def kat_checks():
    assert oracle.vb_index(np.float32(-0.04), np.float32(0.04)) == -2
    Ei, _ = oracle.pose_prepare(POSES["rotated"])
    assert np.allclose(Ei.astype(np.float64) @ POSES["rotated"].astype(np.float64), np.eye(4), atol=1e-5)


def test_plane_integrate_identity():
    CHECKS[0]()


def test_plane_integrate_rotated():
    CHECKS[1]()


def test_plane_raycast():
    CHECKS[2]()


def test_deintegrate_never_integrated():
    CHECKS[3]()


def test_sphere_raycast_surface():
    """Sphere seen from four poses at 320x240: the raycast point of every
    pixel whose ray meets the sphere within 30 degrees of its normal lies on
    the sphere to within 2.5 voxels, median under 0.75 voxel.  (The reference
    samples voxel centres off the ray and interpolates between two of them
    with its crossing formula, GPUVoxelGrid.cpp:660-681; the projective SDF
    of those centres is read from neighbouring pixels, so single points
    scatter by up to about two voxels even on frontal surfaces.)"""
    import superscanner_amd.synth as synth
    c, r = np.array([0.0, 0.0, 0.0]), 0.3
    prims = [(synth.SPHERE, (*c, r))]
    h, w, f, cx, cy = 240, 320, 280.0, 159.5, 119.5
    g = oracle.OracleGrid(VL, TRUNC, h, w)
    g.set_intrinsics(f, f, cx, cy)
    poses = [synth.look_at((1.2 * math.cos(a), 1.2 * math.sin(a), 0.3), c) for a in (0, 0.3, 0.6, 0.9)]
    for T in poses:
        dep = synth.render_depth(prims, T, h, w, (f, f, cx, cy), noise_mm=0.0)
        g.integrate(dep, synth.pose_f32(T))
    T = poses[1]
    res = g.raycast(synth.pose_f32(T))
    i, j = np.nonzero((res["depth"] > 0) & (res["ub"] == 0))
    d = np.stack([(j - cx) / f, (i - (h - cy)) / f, np.ones(len(i))], -1) @ T[:3, :3].T
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = T[:3, 3]
    b = d @ (o - c)
    disc = b * b - (np.dot(o - c, o - c) - r * r)
    ok = disc > 0
    t = -b - np.sqrt(np.where(ok, disc, 0))
    nrm = (o + t[:, None] * d - c) / r
    frontal = ok & (-(d * nrm).sum(1) > math.cos(math.radians(30)))
    assert frontal.sum() > 2000
    err = np.abs(np.linalg.norm(res["points"][i, j].astype(np.float64) - c, axis=1) - r)[frontal]
    assert err.max() < 2.5 * VL and np.median(err) < 0.75 * VL, (err.max(), np.median(err))
    g.close()


def test_empty_depth_is_noop():
    g = make_grid()
    st, cnt = g.integrate(np.zeros((H, W), np.uint16), POSES["identity"])
    assert st == 0 and cnt["n_touched"] == 0 and g.n_bricks() == 0
    r = g.raycast(POSES["identity"])
    assert (r["points"] == 0).all() and r["counts"]["n_hit"] == 0
    g.close()


def test_pool_overflow_is_reported():
    g = make_grid(max_bricks=10)
    st, _ = g.integrate(plane_depth(), POSES["identity"])
    assert st == 3 and g.n_bricks() == 10
    g.close()


@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 5, 6])
def test_planted_mistakes_are_caught(mutation):
    """sdf sign, untransposed pose, floor block index, textbook ci, flipped
    interpolation, deintegrate keeping sdf: each must break a check."""
    oracle.set_mutation(mutation)
    failures = 0
    try:
        for chk in CHECKS + [kat_checks]:
            try:
                chk()
            except AssertionError:
                failures += 1
    finally:
        oracle.set_mutation(0)
    assert failures > 0, f"mutation {mutation} went unnoticed"
