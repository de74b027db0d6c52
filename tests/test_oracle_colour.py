"""Closed-form checks that pin the oracle's colour arithmetic and its weight cap.

The reference keeps a running average of the pixel colour per voxel next to
the SDF (updateSingleVoxelColor / deUpdateSingleVoxelColor,
GPUVoxelGrid.cpp:397-469) and the raycast returns the mean colour of the two
voxels around the zero crossing divided by 255 (:682-684).  On a
fronto-parallel plane seen by an identity camera these reduce to facts that
need no restatement of the update formulas:

- a voxel observed n times in one colour c holds exactly c (the sum c*w + c
  is an exact integer, so (c*w + c)/(w + 1) rounds to c);
- frames of colours a and b give the midpoint (a + b)/2 exactly, three
  colours give (a + b + c)/3 to within one float rounding;
- deintegrating a frame restores the colour before it, and deintegrating the
  last observation empties the voxel (w = 1 -> rgb = 0);
- a voxel takes the colour of the pixel its centre projects to (row, column
  truncated toward zero), computed here in float64;
- along a ray near the optical axis the crossing lies between the voxel
  layers just in front of and just behind the plane, so with a colour that
  depends only on the voxel layer the raycast colour is the mean of the two
  layers' colours / 255.

The max-weight cap (project semantics; the reference never caps,
GPUVoxelGrid.h:14,62) is pinned the same way: at cap m the weight stays m and
the SDF follows the exact running average with divisor m + 1.

Every check is also run against oracle builds with a planted mistake in the
colour or cap arithmetic, which must fail at least one of them.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle
from tests.test_oracle_closed_form import CX, CY, D_MM, FX, FY, H, POSES, TRUNC, VL, W, plane_depth, voxel_centres

E_ID = POSES["identity"]


def make_grid(**kw):
    g = oracle.OracleGrid(VL, TRUNC, H, W, with_color=True, **kw)
    g.set_intrinsics(FX, FY, CX, CY)
    return g


def solid(rgb):
    return np.broadcast_to(np.asarray(rgb, np.uint8), (H, W, 3)).copy()


def pattern_image():
    i, j = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    return np.stack([(3 * j + i) % 256, (7 * i) % 256, (i * j + 11) % 256], -1).astype(np.uint8)


def plane_geometry(coords, depth_m):
    """Per voxel of the given blocks (identity camera, plane at depth_m):
    (updated mask, sdf, row, col, certain) in float64."""
    vm = voxel_centres(coords.astype(np.float64)).reshape(-1, 3)
    row = vm[:, 1] * FY / vm[:, 2] + (H - CY)
    col = vm[:, 0] * FX / vm[:, 2] + CX
    sdf = depth_m - vm[:, 2]
    inside = (row > -1) & (row < H) & (col > -1) & (col < W) & (vm[:, 2] > 0)
    upd = inside & (np.abs(sdf) < TRUNC)
    sure = (np.abs(row - np.round(row)) > 1e-3) & (np.abs(col - np.round(col)) > 1e-3) & \
           (np.abs(np.abs(sdf) - TRUNC) > 1e-5)
    sh = (len(coords), 512)
    return upd.reshape(sh), sdf.reshape(sh), row.reshape(sh), col.reshape(sh), sure.reshape(sh)


def check_constant_colour():
    """n observations of one colour hold that colour exactly, per channel."""
    g = make_grid()
    c = (200, 17, 99)
    for n in range(1, 5):
        g.integrate(plane_depth(), E_ID, rgb=solid(c))
        ex = g.export()
        upd, _, _, _, sure = plane_geometry(ex["coords"], D_MM / 1000.0)
        m = upd & sure
        assert (ex["w"][m] == n).all()
        assert (ex["rgb"][m] == np.asarray(c, np.float32)).all(), f"colour after {n} frames"
        assert (ex["rgb"][~upd & sure] == 0).all()
    g.close()


def check_pixel_colour():
    """Each voxel takes the colour of the pixel its centre projects to."""
    g = make_grid()
    img = pattern_image()
    g.integrate(plane_depth(), E_ID, rgb=img)
    ex = g.export()
    upd, _, row, col, sure = plane_geometry(ex["coords"], D_MM / 1000.0)
    m = upd & sure
    r = np.trunc(row[m]).astype(int)
    c = np.trunc(col[m]).astype(int)
    assert m.sum() > 1000
    assert (ex["rgb"][m] == img[r, c].astype(np.float32)).all(), "voxel colour is not its pixel's"
    g.close()


def check_colour_mean_and_inverse():
    """a, b -> (a+b)/2 exactly; a, b, c -> (a+b+c)/3 within one rounding;
    deintegrating c then b restores (a+b)/2 then a; deintegrating a empties."""
    a, b, c = (10, 201, 77), (31, 6, 250), (128, 99, 1)
    g = make_grid()
    g.integrate(plane_depth(), E_ID, rgb=solid(a))
    g.integrate(plane_depth(), E_ID, rgb=solid(b))
    ex = g.export()
    upd, _, _, _, sure = plane_geometry(ex["coords"], D_MM / 1000.0)
    m = upd & sure
    half = (np.asarray(a, np.float64) + np.asarray(b, np.float64)) / 2
    assert (ex["rgb"][m] == half.astype(np.float32)).all(), "mean of two colours"
    g.integrate(plane_depth(), E_ID, rgb=solid(c))
    ex = g.export()
    for ch in range(3):
        exact = Fraction(a[ch] + b[ch] + c[ch], 3)
        got = ex["rgb"][m][:, ch]
        ulp = np.spacing(np.float32(float(exact)))
        assert (np.abs(got.astype(np.float64) - float(exact)) <= ulp).all(), f"mean of three, channel {ch}"
    # deintegrate c: back to the two-frame mean (exact: 3 * m3 - c recovers 2 * m2
    # to within the rounding of m3, and dividing by 2 lands on m2)
    g.deintegrate(plane_depth(), E_ID, rgb=solid(c))
    ex = g.export()
    assert (np.abs(ex["rgb"][m].astype(np.float64) - half) <= 1e-4).all(), "deintegrate restores the mean"
    g.deintegrate(plane_depth(), E_ID, rgb=solid(b))
    ex = g.export()
    assert (np.abs(ex["rgb"][m].astype(np.float64) - np.asarray(a)) <= 2e-4).all(), "deintegrate restores a"
    assert (ex["w"][m] == 1).all()
    g.deintegrate(plane_depth(), E_ID, rgb=solid(a))
    ex = g.export()
    assert (ex["rgb"][m] == 0).all() and (ex["w"][m] == 0).all(), "last observation removed"
    g.close()


def check_deintegrate_exact():
    """Two frames of one colour, one deintegrated: (c*2 - c)/1 = c exactly."""
    c = (250, 3, 64)
    g = make_grid()
    for _ in range(2):
        g.integrate(plane_depth(), E_ID, rgb=solid(c))
    g.deintegrate(plane_depth(), E_ID, rgb=solid(c))
    ex = g.export()
    upd, _, _, _, sure = plane_geometry(ex["coords"], D_MM / 1000.0)
    m = upd & sure
    assert (ex["w"][m] == 1).all()
    assert (ex["rgb"][m] == np.asarray(c, np.float32)).all(), "deintegrate of one of two equal colours"
    g.close()


def layer_colour(gz):
    """Colour of a voxel by its global z layer only."""
    gz = np.asarray(gz, np.int64)
    return np.stack([np.clip(2 * gz - 150, 0, 255), np.clip(255 - gz, 0, 255), (3 * gz) % 256], -1)


def check_raycast_colour():
    """With voxel colours set per z layer, every central ray's colour is the
    mean of the layer in front of the plane and the layer behind it, / 255."""
    g = make_grid()
    for _ in range(2):
        g.integrate(plane_depth(), E_ID, rgb=solid((1, 2, 3)))
    ex = g.export()
    v = np.arange(512)
    gz = ex["coords"][:, 2:3] * 8 + (v // 64)[None, :]
    g.import_rgb(layer_colour(gz).astype(np.float32))
    r = g.raycast(E_ID)
    D = D_MM / 1000.0
    z_in = int(np.ceil(D / VL - 0.5))  # first layer whose centre is at or behind the plane
    z_out = z_in - 1
    centre = np.zeros((H, W), bool)
    centre[H // 2 - 10:H // 2 + 10, W // 2 - 10:W // 2 + 10] = True
    ok = centre & (r["depth"] > 0) & (r["ub"] == 0)
    assert ok.sum() > 300
    lo, hi = layer_colour(z_out), layer_colour(z_in)
    got = r["colors"][ok].astype(np.float64)
    for ch in range(3):
        exact = float(Fraction(int(lo[ch]) + int(hi[ch]), 2 * 255))
        ulp = np.spacing(np.float32(exact))
        assert (np.abs(got[:, ch] - exact) <= ulp).all(), f"raycast colour channel {ch}"
    g.close()


def check_weight_cap():
    """Cap m = 3: three frames at D1, two at D2.  w stays 3 and the sdf is the
    running average with divisor m + 1 after the cap."""
    m_cap = 3.0
    d1, d2 = D_MM, D_MM + 7  # 7 mm apart: voxels near both planes see both
    g = oracle.OracleGrid(VL, TRUNC, H, W, max_weight=m_cap)
    g.set_intrinsics(FX, FY, CX, CY)
    for _ in range(3):
        g.integrate(np.full((H, W), d1, np.uint16), E_ID)
    for _ in range(2):
        g.integrate(np.full((H, W), d2, np.uint16), E_ID)
    ex = g.export()
    u1, s1, _, _, ok1 = plane_geometry(ex["coords"], d1 / 1000.0)
    u2, s2, _, _, ok2 = plane_geometry(ex["coords"], d2 / 1000.0)
    both = u1 & u2 & ok1 & ok2
    assert both.sum() > 1000
    assert (ex["w"][both] == m_cap).all(), "weight must stay at the cap"
    # exact running average in rationals from the float32 per-frame sdf
    got = ex["sdf"][both].astype(np.float64)
    vz = voxel_centres(ex["coords"].astype(np.float64))[..., 2][both]
    e1 = np.float32(d1 / 1000.0).astype(np.float64) - vz
    e2 = np.float32(d2 / 1000.0).astype(np.float64) - vz
    s = e1  # after three equal frames
    for _ in range(2):
        s = (s * m_cap + e2) / (m_cap + 1.0)
    assert np.abs(got - s).max() < 2e-6, np.abs(got - s).max()
    # a voxel seen only by the first plane keeps w = 3, sdf = its sdf
    only1 = u1 & ~u2 & ok1 & ok2
    assert (ex["w"][only1] == m_cap).all()
    g.close()


CHECKS = [check_constant_colour, check_pixel_colour, check_colour_mean_and_inverse, check_deintegrate_exact,
          check_raycast_colour, check_weight_cap]


@pytest.mark.parametrize("check", CHECKS, ids=lambda f: f.__name__)
def test_colour_and_cap_closed_form(check):
    check()


# 7: colour divisor w + 2; 8: cap test w > m instead of w >= m; 9: raycast
# colour from the inner voxel only; 10: colour read at the transposed pixel;
# 11: deintegrate adds the pixel colour instead of subtracting it
@pytest.mark.parametrize("mutation", [7, 8, 9, 10, 11])
def test_planted_colour_mistakes_are_caught(mutation):
    oracle.set_mutation(mutation)
    failures = 0
    try:
        for chk in CHECKS:
            try:
                chk()
            except AssertionError:
                failures += 1
    finally:
        oracle.set_mutation(0)
    assert failures > 0, f"mutation {mutation} went unnoticed"
