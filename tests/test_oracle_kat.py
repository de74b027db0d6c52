"""Known-answer tests pinning the oracle's scalar kernels.

The reference ships no test of this path (exe/Test.cpp only checks tracker
poses and Kmat4f), so the answers below come from closed forms: wrapping
int32 arithmetic for the hash (GPUVoxelGrid.h:124-128), the truncating
block index with its negative-branch quirk (GPUVoxelGrid.h:162-190), and
rigid-motion algebra for the pose inverse.
"""
import math
import struct

import numpy as np
import pytest

import oracle


def f32(x):
    return struct.unpack("f", struct.pack("f", x))[0]


def ref_hash(x, y, z):
    """Closed form: (73856093x + 19349669y + 83492791z) mod 2^32 read as
    int32, widened to uint64 by sign extension."""
    h = (73856093 * x + 19349669 * y + 83492791 * z) % (1 << 32)
    if h >= 1 << 31:
        h -= 1 << 32
    return h % (1 << 64)


@pytest.mark.parametrize("xyz,expect", [
    ((1, 0, 0), 73856093),
    ((0, 1, 0), 19349669),
    ((0, 0, 1), 83492791),
    ((-1, 0, 0), 0xFFFFFFFFFB990BA3),
    ((10, 10, 10), 1766985530),
    ((20, 20, 20), 0xFFFFFFFFD2A41E74),
    ((-5, 3, -7), 0xFFFFFFFFCA9CFE1D),
    ((152, 951, 211), 0),
    ((0, 0, 0), 0),
])
def test_prime_hash_kat(xyz, expect):
    assert oracle.prime_hash(*xyz) == expect
    assert ref_hash(*xyz) == expect


def test_prime_hash_random_matches_closed_form():
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(-100000, 100000, size=(2000, 3)):
        assert oracle.prime_hash(int(x), int(y), int(z)) == ref_hash(int(x), int(y), int(z))


def test_alias_lattice():
    # (152, 951, 211) is a period of the hash: blocks that far apart collide
    for x, y, z in [(3, -4, 5), (-20, 7, 100), (0, 0, 1)]:
        assert oracle.prime_hash(x, y, z) == oracle.prime_hash(x + 152, y + 951, z + 211)


VL = f32(0.005)
VBL = f32(8 * VL)


@pytest.mark.parametrize("a,expect", [
    (0.0, 0), (-0.0, 0), (f32(0.04), 1), (-1e-9, -1),
    (f32(-0.04), -2),   # reference quirk: (int)(-1 - 1) = -2 where floor gives -1
    (f32(-0.08), -3),   # floor gives -2
    (1.0, 25), (f32(0.039), 0), (f32(-0.039), -1),
])
def test_vb_index_kat(a, expect):
    assert VBL == f32(0.04)
    assert oracle.vb_index(a, VBL) == expect


def test_vb_index_equals_floor_off_the_lattice():
    rng = np.random.default_rng(1)
    a = rng.uniform(-5, 5, 20000).astype(np.float32)
    q = a.astype(np.float64) / np.float64(VBL)
    keep = np.abs(q - np.round(q)) > 1e-4  # away from exact multiples
    for v in a[keep][:5000]:
        assert oracle.vb_index(float(v), VBL) == math.floor(float(v) / float(VBL))


def test_f2i_x86_semantics():
    assert oracle.f2i(2.9) == 2
    assert oracle.f2i(-2.9) == -2
    assert oracle.f2i(float("nan")) == -(1 << 31)
    assert oracle.f2i(3e9) == -(1 << 31)
    assert oracle.f2i(-3e9) == -(1 << 31)
    assert oracle.f2i(-2147483648.0) == -(1 << 31)


def test_vin_vb_interior_and_boundary():
    rng = np.random.default_rng(2)
    for _ in range(2000):
        vb = rng.integers(-50, 50, 3)
        frac = rng.uniform(0.01, 0.99, 3)
        vox = rng.integers(0, 8, 3)
        p = (vb * 8 + vox + frac) * float(VL)
        out = oracle.vin_vb(p.astype(np.float32), vb, VBL, VL)
        assert (out == vox).all(), (p, vb, out, vox)
    # a point on the lower face of a negative block lands in voxel 0
    assert list(oracle.vin_vb([f32(-0.04), 0.0, 0.0], [-1, 0, 0], VBL, VL)) == [0, 0, 0]


def random_rigid(rng, t_scale=3.0):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = rng.uniform(-t_scale, t_scale, 3)
    return T.astype(np.float32)


def test_pose_prepare_closed_form():
    rng = np.random.default_rng(3)
    for _ in range(200):
        E = random_rigid(rng)
        Ei, Ez = oracle.pose_prepare(E)
        R = E[:3, :3].astype(np.float64)
        t = E[:3, 3].astype(np.float64)
        assert (Ei[:3, :3] == E[:3, :3].T).all()            # exact transpose
        assert (Ei[:3, 3] == (-(R.T @ t)).astype(np.float32)).all()  # one rounding
        assert (Ei[3] == np.array([0, 0, 0, 1], np.float32)).all()
        np.testing.assert_allclose(Ei.astype(np.float64) @ E.astype(np.float64), np.eye(4), atol=1e-5)
        c = E[:3, 2].astype(np.float64)
        np.testing.assert_allclose(Ez, c / np.linalg.norm(c), rtol=0, atol=2e-7)


def test_pose_rigidity_checks():
    rng = np.random.default_rng(4)
    E = random_rigid(rng)
    assert oracle.pose_is_rigid(E)
    bad = E.copy(); bad[3, 0] = 0.1
    assert not oracle.pose_is_rigid(bad)
    bad = E.copy(); bad[:3, :3] *= 1.01
    assert not oracle.pose_is_rigid(bad)
    bad = E.copy(); bad[1, 1] = np.nan
    assert not oracle.pose_is_rigid(bad)
