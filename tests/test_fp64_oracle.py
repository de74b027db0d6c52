"""Pins the fp64 integrate oracle (oracle/fp64_oracle.c) used to measure the
fp32 / fp16 GPU grids:

  * closed form on a fronto-parallel plane: sdf = D - z_voxel exactly,
    weight = number of frames;
  * against the binary32 oracle (itself pinned by KATs and closed forms) on
    a scene with curved surfaces: the same voxels observed with the same
    weight agree to fp32 rounding (the few voxels whose projection lands on
    a neighbouring pixel in one of the two precisions excepted);
  * integrate followed by deintegrate of the same frame returns the weights.
"""
import numpy as np

import oracle
from superscanner_amd import synth


def _vox_z(coords, vbl, vl):
    kz = np.arange(512) >> 6
    return coords[:, 2:3] * vbl + (kz[None, :] + 0.5) * vl


def test_plane_closed_form():
    H, W = 120, 160
    g = oracle.Oracle64Grid(0.01, 0.04, H, W, (130.0, 130.0, 79.5, 59.5), (-0.3, -0.3, 0.8), (0.3, 0.3, 1.2))
    for _ in range(3):
        g.integrate(np.full((H, W), 1000, np.uint16), np.eye(4))
    ex = g.export()
    m = ex["w"] > 0
    assert m.sum() > 1000
    assert np.all(ex["w"][m] == 3.0)
    z = _vox_z(ex["coords"], 0.08, 0.01)
    assert np.abs(ex["sdf"][m] - (1.0 - z[m])).max() < 1e-15
    # only blocks meeting the crop are kept
    assert ex["coords"][:, 2].min() >= int(np.floor(0.8 / 0.08))
    assert ex["coords"][:, 2].max() <= int(np.floor(1.2 / 0.08))


def _scene(n=6, H=120, W=160):
    cfg = dict(synth.PLAIN_SMALL)
    intr = synth.scaled_intr(cfg, H, W)
    poses = synth.traj_plain_small(n=24)[:n]
    depth = synth.frames(synth.scene_plain_small(), poses, dict(cfg, intr=intr), H=H, W=W, seed=3)
    return intr, poses, depth


def _index(coords):
    return {tuple(c): k for k, c in enumerate(coords.tolist())}


def test_agrees_with_binary32_oracle_to_rounding():
    H, W = 120, 160
    intr, poses, depth = _scene()
    lo, hi = (-0.6, -0.6, -0.1), (0.7, 0.6, 0.7)
    g64 = oracle.Oracle64Grid(0.01, 0.04, H, W, intr, lo, hi)
    g32 = oracle.OracleGrid(0.01, 0.04, H, W)
    g32.set_intrinsics(*intr)
    for d, T in zip(depth, poses):
        Tf = synth.pose_f32(T)
        g32.integrate(d, Tf)
        g64.integrate(d, Tf.astype(np.float64))
    e64, e32 = g64.export(), g32.export()
    idx = _index(e32["coords"])
    rows = [(k, idx[tuple(c)]) for k, c in enumerate(e64["coords"].tolist()) if tuple(c) in idx]
    assert len(rows) > 0.95 * len(e64["coords"])
    a, b = np.array([r[0] for r in rows]), np.array([r[1] for r in rows])
    w64, w32 = e64["w"][a], e32["w"][b]
    s64, s32 = e64["sdf"][a], e32["sdf"][b]
    same = (w64 > 0) & (w64 == w32)
    assert same.sum() > 0.98 * (w64 > 0).sum()
    err = np.abs(s64[same] - s32[same])
    assert np.percentile(err, 99) < 1e-6     # fp32 rounding of mm-scale values
    assert np.percentile(err, 99.99) < 1e-4  # a neighbouring pixel in one precision


def test_integrate_deintegrate_restores_weights():
    H, W = 120, 160
    intr, poses, depth = _scene(n=3)
    g = oracle.Oracle64Grid(0.01, 0.04, H, W, intr, (-0.6, -0.6, -0.1), (0.7, 0.6, 0.7))
    for d, T in zip(depth[:2], poses[:2]):
        g.integrate(d, T)
    w0 = g.export()["w"].copy()
    g.integrate(depth[2], poses[2])
    g.integrate(depth[2], poses[2], -1)
    w1 = g.export()["w"][: len(w0)]
    assert np.array_equal(w0, w1)
