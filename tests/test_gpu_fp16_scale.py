"""fp16 storage against the fp64 oracle on the multi-room scene.

The GPU grid integrates in fp32 and stores (sdf, w) in fp16 (or fp32); the
fp64 oracle (oracle/fp64_oracle.c) integrates the same frames in double on
the blocks of a crop.  Over voxels observed equally often by both:

  * fp32 storage: error at fp32 rounding (99.9 % below 1e-6 m);
  * fp16 storage: error at fp16 rounding of mm-scale sdf values (a half ulp
    of 8 mm is 3.8e-6 m; 99.9 % below 2e-5 m);
  * the maximum error of either is set by the rare voxel whose fp32
    projection lands on the neighbouring pixel of the fp64 one (reported,
    bounded by the truncation);
  * storage alone (fp16 against fp32 storage, same fp32 measurements, whole
    map): weights identical (integers below 2048), sdf within 3e-5 m;
  * depth rendered from the fp16 map against the fp64 render of the fp64
    map on the crop: RMSE below 5e-5 m.

The measured figures are written to gpurun_out/fp16_accuracy.json when that
folder exists.
"""
import json
import os

import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import dev_u16
from superscanner_amd import diff_render as dr
from superscanner_amd import synth
from superscanner_amd.voxel_grid import GPUVoxelGrid

pytestmark = pytest.mark.gpu

H, W = 360, 640
CROP = ((-3.0, -3.6, 0.0), (-1.9, -0.4, 1.4))  # east side of the first room: wall, door, two objects


def _index(coords):
    return {tuple(c): k for k, c in enumerate(coords.tolist())}


def test_fp16_storage_against_fp64_oracle(gpu):
    cfg = dict(synth.FP16_LARGE)
    intr = synth.scaled_intr(cfg, H, W)
    all_poses = synth.traj_multiroom(2000)
    poses = all_poses[:240:3]  # the first room and its doors
    depth = synth.frames(synth.scene_multiroom(), poses, dict(cfg, intr=intr), H=H, W=W, seed=4)
    vl, tr = cfg["voxel_length"], cfg["truncation"]
    grids = {}
    for st in ("f16", "f32"):
        g = GPUVoxelGrid(1 << 30, tr, vl, H, W, n_buckets=1 << 21, n_bricks=1 << 20, storage=st)
        g.setIntrinsic(*intr)
        grids[st] = g
    o = oracle.Oracle64Grid(vl, tr, H, W, intr, CROP[0], CROP[1])
    for d, T in zip(depth, poses):
        Tf = synth.pose_f32(T)
        for g in grids.values():
            g.integrate(None, dev_u16(d), Tf)
        o.integrate(d, Tf.astype(np.float64))
    e64 = o.export()
    assert len(e64["coords"]) > 2000
    idx64 = _index(e64["coords"])
    report = {"frames": len(poses), "resolution": [W, H], "voxel_length": vl, "crop": CROP,
              "fp64_bricks": int(len(e64["coords"]))}
    ex = {}
    for st, g in grids.items():
        e = g.export_bricks()
        ex[st] = {k: v.cpu().numpy() for k, v in e.items()}
    # fp16 and fp32 storage see the same measurements: identical weights
    i16, i32 = _index(ex["f16"]["coords"]), _index(ex["f32"]["coords"])
    assert set(i16) == set(i32)
    order = np.array([i32[c] for c in i16])
    assert np.array_equal(ex["f16"]["w"], ex["f32"]["w"][order])
    obs = ex["f16"]["w"] > 0
    st_err = np.abs(ex["f16"]["sdf"] - ex["f32"]["sdf"][order])[obs]
    report["storage_f16_vs_f32"] = {"voxels": int(obs.sum()), "max_abs_sdf_err": float(st_err.max()),
                                    "rms_sdf_err": float(np.sqrt((st_err ** 2).mean()))}
    for st in ("f32", "f16"):
        e = ex[st]
        rows = [(k, idx64[c]) for c, k in _index(e["coords"]).items() if c in idx64]
        a = np.array([r[0] for r in rows])
        b = np.array([r[1] for r in rows])
        wg, wo = e["w"][a], e64["w"][b]
        same = (wg > 0) & (wg == wo)
        assert same.sum() > 0.98 * (wo > 0).sum()
        err = np.abs(e["sdf"][a][same] - e64["sdf"][b][same])
        report[st] = {"voxels": int(same.sum()), "max_abs_sdf_err": float(err.max()),
                      "p99_9_abs_sdf_err": float(np.percentile(err, 99.9)),
                      "rms_sdf_err": float(np.sqrt((err ** 2).mean()))}

    # depth of the crop: fp16 map (GPU, fp32 render) vs fp64 map (fp64 render)
    f64 = oracle.OracleField(e64["coords"], e64["sdf"].astype(np.float32), e64["w"].astype(np.float32), vl)
    vh, vw = 180, 320
    vi = synth.scaled_intr(cfg, vh, vw)
    views = np.stack(poses[5:80:15])
    dg = dr.render(grids["f16"], views, vh, vw, vi).cpu().numpy()
    d32 = dr.render(grids["f32"], views, vh, vw, vi).cpu().numpy()
    diffs, diffs32 = [], []
    for k, T in enumerate(views):
        do, _ = f64.render(synth.pose_f32(T).astype(np.float64), vh, vw, vi)
        m = (do > 0) & (dg[k] > 0) & (np.abs(dg[k] - do) < 1e-2)  # same surface, inside the crop
        diffs.append(dg[k][m] - do[m])
        diffs32.append(d32[k][m] - do[m])
    diffs, diffs32 = np.concatenate(diffs), np.concatenate(diffs32)
    assert diffs.size > 5000
    report["render_depth_vs_fp64"] = {
        "pixels": int(diffs.size),
        "rmse_f16": float(np.sqrt((diffs ** 2).mean())), "max_f16": float(np.abs(diffs).max()),
        "rmse_f32": float(np.sqrt((diffs32 ** 2).mean())), "max_f32": float(np.abs(diffs32).max())}
    if os.path.isdir("gpurun_out"):
        with open(os.path.join("gpurun_out", "fp16_accuracy.json"), "w") as fh:
            json.dump(report, fh, indent=1)
    for g in grids.values():
        g.close()
    assert report["storage_f16_vs_f32"]["max_abs_sdf_err"] < 3e-5
    assert report["f32"]["p99_9_abs_sdf_err"] < 1e-6
    assert report["f16"]["p99_9_abs_sdf_err"] < 2e-5
    assert report["f16"]["max_abs_sdf_err"] < tr and report["f32"]["max_abs_sdf_err"] < tr
    assert report["render_depth_vs_fp64"]["rmse_f16"] < 5e-5
