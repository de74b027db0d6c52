"""The benchmark workloads at their own sizes, on the GPU through the C ABI,
against the oracle:

  * room schedule (848x480, vl 4 mm, trunc 20 mm, 2^22 buckets, max weight
    128): the timed sequence's own seeds and call order -- integrate frame i,
    deintegrate frame i - 10 at every 10th pose, raycast -- over its first 70
    poses, and over all 600 (where weights reach the cap), with and without
    colour; grid, counters and raycasts bit-identical;
  * batched reintegration of 64 keyframes (848x480, valid fraction U(0.3, 1)
    per keyframe, poses perturbed by N(1 cm, 0.5 deg)): one
    tsdf_reintegrate_batch (one captured graph) against the oracle's 128
    plain calls in the order of integrate.cpp:108-125, with and without
    colour; bit-identical;
  * fp16 storage at 1280x720 with 2^24 buckets on the multi-room scene
    against the fp64 oracle on a crop (error bounds stated in the test);
  * a batch of 300 ops (more than one staging chunk of 256) against the
    plain calls; the fp16 weight cap; colour refused by a grid without it;
  * the timed workload's 600 frames integrated and deintegrated again: every
    voxel back to zero (a size-independent property at the full size).
"""
import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import assert_grids_equal, assert_raycast_equal, dev_u16, dev_u8, make_pair
from superscanner_amd import _lib, synth
from superscanner_amd.voxel_grid import GPUVoxelGrid

pytestmark = pytest.mark.gpu


def test_room_schedule_first_70_poses(gpu):
    cfg = synth.ROOM_LARGE
    n = 70
    poses64 = synth.traj_random_walk(n=n, seed=2)  # the first 70 poses of the 600-pose walk
    depth = synth.frames(synth.scene_room(seed=1), poses64, cfg, seed=7)
    poses = [synth.pose_f32(T) for T in poses64]
    g, o = make_pair(cfg["voxel_length"], cfg["truncation"], cfg["H"], cfg["W"], cfg["intr"],
                     max_weight=cfg["max_weight"], n_buckets=cfg["n_buckets"], n_bricks=1 << 21)
    tot = dict(n_touched=0, n_new=0, n_upd=0, n_valid=0)
    ray_at = {25, 50, n - 1}
    samples = probes = 0
    before = None
    for i in range(n):
        ops = [(i, 1)] + ([(i - 10, -1)] if i >= 10 and i % 10 == 0 else [])
        for k, s in ops:
            (g.integrate if s > 0 else g.deIntegrate)(None, dev_u16(depth[k]), poses[k])
            st, c = o.integrate(depth[k], poses[k], s)
            assert st == 0
            for key in tot:
                tot[key] += c[key]
        if i in ray_at:
            if before is None:
                g.sync()
                before = g.stats()
            r = assert_raycast_equal(g, o, poses[i])
            samples += r["counts"]["samples"]
            probes += r["counts"]["probes"]
            assert r["counts"]["n_hit"] > 0.5 * cfg["H"] * cfg["W"]
    g.sync()
    s = g.stats()
    assert s["overflow"] == 0 and s["n_bricks"] == o.n_bricks()
    assert s["total_touched"] == tot["n_touched"] and s["total_new"] == tot["n_new"]
    assert s["total_updated"] == tot["n_upd"] and s["total_valid_px"] == tot["n_valid"]
    assert s["ray_samples"] - before["ray_samples"] == samples
    assert s["ray_probes"] - before["ray_probes"] == probes
    assert_grids_equal(g, o)
    g.close()


_ROOM_600 = {}


def room_600():
    """The timed workload's 600 frames (generated once per module)."""
    if not _ROOM_600:
        cfg = synth.ROOM_LARGE
        poses64 = synth.traj_random_walk(n=600, seed=2)
        _ROOM_600["depth"] = synth.frames(synth.scene_room(seed=1), poses64, cfg, seed=7)
        _ROOM_600["poses"] = [synth.pose_f32(T) for T in poses64]
    return synth.ROOM_LARGE, _ROOM_600["poses"], _ROOM_600["depth"]


@pytest.mark.parametrize("color", [False, True])
def test_room_schedule_all_600_poses(gpu, color):
    """The whole timed step (600 poses: integrate frame i, deintegrate frame
    i - 10 at every 10th pose) against the oracle, with raycasts at three
    poses: over 600 frames many voxels reach the weight cap of 128, so the
    cap's rule runs at scale.  Grid, counters and raycasts bit-identical;
    with colour (the reference's default, GPUVoxelGrid.h:99) the running
    colour averages and the raycast colours too."""
    cfg, poses, depth = room_600()
    n = 600
    H, W = cfg["H"], cfg["W"]
    g, o = make_pair(cfg["voxel_length"], cfg["truncation"], H, W, cfg["intr"], with_color=color,
                     max_weight=cfg["max_weight"], n_buckets=cfg["n_buckets"], n_bricks=1 << 17)
    rgbs = None
    if color:  # one colour image per frame, on the host for the oracle
        rgbs = np.random.default_rng(17).integers(0, 256, size=(n, H, W, 3), dtype=np.uint8)
    try:
        tot = dict(n_touched=0, n_new=0, n_upd=0)
        for i in range(n):
            ops = [(i, 1)] + ([(i - 10, -1)] if i >= 10 and i % 10 == 0 else [])
            for k, s in ops:
                rgb = rgbs[k] if color else None
                (g.integrate if s > 0 else g.deIntegrate)(dev_u8(rgb) if color else None, dev_u16(depth[k]),
                                                          poses[k])
                st, c = o.integrate(depth[k], poses[k], s, rgb)
                assert st == 0
                for key in tot:
                    tot[key] += c[key]
            if i in (299, 450, n - 1):
                r = assert_raycast_equal(g, o, poses[i], color=color)
                assert r["counts"]["n_hit"] > 0.5 * cfg["H"] * cfg["W"]
        g.sync()
        s = g.stats()
        assert s["overflow"] == 0 and s["n_bricks"] == o.n_bricks()
        assert s["total_touched"] == tot["n_touched"] and s["total_new"] == tot["n_new"]
        assert s["total_updated"] == tot["n_upd"]
        eg, _ = assert_grids_equal(g, o, color=color)
        assert float(eg["w"].max()) == cfg["max_weight"], "the cap was not reached"
    finally:
        g.close()


def keyframe_workload(keyframes=64):
    """The batched-reintegration workload: keyframes spread over the 600-pose
    walk, each with a random depth dropout, and their pose-graph updates."""
    cfg = synth.ROOM_LARGE
    poses64 = synth.traj_random_walk(n=600, seed=2)
    rng = np.random.default_rng(21)
    kf = np.linspace(0, len(poses64) - 1, keyframes).astype(int)
    keep = rng.uniform(0.3, 1.0, keyframes)
    kdepth = synth.frames(synth.scene_room(seed=1), [poses64[k] for k in kf], cfg, seed=31, keep=keep)
    old = [synth.pose_f32(poses64[k]) for k in kf]
    new = [synth.pose_f32(synth.perturb(poses64[k], rng)) for k in kf]
    return cfg, kdepth, old, new, keep


@pytest.mark.parametrize("color", [False, True])
def test_batched_reintegration_64_keyframes(gpu, color):
    cfg, kdepth, old, new, keep = keyframe_workload(64)
    valid = (kdepth > 0).mean(axis=(1, 2))
    assert valid.min() < 0.4 and valid.max() > 0.8  # uneven valid-pixel counts
    H, W = cfg["H"], cfg["W"]
    g, o = make_pair(cfg["voxel_length"], cfg["truncation"], H, W, cfg["intr"], with_color=color,
                     max_weight=cfg["max_weight"], n_buckets=cfg["n_buckets"], n_bricks=1 << 18)
    kd = [dev_u16(d) for d in kdepth]
    rgb = np.random.default_rng(23).integers(0, 256, size=(len(kd), H, W, 3), dtype=np.uint8) if color else None
    kc = [dev_u8(c) for c in rgb] if color else None
    for k in range(len(kd)):
        g.integrate(kc[k] if color else None, kd[k], old[k])
        o.integrate(kdepth[k], old[k], 1, rgb[k] if color else None)
    g.sync()
    s0 = g.stats()
    g.reintegrate_batch(kd, old, new, rgbs=kc)
    for k in range(len(kd)):  # integrate.cpp:108-125: deIntegrate(old) then integrate(new)
        o.integrate(kdepth[k], old[k], -1, rgb[k] if color else None)
        o.integrate(kdepth[k], new[k], 1, rgb[k] if color else None)
    g.sync()
    s1 = g.stats()
    assert s1["overflow"] == 0 and s1["n_bricks"] == o.n_bricks()
    assert s1["total_updated"] > s0["total_updated"]
    assert_grids_equal(g, o, color=color)
    for k in (0, 31, 63):
        assert_raycast_equal(g, o, new[k], color=color)
    g.close()


def test_fp16_full_resolution_against_fp64(gpu):
    """1280x720 frames, vl 2 mm, trunc 8 mm, 2^24 buckets, fp16 storage,
    against the fp64 oracle on the voxels of a crop that both saw equally
    often.

    Error bound: every update stores the running average rounded to fp16,
    at most half an ulp u of a value below the truncation (|sdf| < 8 mm:
    u = 2^-18 m); the running average scales the error of step k by k / n
    after n observations, so |error| <= u (1 + 2 + ... + n) / n = u (n + 1) / 2
    plus fp32 arithmetic (1e-6 m).  At least 99.9 % of the voxels keep it;
    the rest, and the maximum (below the truncation), are voxels whose fp32
    projection lands on the neighbouring pixel of the fp64 one.  Rendered
    depth RMSE against the fp64 render of the fp64 map: below 5e-5 m."""
    cfg = dict(synth.FP16_LARGE)
    H, W = cfg["H"], cfg["W"]
    poses = synth.traj_multiroom(2000)[:240:4]
    depth = synth.frames(synth.scene_multiroom(), poses, cfg, seed=4)
    vl, tr = cfg["voxel_length"], cfg["truncation"]
    g = GPUVoxelGrid(1 << 30, tr, vl, H, W, n_buckets=cfg["n_buckets"], n_bricks=1 << 21, storage="f16")
    g.setIntrinsic(*cfg["intr"])
    assert g.get_params()["max_weight"] == 2048.0  # fp16 counts weights exactly up to 2048
    crop = ((-3.0, -3.6, 0.0), (-1.9, -0.4, 1.4))
    o = oracle.Oracle64Grid(vl, tr, H, W, cfg["intr"], crop[0], crop[1])
    for d, T in zip(depth, poses):
        Tf = synth.pose_f32(T)
        g.integrate(None, dev_u16(d), Tf)
        o.integrate(d, Tf.astype(np.float64))
    g.sync()
    assert g.stats()["overflow"] == 0
    e64 = o.export()
    assert len(e64["coords"]) > 10000
    e = {k: v.cpu().numpy() for k, v in g.export_bricks().items()}
    idx = {tuple(c): k for k, c in enumerate(e["coords"].tolist())}
    rows = [(idx[tuple(c)], k) for k, c in enumerate(e64["coords"].tolist()) if tuple(c) in idx]
    assert len(rows) == len(e64["coords"])  # every crop block is allocated on the GPU
    a = np.array([r[0] for r in rows])
    b = np.array([r[1] for r in rows])
    wg, wo = e["w"][a], e64["w"][b]
    same = (wg > 0) & (wg == wo)
    assert same.sum() > 0.98 * (wo > 0).sum()
    err = np.abs(e["sdf"][a][same] - e64["sdf"][b][same])
    bound = 2.0 ** -18 * (wg[same] + 1.0) / 2.0 + 1e-6
    within = (err <= bound).mean()
    assert within >= 0.999, within
    assert err.max() < tr
    # depth rendered from the fp16 map (fp32) against the fp64 render of the
    # fp64 map, over pixels that see the crop's surfaces
    from superscanner_amd import diff_render as dr
    f64 = oracle.OracleField(e64["coords"], e64["sdf"].astype(np.float32), e64["w"].astype(np.float32), vl)
    vh, vw = 180, 320
    vi = synth.scaled_intr(cfg, vh, vw)
    views = np.stack(poses[5:60:10])
    dg = dr.render(g, views, vh, vw, vi).cpu().numpy()
    diffs = []
    for k, T in enumerate(views):
        do, _ = f64.render(synth.pose_f32(T).astype(np.float64), vh, vw, vi)
        m = (do > 0) & (dg[k] > 0) & (np.abs(dg[k] - do) < 1e-2)
        diffs.append(dg[k][m] - do[m])
    diffs = np.concatenate(diffs)
    assert diffs.size > 5000
    rmse = float(np.sqrt((diffs ** 2).mean()))
    assert rmse < 5e-5, rmse
    report = {"frames": len(poses), "resolution": [W, H], "n_buckets": cfg["n_buckets"], "crop": crop,
              "voxels": int(same.sum()), "max_abs_sdf_err": float(err.max()),
              "p99_9_abs_sdf_err": float(np.percentile(err, 99.9)), "rms_sdf_err": float(np.sqrt((err ** 2).mean())),
              "within_rounding_bound": float(within), "render_depth_rmse": rmse, "render_pixels": int(diffs.size)}
    import json
    import os
    if os.path.isdir("gpurun_out"):
        with open(os.path.join("gpurun_out", "fp16_accuracy_1280x720.json"), "w") as fh:
            json.dump(report, fh, indent=1)
    g.close()


def test_batch_of_300_ops_equals_plain_calls(gpu):
    """150 keyframes (300 ops: two staging chunks of at most 256 ops, run as
    separate launches) give the plain call sequence bit for bit."""
    cfg = synth.ROOM_LARGE
    H, W = 120, 212
    intr = synth.scaled_intr(cfg, H, W)
    n = 150
    poses64 = synth.traj_random_walk(n=n, seed=5)
    rng = np.random.default_rng(3)
    depth = synth.frames(synth.scene_room(seed=1), poses64, dict(cfg, intr=intr), seed=5, H=H, W=W,
                         keep=rng.uniform(0.3, 1.0, n))
    poses = [synth.pose_f32(T) for T in poses64]
    new = [synth.pose_f32(synth.perturb(T, rng)) for T in poses64]
    a, o = make_pair(cfg["voxel_length"], cfg["truncation"], H, W, intr, max_weight=cfg["max_weight"])
    b, _ = make_pair(cfg["voxel_length"], cfg["truncation"], H, W, intr, max_weight=cfg["max_weight"])
    d = [dev_u16(x) for x in depth]
    for k in range(n):
        a.integrate(None, d[k], poses[k])
        b.integrate(None, d[k], poses[k])
        o.integrate(depth[k], poses[k])
    for k in range(n):
        a.deIntegrate(None, d[k], poses[k])
        a.integrate(None, d[k], new[k])
        o.integrate(depth[k], poses[k], -1)
        o.integrate(depth[k], new[k], 1)
    b.reintegrate_batch(d, poses, new)
    assert_grids_equal(a, o)
    assert_grids_equal(b, o)
    a.close()
    b.close()


def test_fp16_weight_cap(gpu):
    """A fp16 grid created without a cap counts to 2048 and stays there
    (2049 is not a fp16 value); the weights equal the oracle's with cap 2048;
    a cap above 2048 is refused."""
    H, W = 24, 32
    intr = (30.0, 30.0, 15.5, 11.5)
    g = GPUVoxelGrid(1 << 20, 0.02, 0.005, H, W, n_buckets=1 << 14, n_bricks=4096, storage="f16")
    g.setIntrinsic(*intr)
    assert g.get_params()["max_weight"] == 2048.0
    o = oracle.OracleGrid(0.005, 0.02, H, W, max_weight=2048.0)
    o.set_intrinsics(*intr)
    E = np.eye(4, dtype=np.float32)
    rng = np.random.default_rng(1)
    frames = [np.full((H, W), 800, np.uint16) + rng.integers(0, 4, (H, W)).astype(np.uint16) for _ in range(4)]
    dev = [dev_u16(f) for f in frames]
    for k in range(2055):
        g.integrate(None, dev[k % 4], E)
        o.integrate(frames[k % 4], E)
    g.sync()
    from gpu_helpers import sorted_export_gpu, sorted_export_oracle
    eg, eo = sorted_export_gpu(g), sorted_export_oracle(o)
    assert np.array_equal(eg["coords"], eo["coords"])
    assert np.array_equal(eg["w"], eo["w"]) and eg["w"].max() == 2048.0
    # sdf: fp16 storage rounds each of the n = 2055 running averages to half
    # an ulp u of a value below the truncation (20 mm: u = 2^-17 m); the
    # average scales step k's error by k / n, so |error| <= u (n + 1) / 2
    obs = eo["w"] > 0
    assert np.abs(eg["sdf"] - eo["sdf"])[obs].max() <= 2.0 ** -17 * 2056 / 2
    g.close()
    with pytest.raises(_lib.TsdfError) as e:
        GPUVoxelGrid(1 << 20, 0.02, 0.005, H, W, n_bricks=4096, storage="f16", max_weight=4096.0)
    assert e.value.status == _lib.TSDF_BAD_ARG


def test_colour_refused_without_colour_storage(gpu):
    H, W = 24, 32
    g = GPUVoxelGrid(1 << 20, 0.02, 0.005, H, W, n_buckets=1 << 14, n_bricks=4096)
    g.setIntrinsic(30.0, 30.0, 15.5, 11.5)
    d = dev_u16(np.full((H, W), 800, np.uint16))
    c = dev_u8(np.zeros((H, W, 3), np.uint8))
    E = np.eye(4, dtype=np.float32)
    with pytest.raises(_lib.TsdfError) as e:
        g.integrate(c, d, E)
    assert e.value.status == _lib.TSDF_BAD_ARG
    with pytest.raises(_lib.TsdfError) as e:
        g.integrate_batch([d], [E], [1], rgbs=[c])
    assert e.value.status == _lib.TSDF_BAD_ARG
    g.integrate(c, d, E, withColor=False)  # the reference's per-call withColor = false
    g.integrate(None, d, E, fence=True)
    assert g.size() > 0
    g.close()
    torch.cuda.synchronize()


@pytest.mark.parametrize("storage,color", [("f32", False), ("f32", True), ("f16", False)])
def test_room_600_frames_integrate_deintegrate_round_trip(gpu, storage, color):
    """The timed workload's 600 frames (848x480, 2^22 buckets) integrated,
    then deintegrated in reverse order: every voxel returns to w = 0 and
    sdf = 0 (and colour 0) exactly, whatever the float rounding of the
    running averages in between -- inside its band a call moves the weight by
    exactly 1, and the deintegrate at w = 1 sets weight, sdf and colour to 0
    (GPUVoxelGrid.cpp:367-395, 426-469) -- and the raycast then finds no
    surface.  A size-independent property at the full size; weights uncapped
    (fp32) or capped above the 600 calls (fp16, exact integers to 2048)."""
    cfg, poses, depth = room_600()
    n = 600
    H, W = cfg["H"], cfg["W"]
    g = GPUVoxelGrid(1 << 30, cfg["truncation"], cfg["voxel_length"], H, W, n_buckets=cfg["n_buckets"],
                     n_bricks=1 << 21, withColor=color, max_weight=0.0 if storage == "f32" else 2048.0,
                     storage=storage)
    g.setIntrinsic(*cfg["intr"])
    dd = torch.from_numpy(depth.view(np.int16)).cuda()

    colours = None
    if color:  # every frame's colour image, drawn once on the device
        gen = torch.Generator(device="cuda").manual_seed(1000)
        colours = torch.randint(0, 256, (n, H, W, 3), dtype=torch.uint8, device="cuda", generator=gen)

    def rgb(k):
        return colours[k] if color else None

    for k in range(n):
        g.integrate(rgb(k), dd[k], poses[k])
    g.sync()
    mid = g.export_bricks()
    n_bricks = mid["w"].shape[0]
    assert n_bricks > 20000 and float(mid["w"].max()) > 100.0  # voxels seen by many frames
    del mid
    for k in reversed(range(n)):
        g.deIntegrate(rgb(k), dd[k], poses[k])
    g.sync()
    assert g.stats()["overflow"] == 0
    ex = g.export_bricks()
    assert ex["w"].shape[0] == n_bricks
    assert int((ex["w"] != 0).sum()) == 0, "weights left after the round trip"
    sdf_bits = ex["sdf"].float().contiguous().view(torch.int32)
    assert int((sdf_bits != 0).sum()) == 0, "sdf left after the round trip (not +0.0)"
    if color:
        assert int((ex["rgb"] != 0).sum()) == 0, "colour left after the round trip"
    for k in (0, 299, 599):
        g.raycast(poses[k], fence=True)
        ub = g.rayUB.cpu().numpy()
        dep = g.rayDepth.cpu().numpy()
        assert int((dep[ub == 0] != 0).sum()) == 0, "a surface found in an all-zero map"
    g.close()
