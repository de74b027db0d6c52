"""The reintegration scheduler driving the GPU grid, against the oracle
replaying the reference's sequence of calls (oracle/sched.py for the poses,
the C oracle for the grid):

  * a session integrates drifting odometry poses, a loop closure moves the
    world poses, updateReintegrationBuffer picks the frames with
    diff >= treint and the batched drain reintegrates them in one launch: the
    grid equals the oracle's (plain deIntegrate(getT(dofs)), integrate(T) per
    picked frame, in the buffer's order) bit for bit, and so does the raycast;
  * the same session through the reference's one-frame-per-pass loop (step)
    gives the same grid as the batched drain;
  * with colour, and with a frame removed (deintegrated) after the closure.
"""
import numpy as np
import pytest

from gpu_helpers import assert_grids_equal, assert_raycast_equal, dev_u8, dev_u16, make_pair, sorted_export_gpu
from oracle import sched as osched
from superscanner_amd import integration as I
from superscanner_amd import synth

pytestmark = pytest.mark.gpu

H, W = 240, 424


def session(n=24, seed=3):
    cfg = synth.ROOM_LARGE
    intr = synth.scaled_intr(cfg, H, W)
    truth = synth.traj_random_walk(n=n, seed=seed)
    depth = synth.frames(synth.scene_room(seed=1), truth, dict(cfg, intr=intr), seed=seed, H=H, W=W)
    rng = np.random.default_rng(seed + 10)
    odo, drift = [], np.eye(4)
    for T in truth:
        drift = drift @ synth.perturb(np.eye(4), rng, sigma_t=0.004, sigma_deg=0.4)
        odo.append(T @ drift)
    return cfg, intr, truth, odo, depth


def run(cfg, intr, truth, odo, depth, *, batched, color=False, remove=False):
    g, o = make_pair(cfg["voxel_length"], cfg["truncation"], H, W, intr, with_color=color)
    rng = np.random.default_rng(1)
    rgb = rng.integers(0, 256, size=(len(depth), H, W, 3), dtype=np.uint8) if color else None
    s = I.IntegrationScheduler(g, with_color=color)
    frames = []
    for k, (d, T) in enumerate(zip(depth, odo)):
        f = I.Frame(depth=dev_u16(d), rgb=dev_u8(rgb[k]) if color else None, unique_id=k)
        f.setFrametoWorldTrans(T)
        frames.append(f)
        s.push(f)
        o.integrate(d, T.astype(np.float32), 1, rgb[k] if color else None)
    s.run_batched() if batched else [None for _ in iter(s.step, False)]
    # loop closure: the world poses become the true ones
    for f, T in zip(frames, truth):
        f.setFrametoWorldTrans(T)
    s.updateReintegrationBuffer()
    picked = [f.unique_id for f in s.reintegrationBuffer]
    if remove:
        # a frame the closure left alone, so its deintegration commutes with
        # nothing the batch reorders
        remove = next(k for k in range(len(frames)) if k not in picked)
        s.removeFrameFromIntegration(frames[remove])
    if batched:
        s.run_batched()
    else:
        while s.step():
            pass
    # the reference's calls, replayed on the oracle
    for k in picked:
        old = osched.dof_to_mat(osched.mat_to_dof(odo[k])).astype(np.float32)
        o.integrate(depth[k], old, -1, rgb[k] if color else None)
        o.integrate(depth[k], truth[k].astype(np.float32), 1, rgb[k] if color else None)
    if remove is not False:
        old = osched.dof_to_mat(osched.mat_to_dof(odo[remove])).astype(np.float32)
        o.integrate(depth[remove], old, -1, rgb[remove] if color else None)
    return g, o, picked


@pytest.mark.parametrize("color,remove", [(False, False), (True, False), (False, True)])
def test_batched_session_equals_oracle(gpu, color, remove):
    cfg, intr, truth, odo, depth = session()
    g, o, picked = run(cfg, intr, truth, odo, depth, batched=True, color=color, remove=remove)
    assert 4 <= len(picked) < len(depth)
    assert_grids_equal(g, o, color=color)
    assert_raycast_equal(g, o, truth[len(truth) // 2].astype(np.float32), color=color)
    g.close()


def test_loop_and_batched_sessions_agree(gpu):
    cfg, intr, truth, odo, depth = session()
    ga, _, pa = run(cfg, intr, truth, odo, depth, batched=False)
    gb, _, pb = run(cfg, intr, truth, odo, depth, batched=True)
    assert pa == pb
    ea, eb = sorted_export_gpu(ga), sorted_export_gpu(gb)
    assert np.array_equal(ea["coords"], eb["coords"])
    assert np.array_equal(ea["w"].view(np.uint32), eb["w"].view(np.uint32))
    assert np.array_equal(ea["sdf"].view(np.uint32), eb["sdf"].view(np.uint32))
    ga.close()
    gb.close()
