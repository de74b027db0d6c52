"""The oracle's block collection runs on several threads (row bands of the
image, each with its own block set, merged in band order).  The serial loop
of vbCollectionTag (GPUVoxelGrid.cpp:471-514) inserts blocks in pixel order,
and that first-occurrence order fixes the brick ids and, when the pool runs
out, which blocks are refused.  These checks run the same calls on 1 and on
several threads and require identical results: exported coordinates in id
order, keys, sdf / w / colour bits, the per-call counters and the status of
an overflowing call.
"""
import numpy as np
import pytest

import oracle
from superscanner_amd import synth


@pytest.fixture(autouse=True)
def _restore_threads():
    yield
    oracle.set_threads(1)


def run(threads, max_bricks, with_color=False, n=6):
    oracle.set_threads(threads)
    H, W = 240, 424
    cfg = synth.ROOM_LARGE
    intr = synth.scaled_intr(cfg, H, W)
    poses = synth.traj_random_walk(n=n, seed=5)
    depth = synth.frames(synth.scene_room(seed=1), poses, dict(cfg, intr=intr), seed=6, H=H, W=W)
    rng = np.random.default_rng(9)
    o = oracle.OracleGrid(cfg["voxel_length"], cfg["truncation"], H, W, max_bricks=max_bricks,
                          with_color=with_color, max_weight=128.0)
    o.set_intrinsics(*intr)
    log = []
    for k in range(n):
        rgb = rng.integers(0, 256, size=(H, W, 3), dtype=np.uint8) if with_color else None
        log.append(o.integrate(depth[k], synth.pose_f32(poses[k]), 1, rgb))
        if k == 3:
            log.append(o.integrate(depth[1], synth.pose_f32(poses[1]), -1, rgb))
    ex = o.export()
    o.close()
    return log, ex


def assert_same(a, b):
    (la, ea), (lb, eb) = a, b
    assert la == lb, "statuses or counters differ"
    for k in ("coords", "keys", "sdf", "w", "rgb"):
        if ea[k] is None:
            assert eb[k] is None
            continue
        assert ea[k].shape == eb[k].shape, k
        assert np.array_equal(np.ascontiguousarray(ea[k]).view(np.uint8), np.ascontiguousarray(eb[k]).view(np.uint8)), k


@pytest.mark.parametrize("threads", [2, 8])
def test_threads_give_the_serial_result(threads):
    serial = run(1, 1 << 20, with_color=True)
    assert len(serial[1]["keys"]) > 1000
    assert_same(serial, run(threads, 1 << 20, with_color=True))


def test_threads_refuse_the_same_blocks_when_the_pool_fills():
    serial = run(1, 1500)
    statuses = [st for st, _ in serial[0]]
    assert any(st != 0 for st in statuses), "the pool should have filled"
    assert len(serial[1]["keys"]) == 1500
    assert_same(serial, run(8, 1500))
