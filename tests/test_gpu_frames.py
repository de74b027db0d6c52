"""Frame ingest on the GPU (superscanner_amd/frames.py): a recorded sequence
written to disk (16-bit depth PNG, colour PNG, intrinsic.txt) is read back
through the prefetching reader, uploaded from pinned memory on a copy
stream, and integrated; the grid equals the oracle's integration of the
arrays that were written, bit for bit.  MovetoGPU / MovetoCPU round-trip a
frame and keep the device-frame count."""
import numpy as np
import pytest
import torch

from gpu_helpers import assert_grids_equal, make_pair
from superscanner_amd import frames as F
from superscanner_amd import synth
from test_frames import write_sequence

pytestmark = pytest.mark.gpu


def test_sequence_reader_feeds_the_grid(gpu, tmp_path):
    H, W, n = 240, 424, 8
    cfg = synth.ROOM_LARGE
    intr = synth.scaled_intr(cfg, H, W)
    poses = [synth.pose_f32(T) for T in synth.traj_random_walk(n=n, seed=3)]
    depth = synth.frames(synth.scene_room(seed=1), synth.traj_random_walk(n=n, seed=3), dict(cfg, intr=intr),
                         seed=3, H=H, W=W)
    rng = np.random.default_rng(2)
    rgbs = [rng.integers(0, 256, size=(H, W, 3), dtype=np.uint8) for _ in range(n)]
    write_sequence(str(tmp_path), depth, rgbs, padded=True)
    K = np.array([[intr[0], 0, intr[2]], [0, intr[1], intr[3]], [0, 0, 1]])
    F.write_intrinsic(tmp_path / "intrinsic.txt", W, H, K)
    rd = F.SequenceReader(str(tmp_path), 0, n, H, W, prefetch=3, workers=2)
    assert (rd.width, rd.height) == (W, H)
    fx, fy, cx, cy = F.intrinsic_params(rd.K)
    g, o = make_pair(cfg["voxel_length"], cfg["truncation"], H, W, (fx, fy, cx, cy), with_color=True)
    before = F.HostFrame.n_gpu_frames
    seen = []
    for f in rd:
        assert f.dataOnGPU
        g.integrate(f.rgb, f.depth, poses[f.unique_id])
        o.integrate(depth[f.unique_id], poses[f.unique_id], 1, rgbs[f.unique_id])
        seen.append(f.unique_id)
        assert torch.equal(f.depth.cpu(), torch.from_numpy(depth[f.unique_id].view(np.int16)))
        f.MovetoCPU()
    rd.close()
    assert seen == list(range(n)) and F.HostFrame.n_gpu_frames == before
    assert_grids_equal(g, o, color=True)
    g.close()


def test_move_round_trip(gpu):
    rng = np.random.default_rng(0)
    d = rng.integers(0, 65536, size=(37, 53)).astype(np.uint16)
    c = rng.integers(0, 256, size=(37, 53, 3)).astype(np.uint8)
    f = F.HostFrame(d, c, unique_id=4)
    before = F.HostFrame.n_gpu_frames
    s = torch.cuda.Stream()
    f.MovetoGPU(s)
    f.MovetoGPU(s)  # already there: no second copy, no second count
    assert F.HostFrame.n_gpu_frames == before + 1
    f.wait()
    assert np.array_equal(f.depth.cpu().numpy().view(np.uint16), d)
    f.depth[0, 0] = 7  # changed on the device: MovetoCPU brings it back
    f.MovetoCPU()
    assert not f.dataOnGPU and F.HostFrame.n_gpu_frames == before
    assert int(f.depth_h[0, 0]) == 7 and np.array_equal(f.rgb_h.numpy(), c)
    assert np.array_equal(f.depth_h.numpy().view(np.uint16)[1:], d[1:])
