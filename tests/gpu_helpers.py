"""Shared helpers of the GPU parity tests: run the HIP path through the C ABI
(via superscanner_amd.voxel_grid) and the oracle on the same inputs, and
compare the results bit for bit."""
import os

import numpy as np
import torch

import oracle
from superscanner_amd.voxel_grid import GPUVoxelGrid


def oracle_threads() -> int:
    """The host cores this process is allotted (OMP_NUM_THREADS where set):
    the oracle's threaded block collection gives the serial loop's result
    (tests/test_oracle_parallel.py), so the parity tests use them all."""
    env = os.environ.get("OMP_NUM_THREADS")
    return int(env) if env and env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))


def dev_u16(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int16)).cuda()


def dev_u8(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint8)).cuda()


def make_pair(vl, trunc, H, W, intr, *, with_color=False, max_weight=0.0, n_buckets=1 << 18,
              n_bricks=1 << 17, storage="f32"):
    g = GPUVoxelGrid(n_bricks * 512, trunc, vl, H, W, n_buckets=n_buckets, n_bricks=n_bricks,
                     withColor=with_color, max_weight=max_weight, storage=storage)
    g.setIntrinsic(*intr)
    oracle.set_threads(oracle_threads())
    o = oracle.OracleGrid(vl, trunc, H, W, max_bricks=n_bricks, with_color=with_color,
                          max_weight=max_weight)
    o.set_intrinsics(*intr)
    return g, o


def sorted_export_gpu(g):
    ex = {k: v.cpu().numpy() for k, v in g.export_bricks().items()}
    order = np.lexsort((ex["coords"][:, 2], ex["coords"][:, 1], ex["coords"][:, 0]))
    return {k: v[order] for k, v in ex.items()}


def sorted_export_oracle(o):
    ex = o.export()
    order = np.lexsort((ex["coords"][:, 2], ex["coords"][:, 1], ex["coords"][:, 0]))
    return {k: (v[order] if v is not None else None) for k, v in ex.items()}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_bits_equal(got, exp, what):
    """Equal bit for bit, except that any NaN equals any NaN: a 0 / 0 of the
    reference's arithmetic (a zero crossing between two zero sdf samples)
    is a NaN on every platform, but its sign and payload are the
    platform's (x86 gives 0xffc00000, gfx950 0x7fc00000)."""
    got, exp = np.asarray(got), np.asarray(exp)
    assert got.shape == exp.shape, f"{what}: shape {got.shape} vs {exp.shape}"
    bad = (bits(got) != bits(exp)) & ~(np.isnan(got) & np.isnan(exp))
    if bad.any():
        idx = np.argwhere(bad)[:5]
        detail = ", ".join(f"{tuple(i)}: {got[tuple(i)]!r} vs {exp[tuple(i)]!r}" for i in idx)
        raise AssertionError(f"{what}: {int(bad.sum())} of {bad.size} values differ; first: {detail}")


def assert_grids_equal(g, o, color=False):
    eg, eo = sorted_export_gpu(g), sorted_export_oracle(o)
    assert eg["coords"].shape == eo["coords"].shape, (eg["coords"].shape, eo["coords"].shape)
    assert (eg["coords"] == eo["coords"]).all(), "allocated block sets differ"
    assert_bits_equal(eg["w"], eo["w"], "weights")
    assert_bits_equal(eg["sdf"], eo["sdf"], "sdf")
    if color:
        assert_bits_equal(eg["rgb"], eo["rgb"], "rgb")
    return eg, eo


def assert_raycast_equal(g, o, E, color=False):
    g.raycast(E, fence=True)
    r = o.raycast(E)
    ub_g = g.rayUB.cpu().numpy()
    assert (ub_g == r["ub"]).all(), f"UB flags differ at {int((ub_g != r['ub']).sum())} pixels"
    ok = r["ub"] == 0
    assert_bits_equal(g.rayPcd.cpu().numpy()[ok], r["points"][ok], "raycast points")
    assert_bits_equal(g.rayDepth.cpu().numpy()[ok], r["depth"][ok], "raycast depth")
    if color:
        assert_bits_equal(g.rayColorPcd.cpu().numpy()[ok], r["colors"][ok], "raycast colours")
    return r
