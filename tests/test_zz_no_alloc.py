"""tsdf_icp_moments_noalloc -- the ICP iteration sums without any device
allocation (its partial sums live in a static device array) -- against
tsdf_icp_moments, which allocates its partial sums per call: bit-identical
sums on two calls in a row, the second on a smaller image, and a refused
oversized image.  Runs through the C ABI."""
import ctypes as C

import numpy as np
import pytest
import torch

from superscanner_amd import _lib, icp, synth

pytestmark = pytest.mark.gpu


def pair(h, w, seed):
    f = 525.0 * w / 640
    intr = (f, f, w / 2 - 0.5, h / 2 - 0.5)
    room = synth.scene_room(seed=1)
    T1 = synth.look_at((0.5, 0.3, 1.4), (2.4, 1.9, 0.8))
    M = np.eye(4)
    M[:3, :3] = synth.axis_angle((0.3, 1.0, 0.2), np.deg2rad(1.5))
    M[:3, 3] = (0.01, -0.008, 0.012)
    T2 = T1 @ M
    d1 = synth.render_depth(room, T1, h, w, intr, noise_mm=1.0, max_range_m=3.5, seed=seed, keep=0.97)
    d2 = synth.render_depth(room, T2, h, w, intr, noise_mm=1.0, max_range_m=3.5, seed=seed + 1, keep=0.97)
    Txy = np.linalg.inv(T2) @ T1
    x = icp.CreatePcdFromDepthView(d1, intr)
    y = icp.CreatePcdFromDepthView(d2, intr)
    xn = icp.associate(x, y, None, Txy, intr)
    return x, xn, Txy


def test_noalloc_moments_equal_allocating_ones(gpu):
    for (h, w), seed in (((480, 640), 1), ((97, 131), 5)):  # the second call is smaller
        x, xn, Txy = pair(h, w, seed)
        want = icp.moments(x, xn, Txy)
        got = icp.moments_noalloc(x, xn, Txy)
        assert want[1] > 0.5 * want[0] > 0
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (h, w, got - want)
    # the same call again after a different one gives the same bits
    assert np.array_equal(icp.moments_noalloc(x, xn, Txy).view(np.uint64), want.view(np.uint64))
    # on another stream too
    s2 = torch.cuda.Stream()
    with torch.cuda.stream(s2):
        again = icp.moments_noalloc(x, xn, Txy, stream=s2)
    assert np.array_equal(again.view(np.uint64), want.view(np.uint64))


def test_noalloc_moments_refuse_oversized_images(gpu):
    L = _lib.load()
    out = np.zeros(_lib.ICP_MOMENTS, np.float64)
    T = np.eye(4).reshape(16)
    dp = C.POINTER(C.c_double)
    dummy = C.c_void_p(1)  # refused before any device access
    st = L.tsdf_icp_moments_noalloc(dummy, dummy, T.ctypes.data_as(dp), 1 << 15, 1 << 14,
                                    out.ctypes.data_as(dp), None)
    assert st == _lib.TSDF_BAD_ARG
    assert b"2^29" in L.tsdf_last_error()
