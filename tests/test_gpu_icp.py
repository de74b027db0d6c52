"""Point maps, normal maps and projective ICP on the GPU (through the C ABI)
against the oracle (oracle/icp_oracle.c, pinned by test_icp_oracle.py):

  * depth -> points, points -> normals and the association of both types are
    bit-identical to the oracle (NaN normals compared as NaN), on rendered
    frames with holes, at 640 x 480 and at an odd size, through both the
    LDS-staged scan and the global-memory one;
  * the iteration sums agree with the oracle's reference-order sums (counts
    exact, centroids and the centred cross-covariance to 1e-9 relative, the
    RMSE, which the reference sums in float, to 1e-5 relative);
  * kokkosICP follows the oracle's loop (same per-iteration sums, numpy SVD
    there, polar iteration here): transform within 1e-5, RMSE within 1e-4
    relative, fitness within 1e-6, and keeps an exact alignment.
"""
import numpy as np
import pytest
import torch

from oracle import icp as oicp
from superscanner_amd import icp, synth

pytestmark = pytest.mark.gpu


def frames(h=480, w=640, f=525.0, noise=1.0, deg=1.5):
    intr = (f * w / 640, f * w / 640, w / 2 - 0.5, h / 2 - 0.5)
    room = synth.scene_room(seed=1)
    T1 = synth.look_at((0.5, 0.3, 1.4), (2.4, 1.9, 0.8))
    M = np.eye(4)
    M[:3, :3] = synth.axis_angle((0.3, 1.0, 0.2), np.deg2rad(deg))
    M[:3, 3] = (0.01, -0.008, 0.012)
    T2 = T1 @ M
    d1 = synth.render_depth(room, T1, h, w, intr, noise_mm=noise, max_range_m=3.5, seed=1, keep=0.97)
    d2 = synth.render_depth(room, T2, h, w, intr, noise_mm=noise, max_range_m=3.5, seed=2, keep=0.97)
    return d1, d2, T1, T2, intr


def bits_equal_nan(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


@pytest.mark.parametrize("size", [(480, 640), (97, 131)])
def test_points_normals_association_bit_exact(gpu, size):
    h, w = size
    d1, d2, T1, T2, intr = frames(h, w)
    assert (d1 == 0).any() and (d1 > 0).mean() > 0.5
    x = icp.CreatePcdFromDepthView(d1, intr)
    y = icp.CreatePcdFromDepthView(d2, intr)
    ox, oy = oicp.depth_to_points(d1, intr), oicp.depth_to_points(d2, intr)
    assert bits_equal_nan(x.cpu().numpy(), ox) and bits_equal_nan(y.cpu().numpy(), oy)
    yn = icp.CalculateNormalsFromPCDView(y)
    oyn = oicp.normals(oy)
    assert bits_equal_nan(yn.cpu().numpy(), oyn)
    Txy = np.linalg.inv(T2) @ T1
    turn = np.eye(4)
    turn[:3, :3] = synth.axis_angle((1.0, 0.2, -0.4), np.deg2rad(4.0))
    for T in (np.eye(4), Txy, Txy @ turn):
        # n_search 3 and 5 stage every 16 x 16 tile's window box in LDS
        # (at most 1024 candidates); at 8 (a box of 32 x 32 and more) the
        # widely spread tiles of the turned transform scan global memory, and
        # at 12 most tiles do
        for t, ns, md in ((icp.PointToPoint, 5, 0.075), (icp.PointToPlane, 5, 0.075), (icp.PointToPoint, 3, 0.02),
                          (icp.PointToPlane, 8, 0.075), (icp.PointToPoint, 12, 0.1)):
            got = icp.associate(x, y, yn, T, intr, icp_type=t, n_search=ns, max_dist=md).cpu().numpy()
            exp = oicp.associate(ox, oy, oyn, T, intr, icp_type=t, n_search=ns, max_dist=md)
            assert (exp[..., 2] != 0).mean() > 0.1
            assert bits_equal_nan(got, exp), (t, ns, md, int((got != exp).any(-1).sum()))


def test_moments_match_reference_sums(gpu):
    d1, d2, T1, T2, intr = frames()
    Txy = np.linalg.inv(T2) @ T1
    x = icp.CreatePcdFromDepthView(d1, intr)
    y = icp.CreatePcdFromDepthView(d2, intr)
    xn = icp.associate(x, y, None, Txy, intr)
    m = icp.moments(x, xn, Txy)
    s = oicp.sums(x.cpu().numpy(), xn.cpu().numpy(), Txy)
    assert m[0] == s[0] and m[1] == s[1] and m[1] > 1e5
    n = m[1]
    cx, cy = m[2:5] / n, m[5:8] / n
    assert np.abs(cx - s[2:5]).max() < 1e-9 * np.abs(s[2:5]).max()
    assert np.abs(cy - s[5:8]).max() < 1e-9 * np.abs(s[5:8]).max()
    Hm = m[8:17].reshape(3, 3) - n * np.outer(cx, cy)
    Hs = s[8:17].reshape(3, 3)
    assert np.abs(Hm - Hs).max() < 1e-9 * np.abs(Hs).max()
    rmse = np.sqrt(m[17] / n)
    assert abs(rmse - s[17]) < 1e-5 * s[17]
    # the reduction is deterministic
    assert np.array_equal(m, icp.moments(x, xn, Txy))


@pytest.mark.parametrize("icp_type", [0, 1])
def test_kokkos_icp_follows_the_oracle(gpu, icp_type):
    d1, d2, T1, T2, intr = frames(240, 320, noise=1.0)
    start = T1 @ np.array([[1, 0, 0, 0.004], [0, 1, 0, -0.003], [0, 0, 1, 0.002], [0, 0, 0, 1.0]])
    T, rmse, fit = icp.kokkosICP(d1, d2, start, T2, intr, ICPType=icp_type)
    To, rmse_o, fit_o, rec = oicp.kokkosICP(d1, d2, start, T2, intr, icp_type=icp_type)
    assert abs(rmse - rmse_o) < 1e-4 * rmse_o and abs(fit - fit_o) < 1e-6
    # the returned transform is the oracle's iteration of lowest RMSE (or one
    # within the float-summed RMSE's rounding of it)
    near = [r["T"] for r in rec if r["rmse"] <= rmse_o * (1 + 1e-4)]
    assert min(np.abs(T - t).max() for t in near) < 1e-5, [np.abs(T - t).max() for t in near]
    assert 0 < rmse < 0.02 and fit > 0.5


def test_kokkos_icp_keeps_an_exact_alignment(gpu):
    n, f = 239, 200.0
    intr = (f, f, 119.0, 120.0)
    T1 = synth.look_at((1.2, -0.9, 0.9), (0.0, 0.0, 0.2))
    d1 = synth.render_depth(synth.scene_plain_small(), T1, n, n, intr, noise_mm=0.0, max_range_m=4.0)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    d2 = np.zeros_like(d1)
    d2[jj, 238 - ii] = d1
    Rz = np.eye(4)
    Rz[:2, :2] = [[0.0, -1.0], [1.0, 0.0]]
    T, rmse, fit = icp.kokkosICP(d1, d2, Rz, np.eye(4), intr)
    assert np.abs(T - Rz).max() < 1e-6 and rmse < 1e-6 and fit > 0.5


def test_kokkos_icp_workspace_and_streams(gpu):
    """The per-device workspace: released, then grown from a small frame to
    a larger one (both follow the oracle); a call with no iteration returns
    the identity and waits for its own launches; a call on a second stream
    after it, and a call after another release, give the same bits."""
    icp.release()
    runs = []
    for h, w in ((97, 131), (240, 320)):
        d1, d2, T1, T2, intr = frames(h, w, noise=1.0)
        start = T1 @ np.array([[1, 0, 0, 0.004], [0, 1, 0, -0.003], [0, 0, 1, 0.002], [0, 0, 0, 1.0]])
        T, rmse, fit = icp.kokkosICP(d1, d2, start, T2, intr)
        To, rmse_o, fit_o, rec = oicp.kokkosICP(d1, d2, start, T2, intr)
        assert abs(rmse - rmse_o) < 1e-4 * rmse_o and abs(fit - fit_o) < 1e-6, (h, w)
        near = [r["T"] for r in rec if r["rmse"] <= rmse_o * (1 + 1e-4)]
        assert min(np.abs(T - t).max() for t in near) < 1e-5, (h, w)
        runs.append((d1, d2, start, T2, intr, T, rmse, fit))
    d1, d2, start, T2, intr, T, rmse, fit = runs[-1]
    T0, r0, f0 = icp.kokkosICP(d1, d2, start, T2, intr, n_it=0)
    assert np.array_equal(T0, np.eye(4)) and r0 == 1e6 and f0 == 0.0
    s2 = torch.cuda.Stream()
    with torch.cuda.stream(s2):
        Tb, rb, fb = icp.kokkosICP(d1, d2, start, T2, intr, stream=s2)
    assert np.array_equal(Tb, T) and rb == rmse and fb == fit
    icp.release()
    Tc, rc, fc = icp.kokkosICP(d1, d2, start, T2, intr)
    assert np.array_equal(Tc, T) and rc == rmse and fc == fit


def test_argument_errors(gpu):
    from superscanner_amd import _lib
    d = torch.zeros((8, 8), dtype=torch.int16, device=gpu)
    with pytest.raises(ValueError):
        icp.kokkosICP(d, torch.zeros((8, 9), dtype=torch.int16, device=gpu), np.eye(4), np.eye(4), (1, 1, 1, 1))
    with pytest.raises(_lib.TsdfError):
        icp.kokkosICP(d, d, np.eye(4), np.zeros((4, 4)), (1, 1, 1, 1))  # singular ty
    x = icp.CreatePcdFromDepthView(d, (1.0, 1.0, 3.5, 3.5))
    with pytest.raises(_lib.TsdfError):
        icp.associate(x, x, None, np.eye(4), (1.0, 1.0, 3.5, 3.5), icp_type=icp.PointToPlane)
    # no correspondence at all: identity and the 1e6 sentinel, as the reference returns them
    T, rmse, fit = icp.kokkosICP(d, d, np.eye(4), np.eye(4), (1.0, 1.0, 3.5, 3.5))
    assert np.array_equal(T, np.eye(4)) and rmse == 1e6 and fit == 0.0
