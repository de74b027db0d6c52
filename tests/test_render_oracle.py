"""Pins the fp64 CPU reference of the differentiable render (oracle
render_oracle.c) without a GPU:

  * closed forms on a fronto-parallel plane: depth = D, normal = (0, 0, -1),
    dD/dv = (0, 0, -1), dD/dw = (-y, x, 0) at the hit point (x, y, D);
  * central finite differences (h = 1e-4, fp64) of the rendered depth with
    respect to the 6 pose parameters and to single voxel values, on a scene
    with curved surfaces;
  * planted mistakes in the adjoint (rotation sign, missing 1/s) fail them;
  * the fp64 SE(3) helpers (exp, left Jacobian) and the C ABI's
    Gauss-Newton step (Cholesky solve + composition) against numpy.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from superscanner_amd import synth
from superscanner_amd.diff_render import se3_exp, se3_left_jacobian

INTR_V = (65.0, 65.0, 39.5, 29.5)  # 80 x 60 render view


def _plane_field():
    H, W = 120, 160
    g = oracle.OracleGrid(0.01, 0.04, H, W)
    g.set_intrinsics(130.0, 130.0, 79.5, 59.5)
    for _ in range(2):
        g.integrate(np.full((H, W), 1000, np.uint16), np.eye(4, dtype=np.float32))
    ex = g.export()
    return oracle.OracleField(ex["coords"], ex["sdf"], ex["w"], 0.01)


def _scene_field():
    """plain_small scene (plane, sphere, box) fused from 6 views at 160 x 120."""
    cfg = dict(synth.PLAIN_SMALL, noise_mm=0.0)
    H, W = 120, 160
    intr = synth.scaled_intr(cfg, H, W)
    poses = synth.traj_plain_small(n=24)[:6]
    depth = synth.frames(synth.scene_plain_small(), poses, dict(cfg, intr=intr), H=H, W=W)
    g = oracle.OracleGrid(0.01, 0.04, H, W)
    g.set_intrinsics(*intr)
    for d, T in zip(depth, poses):
        g.integrate(d, synth.pose_f32(T))
    ex = g.export()
    return oracle.OracleField(ex["coords"], ex["sdf"], ex["w"], 0.01), poses, ex


def _ray_dirs(h, w, intr):
    fx, fy, cx, cy = intr
    j, i = np.meshgrid(np.arange(w), np.arange(h))
    return (j - cx) / fx, (i - (h - cy)) / fy


def test_plane_closed_form():
    f = _plane_field()
    dep, nrm = f.render(np.eye(4), 60, 80, INTR_V, normals=True)
    hit = dep > 0
    assert hit.mean() > 0.95
    np.testing.assert_allclose(dep[hit], 1.0, atol=1e-6)
    np.testing.assert_allclose(nrm[hit], np.broadcast_to([0.0, 0.0, -1.0], nrm[hit].shape), atol=1e-9)
    jac, _, _ = f.render_grad(np.eye(4), 60, 80, INTR_V, dep)
    rx, ry = _ray_dirs(60, 80, INTR_V)
    x, y = rx * dep, ry * dep
    np.testing.assert_allclose(jac[hit][:, :3], np.broadcast_to([0.0, 0.0, -1.0], (hit.sum(), 3)), atol=1e-7)
    np.testing.assert_allclose(jac[hit][:, 3], -y[hit], atol=1e-7)
    np.testing.assert_allclose(jac[hit][:, 4], x[hit], atol=1e-7)
    np.testing.assert_allclose(jac[hit][:, 5], 0.0, atol=1e-7)


def _cells(T, dep, vl=0.01):
    """Trilinear cell of every hit point (the field is smooth inside one)."""
    rx, ry = _ray_dirs(dep.shape[0], dep.shape[1], INTR_V)
    pc = np.stack([rx * dep, ry * dep, dep], -1)
    pw = pc @ T[:3, :3].T + T[:3, 3]
    return np.floor(pw / vl - 0.5).astype(np.int64)


def _fd_pose(f, T0, h=1e-4):
    """Analytic dD/dxi, its central difference at h and the Richardson
    extrapolation of the central differences at h and h / 2 (which cancels
    their h^2 error term).  `ok` marks pixels that hit in every render and
    whose perturbed hit points stay in the trilinear cell of the unperturbed
    one for every parameter: there the depth is a smooth function of the
    pose and the difference quotients are meaningful."""
    dep0, _ = f.render(T0, 60, 80, INTR_V)
    jac, _, _ = f.render_grad(T0, 60, 80, INTR_V, dep0)
    fd = np.zeros_like(jac)
    rich = np.zeros_like(jac)
    ok = dep0 > 0
    c0 = _cells(T0, dep0)
    for k in range(6):
        q = []
        for hh in (h, h / 2):
            e = np.zeros(6)
            e[k] = hh
            Tp, Tm = se3_exp(e) @ T0, se3_exp(-e) @ T0
            dp, _ = f.render(Tp, 60, 80, INTR_V)
            dm, _ = f.render(Tm, 60, 80, INTR_V)
            q.append((dp - dm) / (2 * hh))
            ok &= (dp > 0) & (dm > 0)
            ok &= (_cells(Tp, dp) == c0).all(-1) & (_cells(Tm, dm) == c0).all(-1)
        fd[..., k] = q[0]
        rich[..., k] = (4 * q[1] - q[0]) / 3
    return dep0, jac, fd, rich, ok


def _view_pose(poses):
    T0 = synth.perturb(poses[2], np.random.default_rng(5), sigma_t=0.01, sigma_deg=1.0)
    return T0


def test_pose_gradient_matches_fd():
    f, poses, _ = _scene_field()
    dep0, jac, fd, rich, ok = _fd_pose(f, _view_pose(poses))
    # most hits are inside one cell for the whole stencil
    assert ok.sum() > 0.6 * (dep0 > 0).sum() and ok.sum() > 1500
    # |dD/dv| = 1 / cos(ray, normal): rays within ~70 deg of the normal
    well = ok & (np.linalg.norm(jac[..., :3], axis=-1) < 3.0)
    assert well.sum() > 0.5 * ok.sum()
    # h = 1e-4 central differences (errors relative to the norm of the
    # pixel's Jacobian row): their O(h^2) truncation grows like 1 / cos^3 and
    # reaches a few per cent on far grazing floor pixels
    scale = np.linalg.norm(jac, axis=-1, keepdims=True)
    rel = (np.abs(fd - jac) / (1e-6 + scale)).max(-1)
    assert np.median(rel[ok]) < 1e-5
    assert np.percentile(rel[well], 99) < 1e-3 and rel[well].max() < 1e-2
    # with the h^2 term cancelled the agreement is 1e-4 on well-conditioned
    # pixels and 1e-3 on all of them
    rel_r = np.abs(rich - jac) / (1e-6 + scale)
    assert rel_r[well].max() < 1e-4
    assert rel_r[ok].max() < 1e-3
    s_fd, s_an = rich[ok].sum(0), jac[ok].sum(0)
    assert np.all(np.abs(s_fd - s_an) <= 1e-4 * np.abs(s_an).max())


@pytest.mark.parametrize("mut", [7])
def test_pose_gradient_mutation_is_caught(mut):
    f, poses, _ = _scene_field()
    oracle.set_mutation(mut)
    try:
        dep0, jac, fd, rich, ok = _fd_pose(f, _view_pose(poses))
    finally:
        oracle.set_mutation(0)
    rel = np.abs(rich - jac) / (1e-6 + np.linalg.norm(jac, axis=-1, keepdims=True))
    assert (rel[ok].max(axis=-1) < 1e-3).mean() < 0.5


def _voxel_fd(f, ex, T0, mut=0, n_vox=12, h=1e-4):
    """For the n_vox voxels with the largest gradient of
    L = sum over well-conditioned hits of g * D (g random): rows of (central
    difference at h, Richardson extrapolation from h and h / 2, analytic
    dL/dV), all over the pixels whose hit stays in the same trilinear cell
    when the voxel moves by up to h."""
    dep0, _ = f.render(T0, 60, 80, INTR_V)
    jac, _, _ = f.render_grad(T0, 60, 80, INTR_V, dep0)
    well = (dep0 > 0) & (np.linalg.norm(jac[..., :3], axis=-1) < 3.0)
    rng = np.random.default_rng(9)
    gd = rng.normal(size=dep0.shape) * well
    _, _, gs = f.render_grad(T0, 60, 80, INTR_V, dep0, gd, want_sdf=True)
    flat = np.argsort(-np.abs(gs).reshape(-1))[:n_vox]
    sdf = ex["sdf"].copy()
    c0 = _cells(T0, dep0)
    out = []
    for fi in flat:
        b, v = divmod(int(fi), 512)
        deps, effs = [], []
        for hh in (h, h / 2):
            for sgn in (1, -1):
                s2 = sdf.copy()
                s2[b, v] = np.float32(sdf[b, v] + sgn * hh)
                f.set_sdf(s2)
                deps.append(f.render(T0, 60, 80, INTR_V)[0])
            effs.append(float(np.float32(sdf[b, v] + hh)) - float(np.float32(sdf[b, v] - hh)))
        f.set_sdf(sdf)
        same = well.copy()
        for d in deps:
            same &= (d > 0) & (_cells(T0, d) == c0).all(-1)
        if not same.any():
            continue
        g_same = gd * same
        q = [((deps[2 * k] - deps[2 * k + 1]) * g_same).sum() / effs[k] for k in range(2)]
        oracle.set_mutation(mut)
        try:
            _, _, gsv = f.render_grad(T0, 60, 80, INTR_V, dep0, g_same, want_sdf=True)
        finally:
            oracle.set_mutation(0)
        out.append((q[0], (4 * q[1] - q[0]) / 3, gsv[b, v]))
    return np.array(out)


def test_voxel_gradient_matches_fd():
    f, poses, ex = _scene_field()
    r = _voxel_fd(f, ex, _view_pose(poses))
    r = r[np.abs(r[:, 2]) > 1e-2]  # voxels whose affected hits all change cell drop out
    assert len(r) >= 8
    fd, rich, an = r[:, 0], r[:, 1], r[:, 2]
    np.testing.assert_allclose(fd, an, rtol=2e-3)      # h = 1e-4 central difference
    np.testing.assert_allclose(rich, an, rtol=1e-4)    # h^2 term cancelled


def test_voxel_gradient_mutation_is_caught():
    f, poses, ex = _scene_field()
    r = _voxel_fd(f, ex, _view_pose(poses), mut=8)
    assert not np.allclose(r[:, 1], r[:, 2], rtol=1e-2)


def test_se3_left_jacobian():
    rng = np.random.default_rng(3)
    for _ in range(5):
        xi = rng.normal(size=6) * np.array([0.3, 0.3, 0.3, 0.6, 0.6, 0.6])
        J = se3_left_jacobian(xi)
        for k in range(6):
            e = np.zeros(6)
            e[k] = 1e-6
            lhs = se3_exp(xi + e)
            rhs = se3_exp(J @ e) @ se3_exp(xi)
            assert np.abs(lhs - rhs).max() < 1e-10


def test_gauss_newton_step_c_abi():
    from superscanner_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(4)
    A = rng.normal(size=(40, 6))
    H = A.T @ A
    b = rng.normal(size=6)
    T = se3_exp(rng.normal(size=6) * 0.2)
    tri = np.array([H[p, q] for p in range(6) for q in range(p, 6)])
    out = np.zeros(16)
    d = np.zeros(6)
    dp = C.POINTER(C.c_double)
    st = L.tsdf_se3_gauss_newton_step(np.ascontiguousarray(tri).ctypes.data_as(dp),
                                      np.ascontiguousarray(b).ctypes.data_as(dp),
                                      np.ascontiguousarray(T.reshape(16)).ctypes.data_as(dp),
                                      out.ctypes.data_as(dp), d.ctypes.data_as(dp))
    assert st == 0
    delta = -np.linalg.solve(H, b)
    np.testing.assert_allclose(d, delta, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(out.reshape(4, 4), se3_exp(delta) @ T, rtol=1e-10, atol=1e-12)
    # singular system -> BAD_ARG, output untouched
    z = np.zeros(21)
    assert L.tsdf_se3_gauss_newton_step(z.ctypes.data_as(dp), np.ascontiguousarray(b).ctypes.data_as(dp),
                                        np.ascontiguousarray(T.reshape(16)).ctypes.data_as(dp),
                                        out.ctypes.data_as(dp), d.ctypes.data_as(dp)) == _lib.TSDF_BAD_ARG
