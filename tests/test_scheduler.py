"""The reintegration scheduler's host side (include/tsdf.h scheduler helpers
and superscanner_amd/integration.py) against closed forms and the numpy
oracle (oracle/sched.py), on the CPU:

  * MattoDof / getT: single-axis rotations and a translation have known
    dofs; getT(MattoDof(T)) = T and MattoDof(getT(d)) = d to rounding;
  * diff: rotation angles count once, translations twice;
  * selection: largest diff first, ties in list order, cut at the first diff
    below treint (a diff equal to treint is taken);
  * the scheduler's buffers, driven through a grid that records its calls:
    the order and poses of the reference's thread loop (integrate.cpp:60-141),
    removeFrameFromIntegration, the batched drain (one reintegrate_batch with
    the same (old, new) pairs the plain loop would issue), and the stop path
    (which integrates the frames waiting for deintegration, integrate.cpp:196-207).
"""
import numpy as np
import pytest

from oracle import sched as osched
from superscanner_amd import integration as I
from superscanner_amd import synth


def rot(axis, a):
    T = np.eye(4)
    T[:3, :3] = synth.axis_angle(axis, a)
    return T


@pytest.mark.parametrize("k", [0, 1, 2])
def test_single_axis_dofs(k):
    axis = np.eye(3)[k]
    d = I.mat_to_dof(rot(axis, 0.3))
    exp = np.zeros(6)
    exp[k] = 0.3
    assert np.abs(d - exp).max() < 1e-15
    T = np.eye(4)
    T[:3, 3] = (1.5, -2.0, 0.25)
    assert np.array_equal(I.mat_to_dof(T), [0, 0, 0, 1.5, -2.0, 0.25])


def test_round_trips():
    rng = np.random.default_rng(0)
    for _ in range(200):
        d = np.concatenate([rng.uniform(-3, 3, 1), rng.uniform(-1.5, 1.5, 1), rng.uniform(-3, 3, 1),
                            rng.uniform(-5, 5, 3)])
        T = I.dof_to_mat(d)
        assert np.abs(T[:3, :3] @ T[:3, :3].T - np.eye(3)).max() < 1e-15
        assert np.abs(I.mat_to_dof(T) - d).max() < 1e-13
        assert np.abs(I.dof_to_mat(I.mat_to_dof(T)) - T).max() < 1e-15
        # the oracle's restatement agrees to rounding
        assert np.abs(osched.dof_to_mat(d) - T).max() < 1e-15
        assert np.abs(osched.mat_to_dof(T) - I.mat_to_dof(T)).max() < 1e-15


def test_diff_weights():
    z = np.zeros(6)
    assert I.reint_diff([0.03, 0, 0, 0, 0, 0], z) == pytest.approx(0.03, abs=1e-17)
    assert I.reint_diff([0, 0, 0, 0, 0.01, 0], z) == pytest.approx(0.02, abs=1e-17)
    assert I.reint_diff([0, 0.03, 0, 0.02, 0, 0], z) == pytest.approx(0.05, abs=1e-16)
    assert I.reint_diff([0.1, 0.2, 0.3, 0.4, 0.5, 0.6], [0.1, 0.2, 0.3, 0.4, 0.5, 0.6]) == 0.0


def test_selection_order_ties_and_cut():
    w = np.zeros((6, 6))
    w[0, 3] = 0.01     # diff 0.02
    w[1, 0] = 0.05     # 0.05
    w[2, 4] = 0.025    # 0.05, tie with 1, later in the list
    w[3, 5] = 0.02     # 0.04 == treint: taken
    w[4, 1] = 0.0399   # below
    w[5, 2] = 0.2      # 0.2
    order, m = I.reint_select(w, np.zeros((6, 6)), 0.04)
    assert order[:4] == [5, 1, 2, 3] and m == 4
    assert (order, m) == osched.select(w, np.zeros((6, 6)), 0.04)
    rng = np.random.default_rng(5)
    for _ in range(50):
        n = int(rng.integers(0, 40))
        a = rng.normal(0, 0.03, (n, 6)).round(3)  # rounding makes ties
        b = np.zeros((n, 6))
        assert I.reint_select(a, b, 0.04) == osched.select(a, b, 0.04)


class RecordingGrid:
    """Stands in for GPUVoxelGrid: records (sign, frame id, float pose)."""
    withColor = False

    def __init__(self):
        self.calls = []

    def integrate(self, rgb, depth, extr, withColor=True):
        self.calls.append((+1, int(depth), np.asarray(extr, np.float32).copy()))

    def deIntegrate(self, rgb, depth, extr, withColor=True):
        self.calls.append((-1, int(depth), np.asarray(extr, np.float32).copy()))

    def reintegrate_batch(self, depths, old, new, rgbs=None):
        for d, o, n in zip(depths, old, new):
            self.calls.append((-1, int(d), np.asarray(o, np.float32).copy()))
            self.calls.append((+1, int(d), np.asarray(n, np.float32).copy()))


def make_frames(n, seed=1):
    rng = np.random.default_rng(seed)
    fr = []
    for k in range(n):
        f = I.Frame(depth=k, unique_id=k)
        T = synth.look_at(rng.uniform(-2, 2, 3) + (0, 0, 3), (0, 0, 0))
        f.setFrametoWorldTrans(T)
        fr.append(f)
    return fr


def same(a, b):
    return len(a) == len(b) and all(x[0] == y[0] and x[1] == y[1] and np.array_equal(x[2], y[2])
                                    for x, y in zip(a, b))


def moved(f, rng, scale):
    M = rot(rng.normal(size=3), rng.uniform(0, scale))
    M[:3, 3] = rng.normal(0, scale, 3)
    return f.frametoworldtrans @ M


def test_thread_loop_order_and_poses():
    g = RecordingGrid()
    s = I.IntegrationScheduler(g)
    fr = make_frames(5)
    for f in fr:
        s.push(f)
    while s.step():
        pass
    assert [c[:2] for c in g.calls] == [(1, k) for k in range(5)]
    assert all(np.array_equal(c[2], fr[k].frametoworldtrans.astype(np.float32)) for k, c in enumerate(g.calls))
    # frames 1 and 3 move far, 4 a little: 3 then 1 are reintegrated
    rng = np.random.default_rng(2)
    old = {k: fr[k].integrateddofs.copy() for k in range(5)}
    fr[1].setFrametoWorldTrans(moved(fr[1], rng, 0.05))
    fr[3].setFrametoWorldTrans(moved(fr[3], rng, 0.2))
    fr[4].setFrametoWorldTrans(moved(fr[4], rng, 0.001))
    d = {k: osched.diff(fr[k].worlddofs, fr[k].integrateddofs) for k in (1, 3, 4)}
    assert d[3] > d[1] > 0.04 > d[4]
    s.updateReintegrationBuffer()
    assert [f.unique_id for f in s.reintegrationBuffer] == [3, 1]
    assert [f.unique_id for f in s.integratedframes][:2] == [3, 1]
    # a new frame arrives and frame 0 is removed: one loop pass does all three
    new = make_frames(6, seed=9)[5]
    new.unique_id = new.depth = 5
    s.push(new)
    s.removeFrameFromIntegration(fr[0])
    g.calls.clear()
    s.step()
    exp = [(1, 5, new.frametoworldtrans.astype(np.float32)),
           (-1, 3, osched.dof_to_mat(old[3]).astype(np.float32)),
           (1, 3, fr[3].frametoworldtrans.astype(np.float32)),
           (-1, 0, osched.dof_to_mat(old[0]).astype(np.float32))]
    assert same(g.calls, exp)
    assert np.array_equal(fr[3].integrateddofs, fr[3].worlddofs) and not fr[0].integrateddofs.any()
    assert all(f is not fr[0] for f in s.integratedframes)
    s.step()
    assert same(g.calls[4:], [(-1, 1, osched.dof_to_mat(old[1]).astype(np.float32)),
                              (1, 1, fr[1].frametoworldtrans.astype(np.float32))])
    assert not s.step()


def test_remove_waiting_frame_never_reaches_the_grid():
    g = RecordingGrid()
    s = I.IntegrationScheduler(g)
    fr = make_frames(3)
    for f in fr:
        s.push(f)
    s.removeFrameFromIntegration(fr[1])
    while s.step():
        pass
    assert [c[1] for c in g.calls] == [0, 2]


def test_batched_drain_issues_the_loop_pairs():
    runs = []
    for batched in (False, True):
        g = RecordingGrid()
        s = I.IntegrationScheduler(g)
        fr = make_frames(12, seed=4)
        for f in fr:
            s.push(f)
        s.run_batched() if batched else [None for _ in iter(s.step, False)]
        r = np.random.default_rng(7)
        for f in fr:
            f.setFrametoWorldTrans(moved(f, r, 0.06))
        s.updateReintegrationBuffer()
        n = len(s.reintegrationBuffer)
        assert n >= 3
        g.calls.clear()
        if batched:
            out = s.run_batched()
            assert out["reintegrated"] == n
        else:
            while s.step():
                pass
        runs.append(g.calls)
    assert same(runs[0], runs[1])


def test_stop_path_integrates_waiting_deintegrations():
    g = RecordingGrid()
    s = I.IntegrationScheduler(g)
    fr = make_frames(3)
    for f in fr[:2]:
        s.push(f)
    while s.step():
        pass
    s.push(fr[2])
    s.removeFrameFromIntegration(fr[0])
    g.calls.clear()
    s.finish()
    assert [c[:2] for c in g.calls] == [(1, 2), (1, 0)]  # frame 0 is integrated again, as written
    assert np.array_equal(g.calls[1][2], osched.dof_to_mat(fr[0].worlddofs).astype(np.float32))
