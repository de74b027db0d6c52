"""Launch-shape autotuner for the TSDF kernels (no reference counterpart: the
reference launches fixed Kokkos MDRange policies, GPUVoxelGrid.cpp:59-143).

The tunables are the fields of ``tsdf_tuning`` (include/tsdf.h):

  integrate_wg_per_cu   resident brick-update workgroups per CU (occupancy
                        of the persistent k_integrate / k_integrate_batch)
  batch_ops_per_pass    ops per collection + update pass of a batch (the
                        pass's point maps, ops * H * W * 16 B, should stay in
                        the last-level cache; more ops per pass means fewer
                        payload round trips)
  batch_graph           a batch of <= 256 ops as one captured graph
  raycast_xcd_map       XCD-banded tile order of the raycast
  raycast_skip          exact empty-space skip of the raycast (0 none, 1 distance
                        field over superblocks, 2 depth bounds of image tiles)
  raycast_budget        march rounds of a ray before a second kernel carries it
                        with 16 lanes (0 never)
  raycast_order         dispatch order of the raycast's tiles (slowest of the
                        previous raycast first, reversed, or the hardware's)

Each tuner times candidates on the caller's own representative workload with
HIP events on the grid's stream and returns the fastest setting together with
the measured table.  A batch changes the map, so every reintegration
candidate runs on a grid rebuilt by the caller's ``make_grid``.
"""
from __future__ import annotations

import itertools

import torch

INTEGRATE_WG_PER_CU = (2, 3, 4, 6, 8)
BATCH_OPS_PER_PASS = (4, 8, 16, 32, 64, 128)


def _timed(fn, stream=None):
    s = stream or torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(s)
    fn()
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b)


def tune_reintegration(make_grid, depths, old_poses, new_poses, *, ops_per_pass=BATCH_OPS_PER_PASS,
                       wg_per_cu=INTEGRATE_WG_PER_CU, repeats=1):
    """Picks (batch_ops_per_pass, integrate_wg_per_cu) for
    ``GPUVoxelGrid.reintegrate_batch(depths, old_poses, new_poses)``.

    make_grid() must return a grid in the state the batch starts from.
    Returns (best tuning dict, [(ops_per_pass, wg_per_cu, ms), ...])."""
    n_ops = 2 * len(depths)
    table = []
    for per, wg in itertools.product(ops_per_pass, wg_per_cu):
        if per > max(n_ops, 1) and per != min(ops_per_pass):
            continue
        best = float("inf")
        for _ in range(repeats):
            g = make_grid()
            try:
                g.set_tuning(batch_ops_per_pass=per, integrate_wg_per_cu=wg)
                g.reserve_batch(n_ops)
                torch.cuda.synchronize()
                best = min(best, _timed(lambda: g.reintegrate_batch(depths, old_poses, new_poses)))
            finally:
                g.close()
        table.append((per, wg, best))
    per, wg, _ = min(table, key=lambda r: r[2])
    return {"batch_ops_per_pass": per, "integrate_wg_per_cu": wg}, table


def tune_integrate(grid, depths, poses, *, wg_per_cu=INTEGRATE_WG_PER_CU, rounds=3):
    """Picks integrate_wg_per_cu for plain integrate calls on ``grid``.  Each
    candidate integrates every frame and deintegrates it again, which leaves
    the weights as they were (the sdf values may differ in the last bits).
    The candidates are timed in `rounds` interleaved rounds and each keeps
    its best time, so one disturbed sample does not decide.
    Returns (best tuning dict, [(wg_per_cu, ms), ...])."""
    best = {wg: float("inf") for wg in wg_per_cu}
    for _ in range(rounds):
        for wg in wg_per_cu:
            grid.set_tuning(integrate_wg_per_cu=wg)

            def run():
                for d, E in zip(depths, poses):
                    grid.integrate(None, d, E)
                for d, E in zip(depths, poses):
                    grid.deIntegrate(None, d, E)

            run()  # warm
            best[wg] = min(best[wg], _timed(run))
    table = [(wg, best[wg]) for wg in wg_per_cu]
    wg = min(table, key=lambda r: r[1])[0]
    grid.set_tuning(integrate_wg_per_cu=wg)
    return {"integrate_wg_per_cu": wg}, table


def tune_raycast(grid, poses, *, xcd_map=(0,), skip=(0, 1, 2), order=(0, 2), budget=(0, 12, 16), rounds=3):
    """Picks raycast_xcd_map, raycast_skip, raycast_order and raycast_budget
    for ``grid.raycast`` (the group march replays no superblock field, so
    skip 1 runs without a budget).  The candidates are timed in `rounds`
    interleaved rounds and each keeps its best time.  Returns (best tuning
    dict, [(xcd_map, skip, order, budget, ms), ...])."""
    cands = [c for c in itertools.product(xcd_map, skip, order, budget) if not (c[1] == 1 and c[3])]
    best = {c: float("inf") for c in cands}

    def run():
        for E in poses:
            grid.raycast(E)

    for _ in range(rounds):
        for xm, sk, od, bu in cands:
            grid.set_tuning(raycast_xcd_map=xm, raycast_skip=sk, raycast_order=od, raycast_budget=bu)
            run()  # warm (and, for order 2, the costs the first timed raycast orders by)
            best[(xm, sk, od, bu)] = min(best[(xm, sk, od, bu)], _timed(run))
    table = [(*c, best[c]) for c in cands]
    xm, sk, od, bu, _ = min(table, key=lambda r: r[4])
    grid.set_tuning(raycast_xcd_map=xm, raycast_skip=sk, raycast_order=od, raycast_budget=bu)
    return {"raycast_xcd_map": xm, "raycast_skip": sk, "raycast_order": od, "raycast_budget": bu}, table
