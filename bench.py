#!/usr/bin/env python3
"""TSDF integrate + raycast throughput on one MI355X.

Workload ("room_large"): 600 synthetic 848x480 uint16 depth frames of a seeded
random room (floor/walls, 6 boxes, 2 spheres) along a random walk of 5 mm /
0.5 deg per step; voxel 4 mm, truncation 20 mm (5 voxels), 2^22 hash buckets,
max weight 128, fp32 sdf + fp32 weight.  One step is one pass over the whole
sequence from an empty map: for pose i, integrate frame i; when i >= 10 and
i % 10 == 0 deintegrate frame i - 10 at its pose; raycast pose i (world points
+ depth).  value = 600 * 848 * 480 / step time, in Mpx/s.  Depth frames are
resident in HBM before the timed region starts.

Besides the headline line the JSON carries:
  roofline      the dominant kernel's algorithmic bytes per launch / its
                average duration (HIP events around every launch on its
                stream, over a second timed pass), against 8 TB/s
  kernels       the same figures for every kernel of the path
  cpu_baseline  the oracle (C restatement of the reference) on the first
                frames of the same sequence, on the host cores; and on the
                whole CPU-runnable plain workload (plain_small)
  plain_small   the CPU-runnable plain workload (30 frames of 640x480, the
                analytic scene): integrate all, raycast one view
  batched_reintegration
                64 keyframes re-fused at perturbed poses (valid fraction
                U(0.3, 1) per keyframe): one batched pass vs 128 plain calls,
                with the resulting TSDFs compared for identity
  maintenance   mesh extraction and cleanMemory on the map of the last step
  icp           projective association (both types) on two 640x480 frames,
                HIP-event timed, and a whole kokkosICP call
  ingest        PNG sequence -> pinned host -> HBM frame rate, and the pinned
                host -> device copy rate (PCIe-inclusive, not the headline)
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
S_V = 8        # bytes per voxel: fp32 sdf + fp32 weight
S_H = 8        # bytes per touched hash slot
S_OUT = 16     # raycast output per pixel: world point (12) + depth (4)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=600)
    ap.add_argument("--cpu-frames", type=int, default=60,
                    help="frames of the sequence the CPU baseline processes")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-reintegration", action="store_true")
    ap.add_argument("--keyframes", type=int, default=64)
    ap.add_argument("--no-diff-render", action="store_true")
    ap.add_argument("--no-maintenance", action="store_true")
    ap.add_argument("--no-icp", action="store_true")
    ap.add_argument("--no-ingest", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--fp16-frames", type=int, default=2000)
    ap.add_argument("--footprint-samples", type=int, default=12)
    ap.add_argument("--tuning", type=str, default=None, metavar="JSON",
                    help="launch shapes to use instead of the autotuner's (a tsdf_tuning dict, e.g. for a "
                         "profile of the steps alone); the raycast order is not re-settled")
    ap.add_argument("--pmc-workload", action="store_true",
                    help="run 2 steps of the workload plus a 256 MiB calibration copy and exit "
                         "(for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--pmc-batch", action="store_true",
                    help="run one batched reintegration of 64 keyframes and the same 128 plain calls, and exit "
                         "(for rocprofv3 --pmc)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="turn two rocprofv3 counter_collection.csv files of --pmc-workload runs into "
                         "profiles/pmc_traffic.json and profiles/pmc_traffic_room_large.txt")
    ap.add_argument("--summarize-sq", nargs="+", metavar="DB",
                    help="print per-kernel medians of the counters in rocprofv3 --pmc result databases "
                         "(run_results.db of --pmc-workload runs) as JSON")
    return ap.parse_args()


def workload(frames):
    from superscanner_amd import synth
    cfg = dict(synth.ROOM_LARGE)
    poses64 = synth.traj_random_walk(n=frames, seed=2)
    prims = synth.scene_room(seed=1)
    t0 = time.time()
    depth = synth.frames(prims, poses64, cfg, seed=7)
    gen_s = time.time() - t0
    poses = np.stack([synth.pose_f32(T) for T in poses64])
    return cfg, prims, poses64, poses, depth, gen_s


def schedule(frames):
    """(op, frame) list of one step: ('i', k) integrate, ('d', k) deintegrate,
    ('r', k) raycast."""
    ops = []
    for i in range(frames):
        ops.append(("i", i))
        if i >= 10 and i % 10 == 0:
            ops.append(("d", i - 10))
        ops.append(("r", i))
    return ops


class Runner:
    """Issues one step straight through the C ABI with prebuilt arguments."""

    def __init__(self, grid, depth_dev, poses, frames):
        import torch
        self.g = grid
        self.L = grid._L
        self.h = grid._h
        self.frames = frames
        H, W = grid.resi, grid.resj
        self.pitch = W
        self.base = depth_dev.data_ptr()
        self.fsz = H * W * 2
        self.E = np.ascontiguousarray(poses.reshape(-1, 16).astype(np.float32))
        self.Ep = [self.E[k].ctypes.data_as(C.POINTER(C.c_float)) for k in range(len(self.E))]
        self.pts = torch.empty((H, W, 3), dtype=torch.float32, device=depth_dev.device)
        self.dep = torch.empty((H, W), dtype=torch.float32, device=depth_dev.device)
        self.stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        self.ops = schedule(frames)

    def step(self):
        L, h, s = self.L, self.h, self.stream
        rc = L.tsdf_reset(h, s)
        pts, dep = C.c_void_p(self.pts.data_ptr()), C.c_void_p(self.dep.data_ptr())
        for op, k in self.ops:
            if op == "r":
                rc |= L.tsdf_raycast(h, self.Ep[k], pts, None, dep, None, s)
            else:
                rc |= L.tsdf_integrate(h, C.c_void_p(self.base + k * self.fsz), self.pitch, None, self.Ep[k],
                                       1 if op == "i" else -1, s)
        if rc:
            raise RuntimeError("a TSDF call failed: " + L.tsdf_last_error().decode())


class HostRunner(Runner):
    """The same step with the depth frames starting in pinned host memory
    (the reference's Frame::MovetoGPU before each integrate, Frame.cpp:44-60):
    every integrate / deintegrate uploads its frame inside the timed region.
    overlap=False: upload, then integrate, in stream order.  overlap=True:
    a second stream uploads up to `ring` - 1 frames ahead into a ring of
    device buffers, events ordering each upload after the integrate that
    last read its buffer.  d2h=True also copies every raycast's points and
    depth back to pinned host memory after the raycast."""

    def __init__(self, grid, depth_host, poses, frames, ring=4):
        import torch
        super().__init__(grid, torch.empty(1, dtype=torch.int16, device=grid.device), poses, frames)
        H, W = grid.resi, grid.resj
        self.host = depth_host  # [frames, H, W] int16, pinned
        self.ring = ring
        self.slots = [torch.empty((H, W), dtype=torch.int16, device=grid.device) for _ in range(ring)]
        self.pts_h = torch.empty((H, W, 3), dtype=torch.float32, pin_memory=True)
        self.dep_h = torch.empty((H, W), dtype=torch.float32, pin_memory=True)
        self.copy = torch.cuda.Stream(device=grid.device)
        self.ready = [torch.cuda.Event() for _ in range(ring)]
        self.free = [torch.cuda.Event() for _ in range(ring)]

    def step_host(self, overlap, d2h=False):
        import torch
        L, h, s = self.L, self.h, self.stream
        comp = torch.cuda.current_stream()
        pts, dep = C.c_void_p(self.pts.data_ptr()), C.c_void_p(self.dep.data_ptr())
        uploads = [k for op, k in self.ops if op != "r"]
        R, D = self.ring, self.ring - 1

        def upload(n):
            slot = n % R
            with torch.cuda.stream(self.copy):
                self.copy.wait_event(self.free[slot])
                self.slots[slot].copy_(self.host[uploads[n]], non_blocking=True)
                self.ready[slot].record(self.copy)

        rc = L.tsdf_reset(h, s)
        if overlap:
            for slot in range(R):
                self.free[slot].record(comp)
            for n in range(min(D, len(uploads))):
                upload(n)
        n = 0
        for op, k in self.ops:
            if op == "r":
                rc |= L.tsdf_raycast(h, self.Ep[k], pts, None, dep, None, s)
                if d2h:
                    self.pts_h.copy_(self.pts, non_blocking=True)
                    self.dep_h.copy_(self.dep, non_blocking=True)
                continue
            slot = n % R
            if overlap:
                comp.wait_event(self.ready[slot])
            else:
                self.slots[slot].copy_(self.host[k], non_blocking=True)
            rc |= L.tsdf_integrate(h, C.c_void_p(self.slots[slot].data_ptr()), self.pitch, None, self.Ep[k],
                                   1 if op == "i" else -1, s)
            if overlap:
                self.free[slot].record(comp)
                if n + D < len(uploads):
                    upload(n + D)
            n += 1
        if rc:
            raise RuntimeError("a TSDF call failed: " + L.tsdf_last_error().decode())


def host_inclusive(grid, depth, poses, frames, ms_resident, steps=2):
    """The headline with the frames' host -> device copies inside the timed
    region (and, in one variant, the raycast outputs copied back)."""
    import torch
    host = torch.from_numpy(depth.view(np.int16)).pin_memory()
    run = HostRunner(grid, host, poses, frames)
    px = grid.resi * grid.resj
    out = {"resident_ms_per_step": round(ms_resident, 3)}
    for name, overlap, d2h in (("h2d_serialized", False, False), ("h2d_overlapped", True, False),
                               ("h2d_d2h_serialized", False, True)):
        run.step_host(overlap, d2h)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            run.step_host(overlap, d2h)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        out[name] = {"ms_per_step": round(ms, 3), "Mpx_per_s": round(frames * px / (ms * 1e-3) / 1e6, 2)}
    out["note"] = ("h2d_*: every integrate / deintegrate uploads its uint16 frame from pinned host memory in the "
                   "timed region (serialized: in stream order; overlapped: a copy stream runs 3 frames ahead); "
                   "h2d_d2h_serialized also copies each raycast's points + depth (16 B/px) back to pinned host "
                   "memory; wall clock over %d steps" % steps)
    return out


def device_copy(dev, nbytes=1 << 30, reps=10):
    """Measured device-to-device copy rate (bytes read + written per second):
    hipMemcpy D2D through torch and a plain elementwise copy kernel."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    out = {"bytes": nbytes}
    for name, fn in (("memcpy_d2d", lambda: b.copy_(a)), ("copy_kernel", lambda: torch.mul(a, 1.0, out=b))):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        out[name + "_GBs"] = round(2 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9, 1)
    out["device_copy_GBs"] = max(out["memcpy_d2d_GBs"], out["copy_kernel_GBs"])
    out["peak_GBs"] = HBM_PEAK_GBS
    out["copy_frac_of_peak"] = round(out["device_copy_GBs"] / HBM_PEAK_GBS, 3)
    del a, b
    return out


def autotune_plain(cfg, depth_dev, poses, frames):
    """Launch shapes of the plain integrate and raycast, picked by the
    autotuner on a scratch grid holding the first part of the sequence."""
    from superscanner_amd import autotune
    from superscanner_amd.voxel_grid import GPUVoxelGrid
    g = GPUVoxelGrid(1 << 30, cfg["truncation"], cfg["voxel_length"], cfg["H"], cfg["W"],
                     n_buckets=cfg["n_buckets"], n_bricks=1 << 21, max_weight=cfg["max_weight"])
    g.setIntrinsic(*cfg["intr"])
    n = min(frames, 80)
    for k in range(n):
        g.integrate(None, depth_dev[k], poses[k])
    lo = max(0, n - 20)
    best_i, _ = autotune.tune_integrate(g, [depth_dev[k] for k in range(lo, n)], [poses[k] for k in range(lo, n)])
    best_r, _ = autotune.tune_raycast(g, [poses[k] for k in range(lo, n)])
    g.close()
    return {**best_i, **best_r}


def footprint_pass(runner, n_samples):
    """Re-runs the sequence untimed and measures the raycast footprint
    (distinct voxels / keys) at n_samples evenly spaced poses."""
    g = runner.g
    L, h, s = runner.L, runner.h, runner.stream
    want = set(np.linspace(0, runner.frames - 1, n_samples).astype(int).tolist())
    L.tsdf_reset(h, s)
    u_vox, u_blk = [], []
    for op, k in runner.ops:
        if op == "r":
            if k in want:
                fp = g.raycast_footprint(runner.E[k].reshape(4, 4))
                u_vox.append(fp["u_vox"])
                u_blk.append(fp["u_blk"])
        else:
            L.tsdf_integrate(h, C.c_void_p(runner.base + k * runner.fsz), runner.pitch, None, runner.Ep[k],
                             1 if op == "i" else -1, s)
    g.sync()
    return float(np.mean(u_vox)), float(np.mean(u_blk))


def host_cpu():
    """nproc, the cores this process may run on, and the CPU model."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model}


def cpu_baseline(cfg, poses, depth, n, runs=5):
    """The oracle over the first n frames of the sequence on the host cores
    this process is allotted (OMP_NUM_THREADS where set, else its CPU
    affinity): one warm-up pass, then the median of `runs` passes."""
    import statistics
    import oracle
    host = host_cpu()
    env = os.environ.get("OMP_NUM_THREADS")
    threads = int(env) if env and env.isdigit() and int(env) > 0 else host["affinity"]
    oracle.set_threads(threads)
    ops = schedule(n)

    def one():
        o = oracle.OracleGrid(cfg["voxel_length"], cfg["truncation"], cfg["H"], cfg["W"], max_bricks=1 << 22,
                              max_weight=cfg["max_weight"])
        o.set_intrinsics(*cfg["intr"])
        t0 = time.perf_counter()
        for op, k in ops:
            if op == "r":
                o.raycast(poses[k])
            else:
                o.integrate(depth[k], poses[k], 1 if op == "i" else -1)
        t = time.perf_counter() - t0
        o.close()
        return t

    one()
    times = [one() for _ in range(runs)]
    t = statistics.median(times)

    # the CPU-runnable plain workload, whole, as plain_small_leg runs it
    pcfg, pposes, pdepth = plain_small_workload()

    def plain():
        o = oracle.OracleGrid(pcfg["voxel_length"], pcfg["truncation"], pcfg["H"], pcfg["W"], max_bricks=1 << 19)
        o.set_intrinsics(*pcfg["intr"])
        t0 = time.perf_counter()
        for k in range(pcfg["frames"]):
            o.integrate(pdepth[k], pposes[k], 1)
        o.raycast(pposes[pcfg["frames"] - 1])
        dt = time.perf_counter() - t0
        o.close()
        return dt

    plain()
    ptimes = [plain() for _ in range(runs)]
    pt = statistics.median(ptimes)
    plain_out = {"ms": round(pt * 1e3, 1),
                 "Mpx_per_s": round(pcfg["frames"] * pcfg["H"] * pcfg["W"] / pt / 1e6, 3),
                 "runs_s": [round(x, 3) for x in ptimes]}
    return {"value": round(n * cfg["H"] * cfg["W"] / t / 1e6, 3), "unit": "Mpx/s", "cores": threads,
            "kind": "port",
            "sample": f"first {n} of the 600 poses of the same sequence (integrate, deintegrate every 10th, "
                      f"raycast each pose): median of {runs} passes after one warm-up, {t:.2f} s per pass, on "
                      f"{threads} OpenMP threads",
            "runs_s": [round(x, 3) for x in times], "host": host,
            "threads_from": "OMP_NUM_THREADS" if env else "sched_getaffinity",
            "plain_small": plain_out}


def plain_small_workload():
    """The CPU-runnable plain workload: 30 frames of 640x480 uint16 depth of
    the analytic scene (plane, sphere, box) along the circular trajectory,
    vl 5 mm, trunc 20 mm, 2^20 buckets, fp32 sdf + fp32 weight."""
    from superscanner_amd import synth
    cfg = dict(synth.PLAIN_SMALL)
    poses64 = synth.traj_plain_small(n=cfg["frames"])
    depth = synth.frames(synth.scene_plain_small(), poses64, cfg, seed=1)
    poses = np.stack([synth.pose_f32(T) for T in poses64])
    return cfg, poses, depth


def plain_small_leg(dev, reps=5):
    """The plain workload on the GPU: integrate its 30 frames, then raycast
    the last pose, from a reset grid; HIP events around the whole sequence,
    the median of `reps` after one warm-up.  Mpx/s counts the integrated
    pixels (30 x 640 x 480) over that time."""
    import statistics
    import torch
    from superscanner_amd.voxel_grid import GPUVoxelGrid
    cfg, poses, depth = plain_small_workload()
    n, H, W = cfg["frames"], cfg["H"], cfg["W"]
    dd = torch.from_numpy(depth.view(np.int16)).to(dev)
    g = GPUVoxelGrid(1 << 28, cfg["truncation"], cfg["voxel_length"], H, W, n_buckets=cfg["n_buckets"],
                     n_bricks=1 << 19)
    g.setIntrinsic(*cfg["intr"])
    run = Runner(g, dd, poses, n)
    run.ops = [("i", k) for k in range(n)] + [("r", n - 1)]
    run.step()
    times = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        run.step()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b))
    g.sync()
    bricks = g.stats()["n_bricks"]
    g.close()
    ms = statistics.median(times)
    return {"workload": f"{n} frames {W}x{H} of the analytic plain scene, vl 5 mm, trunc 20 mm, 2^20 buckets: "
                        "integrate every frame, then raycast the last pose",
            "ms": round(ms, 3), "Mpx_per_s": round(n * H * W / ms / 1e3, 1), "bricks": bricks,
            "runs_ms": [round(t, 3) for t in times]}


def reintegration(cfg, prims, poses64, keyframes, dev):
    """Batched vs plain re-fusion of `keyframes` keyframes."""
    import torch
    from superscanner_amd import autotune, synth
    from superscanner_amd.voxel_grid import GPUVoxelGrid
    rng = np.random.default_rng(21)
    n = len(poses64)
    kf = np.linspace(0, n - 1, keyframes).astype(int)
    keep = rng.uniform(0.3, 1.0, keyframes)
    kdepth = synth.frames(prims, [poses64[k] for k in kf], cfg, seed=31, keep=keep)
    old = np.stack([synth.pose_f32(poses64[k]) for k in kf])
    new = np.stack([synth.pose_f32(synth.perturb(poses64[k], rng)) for k in kf])
    kd = torch.from_numpy(kdepth.view(np.int16)).to(dev)
    H, W = cfg["H"], cfg["W"]

    def fresh():
        g = GPUVoxelGrid(1 << 30, cfg["truncation"], cfg["voxel_length"], H, W, n_buckets=cfg["n_buckets"],
                         n_bricks=1 << 18, max_weight=cfg["max_weight"])
        g.setIntrinsic(*cfg["intr"])
        for k in range(keyframes):
            g.integrate(None, kd[k], old[k])
        g.sync()
        return g

    kds = [kd[k] for k in range(keyframes)]
    # launch shapes of the batch picked by the autotuner on this workload
    tuned, table = autotune.tune_reintegration(fresh, kds, old, new)
    a, b = fresh(), fresh()
    b.set_tuning(**tuned)
    b.reserve_batch(2 * keyframes)  # set-up, like the pool itself: the timed batch allocates nothing
    torch.cuda.synchronize()
    # plain: the integrate.cpp:108-125 loop, one deIntegrate + integrate per keyframe
    t0 = time.perf_counter()
    for k in range(keyframes):
        a.deIntegrate(None, kd[k], old[k])
        a.integrate(None, kd[k], new[k])
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    # plain with the reference's fence after every call (GPUVoxelGrid.cpp:79,133)
    c = fresh()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(keyframes):
        c.deIntegrate(None, kd[k], old[k], fence=True)
        c.integrate(None, kd[k], new[k], fence=True)
    t_fenced = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.reintegrate_batch(kds, old, new)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    # the same batch once more on a fresh grid with per-kernel events
    p = fresh()
    p.set_tuning(**tuned)
    p.reserve_batch(2 * keyframes)
    p.profile(True)
    p.profile_read(reset=True)
    p.reintegrate_batch(kds, old, new)
    kt = p.profile_read(reset=True)
    p.close()
    ea, eb = a.export_bricks(), b.export_bricks()
    oa = torch.argsort(ea["coords"][:, 0] * 4194304 + ea["coords"][:, 1] * 2048 + ea["coords"][:, 2])
    ob = torch.argsort(eb["coords"][:, 0] * 4194304 + eb["coords"][:, 1] * 2048 + eb["coords"][:, 2])
    same = bool(torch.equal(ea["coords"][oa], eb["coords"][ob]) and
                torch.equal(ea["sdf"][oa].view(torch.int32), eb["sdf"][ob].view(torch.int32)) and
                torch.equal(ea["w"][oa].view(torch.int32), eb["w"][ob].view(torch.int32)))
    for g in (a, b, c):
        g.close()
    return {"keyframes": keyframes, "plain_calls": 2 * keyframes, "ms_plain": round(t_plain * 1e3, 3),
            "ms_plain_fenced": round(t_fenced * 1e3, 3), "ms_batched": round(t_batch * 1e3, 3),
            "speedup_vs_plain": round(t_plain / t_batch, 2), "speedup_vs_fenced": round(t_fenced / t_batch, 2),
            "tuning": tuned,
            "autotune_ms": {f"{per}x{wg}": round(ms, 3) for per, wg, ms in table},
            "kernel_ms": {"collect_batch": round(kt["collect_batch"]["ms"], 3),
                          "integrate_batch": round(kt["integrate_batch"]["ms"], 3)},
            "bit_identical": same}


def diff_render_leg(dev, n_views=16, vh=240, vw=320, iters=20):
    """Differentiable raycast over the plain_small TSDF (640x480 x 30 frames,
    vl 5 mm): 16 views at 320x240, forward depth, adjoint (pose + voxel
    gradients) and 20 Gauss-Newton tracking iterations per view."""
    import torch
    from superscanner_amd import diff_render as dr
    from superscanner_amd import synth
    from superscanner_amd.voxel_grid import GPUVoxelGrid
    cfg = dict(synth.PLAIN_SMALL)
    poses = synth.traj_plain_small(n=cfg["frames"])
    depth = synth.frames(synth.scene_plain_small(), poses, cfg, seed=1)
    g = GPUVoxelGrid(1 << 30, cfg["truncation"], cfg["voxel_length"], cfg["H"], cfg["W"],
                     n_buckets=cfg["n_buckets"], n_bricks=1 << 18)
    g.setIntrinsic(*cfg["intr"])
    dd = torch.from_numpy(depth.view(np.int16)).to(dev)
    for k in range(len(poses)):
        g.integrate(None, dd[k], synth.pose_f32(poses[k]))
    g.sync()
    vi = synth.scaled_intr(cfg, vh, vw)
    rng = np.random.default_rng(14)
    true = np.stack([synth.perturb(poses[(2 * k) % len(poses)], rng, 0.01, 1.0) for k in range(n_views)])
    init = np.stack([synth.perturb(t, rng, 0.01, 1.0) for t in true])
    sdf = g.sdf_tensor()

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / reps

    dep = dr.render(g, true, vh, vw, vi, sdf=sdf)
    gd = torch.randn_like(dep) * (dep > 0)
    ms_fwd = timed(lambda: dr.render(g, true, vh, vw, vi, sdf=sdf))
    ms_bwd = timed(lambda: dr.render_backward(g, true, vh, vw, vi, dep, gd, sdf=sdf))
    obs = dr.render(g, true, vh, vw, vi)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    T, hist = dr.track(g, obs, init, vi, iters=iters, max_residual=0.1)
    t_track = time.perf_counter() - t0
    err_t = [float(np.linalg.norm(a[:3, 3] - b[:3, 3])) for a, b in zip(T, true)]
    err0 = [float(np.linalg.norm(a[:3, 3] - b[:3, 3])) for a, b in zip(init, true)]
    g.close()
    rays = n_views * vh * vw
    return {"workload": f"{n_views} views {vw}x{vh} over the plain_small TSDF (640x480 x 30 frames, vl 5 mm)",
            "hit_fraction": round(float((dep > 0).float().mean()), 3),
            "ms_forward": round(ms_fwd, 3), "Mrays_per_s_forward": round(rays / ms_fwd / 1e3, 1),
            "ms_backward": round(ms_bwd, 3),
            "gauss_newton": {"iterations": iters, "ms_total": round(t_track * 1e3, 2),
                             "ms_per_iteration_all_views": round(t_track * 1e3 / iters, 3),
                             "median_translation_error_mm_init": round(1e3 * float(np.median(err0)), 3),
                             "median_translation_error_mm_final": round(1e3 * float(np.median(err_t)), 4)}}


def fp16_scale_leg(dev, frames):
    """fp16 storage at scale: the multi-room scene (12 m x 8 m), 1280x720
    uint16 depth, vl 2 mm, 2^24 buckets; integrate each frame and raycast
    each pose with fp16 (sdf, w), then the same sequence with fp32 storage.
    Reports the fp16 throughput and the fp16-vs-fp32 difference of the two
    maps (storage error at scale; the error against an fp64 oracle is
    measured on a crop by tests/test_gpu_fp16_scale.py)."""
    import torch
    from superscanner_amd import synth
    from superscanner_amd.voxel_grid import GPUVoxelGrid
    cfg = dict(synth.FP16_LARGE)
    H, W = cfg["H"], cfg["W"]
    poses64 = synth.traj_multiroom(frames)
    t0 = time.time()
    depth = synth.frames(synth.scene_multiroom(), poses64, cfg, seed=12)
    gen_s = time.time() - t0
    poses = np.stack([synth.pose_f32(T) for T in poses64])
    dd = torch.from_numpy(depth.view(np.int16)).to(dev)
    del depth
    out = {"workload": f"multi-room 12x8 m, {frames} frames {W}x{H}, vl 2 mm, trunc 8 mm, 2^24 buckets; "
                       "integrate each frame, raycast each pose",
           "frames": frames, "generation_s": round(gen_s, 1)}
    maps = {}
    for st in ("f16", "f32"):
        g = GPUVoxelGrid(1 << 30, cfg["truncation"], cfg["voxel_length"], H, W, n_buckets=cfg["n_buckets"],
                         n_bricks=1 << 23, storage=st)
        g.setIntrinsic(*cfg["intr"])
        run = Runner(g, dd, poses, frames)
        run.ops = [o for k in range(frames) for o in (("i", k), ("r", k))]
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        run.step()
        b.record()
        b.synchronize()
        ms = a.elapsed_time(b)
        stats = g.stats()
        g.sync()
        bpv = 2 if st == "f16" else 4
        out[st] = {"ms": round(ms, 1), "ms_per_frame": round(ms / frames, 3),
                   "Mpx_per_s": round(frames * H * W / ms / 1e3, 1), "bricks": stats["n_bricks"],
                   "payload_GB": round(stats["n_bricks"] * 512 * 2 * bpv / 1e9, 2)}
        if st == "f16":
            # the same sequence again (untimed above) with an event pair
            # around every launch: per-kernel times, and the update's
            # algorithmic bytes (touched bricks' payload read, updated
            # voxels written, slots) at 4 B per fp16 voxel
            g.profile(True)
            g.profile_read(reset=True)
            run.step()
            torch.cuda.synchronize()
            kt = g.profile_read(reset=True)
            g.profile(False)
            g.sync()
            s2 = g.stats()
            n_int = frames
            touched = s2["total_touched"] / n_int
            upd = s2["total_updated"] / n_int
            kern = {}
            for name in ("collect", "integrate", "raycast"):
                k = kt[name]
                if k["launches"]:
                    kern[name] = {"avg_us": round(k["ms"] * 1e3 / k["launches"], 2)}
            ib = touched * 512 * 2 * bpv + upd * 2 * bpv + touched * S_H
            if "integrate" in kern:
                gbs = ib / (kern["integrate"]["avg_us"] * 1e-6) / 1e9
                kern["integrate"].update({"bytes_per_launch": int(ib), "achieved_GBs": round(gbs, 1),
                                          "frac": round(gbs / HBM_PEAK_GBS, 4)})
            out[st]["kernels"] = kern
        e = g.export_bricks()
        c = e["coords"].long()
        key = ((c[:, 0] + (1 << 20)) << 42) | ((c[:, 1] + (1 << 20)) << 21) | (c[:, 2] + (1 << 20))
        order = torch.argsort(key)
        maps[st] = {"key": key[order], "sdf": e["sdf"][order], "w": e["w"][order]}
        # depth of the last pose, for the map comparison
        g.raycast(poses[frames - 1])
        maps[st]["depth"] = g.rayDepth.clone()
        del e, c, key, order
        g.close()
        print(f"[bench] fp16 leg: {st} storage done", file=sys.stderr, flush=True)
    a16, a32 = maps["f16"], maps["f32"]
    same_keys = bool(torch.equal(a16["key"], a32["key"]))
    out["same_brick_set"] = same_keys
    if same_keys:
        obs = a16["w"] > 0
        d = (a16["sdf"] - a32["sdf"]).abs()[obs]
        out["f16_vs_f32"] = {"weights_identical": bool(torch.equal(a16["w"], a32["w"])),
                             "observed_voxels": int(obs.sum()),
                             "max_abs_sdf_diff_m": float(d.max()),
                             "rms_sdf_diff_m": float(d.pow(2).mean().sqrt())}
    both = (a16["depth"] > 0) & (a32["depth"] > 0)
    dz = (a16["depth"] - a32["depth"])[both]
    out.setdefault("f16_vs_f32", {})["raycast_depth_rmse_m"] = float(dz.pow(2).mean().sqrt()) if dz.numel() else None
    out["speedup_f16_vs_f32"] = round(out["f32"]["ms"] / out["f16"]["ms"], 3)
    return out


CAL_BYTES = 256 << 20   # calibration copy: 256 MiB read + 256 MiB written
KERNEL_NAMES = {"collect": "k_collect(", "integrate": "k_integrate<", "ray_bounds": "k_ray_bounds<",
                "ray_march": "k_raycast<", "ray_tail": "k_raycast_tail<"}
RAYCAST_PARTS = ("ray_bounds", "ray_march", "ray_tail")  # the kernels of one tsdf_raycast


def maintenance_leg(grid):
    """cleanMemory and mesh extraction on the map the timed steps built:
    marching cubes (the reference's marchingCubesHost) and the marching
    tetrahedra option."""
    import torch
    torch.cuda.synchronize()
    out = {"bricks": grid.size()}
    reps = 5
    # marching cubes: indexed mesh with normals, device buffers sized once
    t0 = time.perf_counter()
    m = grid.extract_mesh(method="cubes")
    out["cubes_first_call_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    V, T = int(m["vertices"].shape[0]), int(m["faces"].shape[0])
    dev = grid.device
    verts = torch.empty((V, 3), dtype=torch.float64, device=dev)
    nrm = torch.empty((V, 3), dtype=torch.float64, device=dev)
    faces = torch.empty((T, 3), dtype=torch.int32, device=dev)
    nv, nt = C.c_uint64(0), C.c_uint64(0)
    t0 = time.perf_counter()
    for _ in range(reps):
        grid._L.tsdf_marching_cubes(grid._h, C.c_void_p(verts.data_ptr()), None, C.c_void_p(nrm.data_ptr()),
                                    C.c_void_p(faces.data_ptr()), None, V, T, C.byref(nv), C.byref(nt), None)
    cubes_ms = (time.perf_counter() - t0) * 1e3 / reps
    out["cubes"] = {"triangles": T, "vertices": V, "ms": round(cubes_ms, 3),
                    "Mtri_per_s": round(T / (cubes_ms * 1e-3) / 1e6, 1),
                    "note": "tsdf_marching_cubes per call, synchronous: vertices (double), faces, vertex "
                            "normals in device buffers"}
    del verts, nrm, faces
    t0 = time.perf_counter()
    m = grid.extract_mesh(method="tetrahedra")
    torch.cuda.synchronize()
    out["mesh_first_call_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    n_tri = int(m["triangles"].shape[0])
    cap = n_tri + 1
    tris = torch.empty((cap, 3, 3), dtype=torch.float32, device=grid.device)
    cnt = C.c_uint64(0)
    t0 = time.perf_counter()
    for _ in range(reps):
        grid._L.tsdf_extract_mesh(grid._h, C.c_void_p(tris.data_ptr()), None, None, cap, C.byref(cnt), None)
    mesh_ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        grid.extract_mesh(method="tetrahedra")
    indexed_ms = (time.perf_counter() - t0) * 1e3 / reps
    out["tetrahedra"] = {"triangles": n_tri, "vertices": int(m["vertices"].shape[0]), "mesh_ms": round(mesh_ms, 3),
                         "mesh_Mtri_per_s": round(n_tri / (mesh_ms * 1e-3) / 1e6, 1),
                         "indexed_mesh_ms": round(indexed_ms, 2),
                         "note": "mesh_ms: tsdf_extract_mesh (triangle soup) per call, synchronous; "
                                 "indexed_mesh_ms: GPUVoxelGrid.extract_mesh(method='tetrahedra') end to end (soup + "
                                 "keys, welding and vertex normals on the GPU, copy of the indexed mesh to the host)"}
    t0 = time.perf_counter()
    removed = grid.cleanMemory()
    out["clean_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
    out["clean_removed"] = removed
    return out


def icp_leg(dev, reps=50):
    """Projective ICP between two 640x480 frames of the room: the
    association kernel timed with HIP events on its stream, and the whole
    kokkosICP (10 iterations) per call."""
    import torch
    from superscanner_amd import icp, synth
    h, w = 480, 640
    intr = (525.0, 525.0, 319.5, 239.5)
    room = synth.scene_room(seed=1)
    T1 = synth.look_at((0.5, 0.3, 1.4), (2.4, 1.9, 0.8))
    M = np.eye(4)
    M[:3, :3] = synth.axis_angle((0.3, 1.0, 0.2), np.deg2rad(1.5))
    M[:3, 3] = (0.01, -0.008, 0.012)
    T2 = T1 @ M
    d1 = synth.render_depth(room, T1, h, w, intr, noise_mm=1.0, max_range_m=6.0, seed=1)
    d2 = synth.render_depth(room, T2, h, w, intr, noise_mm=1.0, max_range_m=6.0, seed=2)
    x = icp.CreatePcdFromDepthView(d1, intr)
    y = icp.CreatePcdFromDepthView(d2, intr)
    yn = icp.CalculateNormalsFromPCDView(y)
    Txy = np.linalg.inv(T2) @ T1
    s = torch.cuda.current_stream()
    out = {"frame": [w, h]}
    for name, t in (("point_to_point", icp.PointToPoint), ("point_to_plane", icp.PointToPlane)):
        icp.associate(x, y, yn, Txy, intr, icp_type=t)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        for _ in range(reps):
            xn = icp.associate(x, y, yn, Txy, intr, icp_type=t)
        b.record(s)
        b.synchronize()
        us = a.elapsed_time(b) * 1e3 / reps
        # compulsory bytes: x and the output (12 B each), y once (12 B), and
        # the normals of y once for point to plane
        per_px = 36 + (12 if t == icp.PointToPlane else 0)
        gbs = per_px * h * w / (us * 1e-6) / 1e9
        out[name] = {"associate_us": round(us, 2), "bytes_per_px": per_px, "achieved_GBs": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "matched_frac": round(float((xn[..., 2] != 0).float().mean()), 3)}
    dd1 = torch.from_numpy(d1.view(np.int16)).to(dev)
    dd2 = torch.from_numpy(d2.view(np.int16)).to(dev)
    icp.kokkosICP(dd1, dd2, T1, T2, intr)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        T, rmse, fit = icp.kokkosICP(dd1, dd2, T1, T2, intr)
    out["kokkosICP_ms"] = round((time.perf_counter() - t0) * 1e3 / 5, 3)
    out["kokkosICP_rmse"], out["kokkosICP_fitness"] = round(rmse, 6), round(fit, 4)
    return out


def ingest_leg(dev, n=32, H=480, W=848):
    """Recorded frames from disk to HBM: 16-bit PNG depth + PNG colour decoded
    by a thread pool, staged in pinned memory, uploaded on a copy stream; and
    the pinned host -> device copy rate alone (PCIe-inclusive figures)."""
    import tempfile
    import torch
    from PIL import Image
    from superscanner_amd import frames as F
    rng = np.random.default_rng(0)
    out = {"frames": n, "size": [W, H]}
    with tempfile.TemporaryDirectory() as root:
        os.makedirs(os.path.join(root, "depth"))
        os.makedirs(os.path.join(root, "color"))
        base = (rng.integers(400, 5000, size=(H, W))).astype(np.uint16)
        col = rng.integers(0, 256, size=(H, W, 3), dtype=np.uint8)
        for k in range(n):
            Image.fromarray(np.roll(base, k, axis=1)).save(os.path.join(root, "depth", F.getPicNumberString(k) + ".png"))
            Image.fromarray(np.roll(col, k, axis=1)).save(os.path.join(root, "color", F.getPicNumberString(k) + ".png"))
        F.write_intrinsic(os.path.join(root, "intrinsic.txt"), W, H,
                          [[430.0, 0, 423.5], [0, 430.0, 239.5], [0, 0, 1]])
        for with_color in (False, True):
            rd = F.SequenceReader(root, 0, n, H, W, with_color=with_color, prefetch=8, workers=8)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for f in rd:
                pass
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            rd.close()
            key = "depth_and_color" if with_color else "depth_only"
            out[key] = {"frames_per_s": round(n / dt, 1),
                        "MB_per_s": round(n * H * W * (5 if with_color else 2) / dt / 1e6, 1)}
    nbytes = 256 << 20
    h = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
    d = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    d.copy_(h, non_blocking=True)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(5):
        d.copy_(h, non_blocking=True)
    b.record()
    b.synchronize()
    out["pinned_h2d_GBs"] = round(5 * nbytes / (a.elapsed_time(b) * 1e-3) / 1e9, 1)
    out["depth_frame_h2d_us"] = round(H * W * 2 / (out["pinned_h2d_GBs"] * 1e9) * 1e6, 1)
    return out


def pmc_workload(args):
    """Two steps of the workload and one plain copy kernel of known size."""
    import torch
    from superscanner_amd.voxel_grid import GPUVoxelGrid
    cfg, prims, poses64, poses, depth, _ = workload(args.frames)
    dev = torch.device("cuda", 0)
    depth_dev = torch.from_numpy(depth.view(np.int16)).to(dev)
    grid = GPUVoxelGrid(1 << 30, cfg["truncation"], cfg["voxel_length"], cfg["H"], cfg["W"],
                        n_buckets=cfg["n_buckets"], n_bricks=1 << 21, max_weight=cfg["max_weight"])
    grid.setIntrinsic(*cfg["intr"])
    run = Runner(grid, depth_dev, poses, args.frames)
    for _ in range(2):
        run.step()
    a = torch.ones(CAL_BYTES // 4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    torch.mul(a, 1.0, out=b)  # the calibration copy: one elementwise kernel
    torch.cuda.synchronize()
    grid.close()
    print("[bench] pmc workload done", file=sys.stderr, flush=True)


def pmc_batch(args):
    """One batched reintegration of 64 keyframes (the batched_reintegration
    workload at the default tuning) and the same 128 plain calls."""
    import torch
    from superscanner_amd import synth
    from superscanner_amd.voxel_grid import GPUVoxelGrid
    cfg = dict(synth.ROOM_LARGE)
    poses64 = synth.traj_random_walk(n=args.frames, seed=2)
    prims = synth.scene_room(seed=1)
    rng = np.random.default_rng(21)
    n = len(poses64)
    kf = np.linspace(0, n - 1, args.keyframes).astype(int)
    keep = rng.uniform(0.3, 1.0, args.keyframes)
    kdepth = synth.frames(prims, [poses64[k] for k in kf], cfg, seed=31, keep=keep)
    old = np.stack([synth.pose_f32(poses64[k]) for k in kf])
    new = np.stack([synth.pose_f32(synth.perturb(poses64[k], rng)) for k in kf])
    dev = torch.device("cuda", 0)
    kd = torch.from_numpy(kdepth.view(np.int16)).to(dev)
    grids = []
    for _ in range(2):
        g = GPUVoxelGrid(1 << 30, cfg["truncation"], cfg["voxel_length"], cfg["H"], cfg["W"],
                         n_buckets=cfg["n_buckets"], n_bricks=1 << 18, max_weight=cfg["max_weight"])
        g.setIntrinsic(*cfg["intr"])
        for k in range(args.keyframes):
            g.integrate(None, kd[k], old[k])
        grids.append(g)
    a, b = grids
    b.reserve_batch(2 * args.keyframes)
    torch.cuda.synchronize()
    b.reintegrate_batch([kd[k] for k in range(args.keyframes)], old, new)
    for k in range(args.keyframes):
        a.deIntegrate(None, kd[k], old[k])
        a.integrate(None, kd[k], new[k])
    torch.cuda.synchronize()
    for g in grids:
        g.close()
    print("[bench] pmc batch done", file=sys.stderr, flush=True)


def summarize_pmc(fetch_csv, write_csv):
    """HBM bytes per launch of each kernel from FETCH_SIZE / WRITE_SIZE (KiB
    per dispatch), each scaled so that the calibration copy reads exactly its
    known size."""
    import csv
    import statistics

    def load(path, counter):
        per = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            name = r["Kernel_Name"]
            key = next((k for k, pat in KERNEL_NAMES.items() if pat in name), None)
            if key is None and "MulFunctor" in name:
                key = "copy"
            if key:
                per.setdefault(key, []).append(float(r["Counter_Value"]) * 1024.0)
        return {k: statistics.median(v) for k, v in per.items()}

    fetch, write = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    f_scale, w_scale = CAL_BYTES / fetch["copy"], CAL_BYTES / write["copy"]
    out = {"calibration": {"copy_bytes": CAL_BYTES, "fetch_scale": round(f_scale, 4),
                           "write_scale": round(w_scale, 4)}}
    lines = ["HBM traffic per launch (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE; median over the",
             "launches of two room_large steps), each counter scaled by the factor that makes a plain",
             f"256 MiB copy kernel read / write exactly 256 MiB (fetch x{f_scale:.3f}, write x{w_scale:.3f}).",
             "Commands:",
             "  rocprofv3 --pmc FETCH_SIZE --kernel-trace --output-format csv -d <dir> -o p -- python3 bench.py "
             "--pmc-workload",
             "  rocprofv3 --pmc WRITE_SIZE --kernel-trace --output-format csv -d <dir> -o p -- python3 bench.py "
             "--pmc-workload",
             "  python3 bench.py --summarize-pmc <fetch dir>/p_counter_collection.csv "
             "<write dir>/p_counter_collection.csv", ""]
    for k in KERNEL_NAMES:
        if k in fetch and k in write:
            fb, wb = fetch[k] * f_scale, write[k] * w_scale
            out[k] = {"fetch_bytes_per_launch": int(fb), "write_bytes_per_launch": int(wb),
                      "hbm_bytes_per_launch": int(fb + wb)}
            lines.append(f"{k:10s} fetch {fb / 1e6:9.2f} MB  write {wb / 1e6:9.2f} MB  total {(fb + wb) / 1e6:9.2f} MB")
    parts = [out[k] for k in RAYCAST_PARTS if k in out]
    if parts:  # one tsdf_raycast: the bounds, the per-lane march and the deferred-ray march
        out["raycast"] = {f: sum(p[f] for p in parts) for f in parts[0]}
        r = out["raycast"]
        lines.append(f"{'raycast':10s} fetch {r['fetch_bytes_per_launch'] / 1e6:9.2f} MB  write "
                     f"{r['write_bytes_per_launch'] / 1e6:9.2f} MB  total {r['hbm_bytes_per_launch'] / 1e6:9.2f} MB"
                     "  (the three kernels of one call)")
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    with open(os.path.join(ROOT, "profiles", "pmc_traffic.json"), "w") as f:
        json.dump(out, f, indent=1)
    with open(os.path.join(ROOT, "profiles", "pmc_traffic_room_large.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def summarize_sq(dbs):
    """Per-kernel medians (over dispatches) of every counter in rocprofv3's
    SQLite results, with the median dispatch duration."""
    import collections
    import sqlite3
    import statistics
    out = {}
    for db in dbs:
        c = sqlite3.connect(db)
        rows = c.execute("select kernel_name, counter_name, value, duration, dispatch_id from counters_collection")
        per = collections.defaultdict(lambda: collections.defaultdict(list))
        dur = collections.defaultdict(dict)
        for k, cn, v, d, disp in rows:
            short = k.replace("void ", "").replace("(anonymous namespace)::", "").split("(")[0]
            per[short][cn].append(v)
            dur[short][disp] = d
        for k in per:
            e = out.setdefault(k, {})
            for cn, vals in per[k].items():
                e[cn] = statistics.median(vals)
            e.setdefault("dur_us", statistics.median(dur[k].values()) / 1e3)
            e.setdefault("dispatches", len(dur[k]))
    print(json.dumps(out, indent=1))


def main():
    args = parse()
    if args.summarize_sq:
        summarize_sq(args.summarize_sq)
        return
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build has no "
              "multi-GPU path)", file=sys.stderr)
        sys.exit(2)
    import torch
    from superscanner_amd.voxel_grid import GPUVoxelGrid
    assert torch.cuda.is_available(), "bench.py needs an MI355X"
    if args.pmc_workload:
        pmc_workload(args)
        return
    if args.pmc_batch:
        pmc_batch(args)
        return
    dev = torch.device("cuda", 0)
    frames = args.frames
    cfg, prims, poses64, poses, depth, gen_s = workload(frames)
    H, W = cfg["H"], cfg["W"]
    print(f"[bench] generated {frames} frames {W}x{H} in {gen_s:.1f} s", file=sys.stderr, flush=True)
    depth_dev = torch.from_numpy(depth.view(np.int16)).to(dev)
    grid = GPUVoxelGrid(1 << 30, cfg["truncation"], cfg["voxel_length"], H, W, n_buckets=cfg["n_buckets"],
                        n_bricks=1 << 21, max_weight=cfg["max_weight"])
    grid.setIntrinsic(*cfg["intr"])
    if args.tuning:
        grid.set_tuning(**json.loads(args.tuning))
        tuning = grid.get_tuning()
        print(f"[bench] given launch shapes: {tuning}", file=sys.stderr, flush=True)
    else:
        tuning = autotune_plain(cfg, depth_dev, poses, frames)
        grid.set_tuning(**tuning)
        print(f"[bench] autotuned launch shapes: {tuning}", file=sys.stderr, flush=True)
    run = Runner(grid, depth_dev, poses, frames)

    for _ in range(args.warmup):
        run.step()
    # The raycast's tile order, settled on the timed workload itself (untimed):
    # the tuner's sample, 20 poses over an 80-frame map, has few of the slow
    # tiles that slowest-first dispatch puts first, and splits between the
    # two orders, while over the whole sequence slowest-first is ~13 % faster.
    order_ms = {}
    for od in ((2, 0) if not args.tuning else ()):
        grid.set_tuning(raycast_order=od)
        run.step()  # builds the tile costs the order is taken from
        oa, ob = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        oa.record()
        run.step()
        ob.record()
        ob.synchronize()
        order_ms[od] = round(oa.elapsed_time(ob), 2)
    if order_ms:
        tuning["raycast_order"] = min(order_ms, key=order_ms.get)
        grid.set_tuning(raycast_order=tuning["raycast_order"])
        run.step()
    torch.cuda.synchronize()
    print(f"[bench] raycast order on the timed workload: step ms {order_ms}", file=sys.stderr, flush=True)
    st = grid.stats()  # counters of one step (each step starts from reset)
    print(f"[bench] warmup done: {st['n_bricks']} bricks", file=sys.stderr, flush=True)

    # timed region
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        run.step()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    print(f"[bench] timed {args.steps} steps: {ms:.2f} ms/step (wall {wall * 1e3 / args.steps:.2f})",
          file=sys.stderr, flush=True)

    # second timed pass with an event pair around every launch
    grid.profile(True)
    grid.profile_read(reset=True)
    torch.cuda.synchronize()
    p0 = time.perf_counter()
    for _ in range(args.steps):
        run.step()
    torch.cuda.synchronize()
    ms_prof = (time.perf_counter() - p0) * 1e3 / args.steps
    kt = grid.profile_read(reset=True)
    grid.profile(False)

    hbm = device_copy(dev)
    print(f"[bench] device copy {hbm['device_copy_GBs']} GB/s", file=sys.stderr, flush=True)
    host_incl = host_inclusive(grid, depth, poses, frames, ms)
    print(f"[bench] host-inclusive steps done", file=sys.stderr, flush=True)

    u_vox, u_blk = footprint_pass(run, args.footprint_samples)
    n_int = sum(1 for o, _ in run.ops if o in "id")
    n_ray = sum(1 for o, _ in run.ops if o == "r")
    px = H * W
    # algorithmic bytes per launch, from the step's own counters
    touched = st["total_touched"] / n_int
    upd = st["total_updated"] / n_int
    new = st["total_new"] / n_int
    valid = st["total_valid_px"] / n_int
    bytes_per = {
        "collect": 2 * px + (touched + new) * S_H,
        "integrate": touched * 512 * S_V + upd * S_V + touched * S_H,
        "raycast": u_vox * S_V + u_blk * S_H + px * S_OUT,
    }
    kernels = {}
    for name in ("collect", "integrate", "raycast"):
        k = kt[name]
        if k["launches"] == 0:
            continue
        avg_ms = k["ms"] / k["launches"]
        gbs = bytes_per[name] / (avg_ms * 1e-3) / 1e9
        kernels[name] = {"avg_us": round(avg_ms * 1e3, 2), "launches_per_step": k["launches"] // args.steps,
                         "bytes_per_launch": int(bytes_per[name]), "achieved_GBs": round(gbs, 1),
                         "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "share_of_step": round(k["ms"] / args.steps / ms_prof, 3)}
    dom = max(kernels, key=lambda n: kt[n]["ms"])
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get(dom, {}).get("hbm_bytes_per_launch")
    roof = {"bound": "hbm", "kernel": dom, "achieved": kernels[dom]["achieved_GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": kernels[dom]["frac"], "traffic": traffic,
            "bytes_per_launch": kernels[dom]["bytes_per_launch"], "avg_us": kernels[dom]["avg_us"],
            "measured_copy_GBs": hbm["device_copy_GBs"],
            "frac_of_measured_copy": round(kernels[dom]["achieved_GBs"] / hbm["device_copy_GBs"], 4)}

    value = frames * px / (ms * 1e-3) / 1e6
    out = {
        "metric": "TSDF integrate+raycast Mpx/s, % HBM roofline; batched vs plain reintegrate",
        "value": round(value, 2), "unit": "Mpx/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "room_large: 600 frames 848x480 uint16 depth, random room + random walk, "
                               "vl 4 mm, trunc 20 mm, 2^22 buckets, wmax 128; integrate each frame, "
                               "deintegrate every 10th, raycast each pose",
                   "frames": frames, "height": H, "width": W, "voxel_length": cfg["voxel_length"],
                   "truncation": cfg["truncation"], "n_buckets": cfg["n_buckets"],
                   "integrate_calls": n_int, "raycast_calls": n_ray, "bricks": st["n_bricks"]},
        "roofline": roof,
        "kernels": kernels,
        "ms_per_step_profiled": round(ms_prof, 3),
        "raycast_footprint": {"u_vox": round(u_vox), "u_blk": round(u_blk), "samples": args.footprint_samples},
        "tuning": tuning,
        "raycast_order_step_ms": order_ms,
        "hbm": hbm,
        "host_inclusive": host_incl,
    }
    if not args.no_maintenance:
        out["maintenance"] = maintenance_leg(grid)
        print("[bench] mesh + clean done", file=sys.stderr, flush=True)
    out["plain_small"] = plain_small_leg(dev)
    print("[bench] plain workload done", file=sys.stderr, flush=True)
    if not args.no_icp:
        out["icp"] = icp_leg(dev)
        print("[bench] icp done", file=sys.stderr, flush=True)
    if not args.no_ingest:
        out["ingest"] = ingest_leg(dev)
        print("[bench] ingest done", file=sys.stderr, flush=True)
    if not args.no_reintegration:
        out["batched_reintegration"] = reintegration(cfg, prims, poses64, args.keyframes, dev)
        print("[bench] reintegration done", file=sys.stderr, flush=True)
    if not args.no_diff_render:
        out["diff_raycast"] = diff_render_leg(dev)
        print("[bench] differentiable raycast done", file=sys.stderr, flush=True)
    if not args.no_fp16:
        out["fp16_large"] = fp16_scale_leg(dev, args.fp16_frames)
        print("[bench] fp16 at scale done", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, poses, depth, min(args.cpu_frames, frames))
        print("[bench] cpu baseline done", file=sys.stderr, flush=True)
    grid.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
