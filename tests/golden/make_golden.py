"""Writes tests/golden/plain_tiny.npz: a small fixture of the TSDF path.

The reference ships no fixture for this path and cannot be built here (its
TSDF code is Kokkos-CUDA only), so the expected values come from the oracle
(oracle/tsdf_oracle.c), which the closed-form tests pin.  The fixture freezes
them: the CPU suite checks the oracle still reproduces it, the GPU suite
checks the HIP path reproduces it bit for bit.

Scene: the plain_small scene (plane, sphere, box) at 80x60 with the same
field of view, vl 2 cm, trunc 8 cm, colour on.  Ops: integrate f0, f1, f2,
deintegrate f1, integrate f3; raycast from poses 0 and 2.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from superscanner_amd import synth  # noqa: E402

H, W = 60, 80
VL, TRUNC = 0.02, 0.08
OPS = [(0, 1), (1, 1), (2, 1), (1, -1), (3, 1)]
RAY_POSES = [0, 2]


def inputs():
    cfg = synth.PLAIN_SMALL
    intr = synth.scaled_intr(cfg, H, W)
    poses = synth.traj_plain_small(n=8)[:4]
    prims = synth.scene_plain_small()
    depth = np.stack([synth.render_depth(prims, T, H, W, intr, 1.0, 4.0, seed=77 + k)
                      for k, T in enumerate(poses)])
    rng = np.random.default_rng(5)
    rgb = rng.integers(0, 256, size=(len(poses), H, W, 3), dtype=np.uint8)
    return intr, np.stack([synth.pose_f32(T) for T in poses]), depth, rgb


def run_oracle(intr, poses, depth, rgb):
    g = oracle.OracleGrid(VL, TRUNC, H, W, with_color=True)
    g.set_intrinsics(*intr)
    counts = []
    for f, s in OPS:
        st, c = g.integrate(depth[f], poses[f], s, rgb[f])
        assert st == 0
        counts.append([c["n_touched"], c["n_new"], c["n_upd"], c["n_valid"]])
    ex = g.export()
    order = np.lexsort((ex["coords"][:, 2], ex["coords"][:, 1], ex["coords"][:, 0]))
    rays = [g.raycast(poses[k]) for k in RAY_POSES]
    out = dict(coords=ex["coords"][order], keys=ex["keys"][order], sdf=ex["sdf"][order],
               w=ex["w"][order], vrgb=ex["rgb"][order], counts=np.array(counts, np.int64))
    for n, r in zip(RAY_POSES, rays):
        for k in ("points", "colors", "depth", "ub"):
            out[f"ray{n}_{k}"] = r[k]
    g.close()
    return out


def main():
    intr, poses, depth, rgb = inputs()
    exp = run_oracle(intr, poses, depth, rgb)
    np.savez_compressed(os.path.join(HERE, "plain_tiny.npz"), intr=np.array(intr, np.float32),
                        poses=poses, depth=depth, rgb=rgb, ops=np.array(OPS, np.int32),
                        vl=np.float32(VL), trunc=np.float32(TRUNC),
                        ray_poses=np.array(RAY_POSES, np.int32), **exp)


if __name__ == "__main__":
    main()
