"""The oracle still reproduces the committed fixture tests/golden/plain_tiny.npz
(see tests/golden/make_golden.py for how it was made)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402


def test_oracle_reproduces_fixture():
    d = np.load(os.path.join(HERE, "golden", "plain_tiny.npz"))
    intr, poses, depth, rgb = make_golden.inputs()
    assert (depth == d["depth"]).all() and (rgb == d["rgb"]).all()
    assert (poses == d["poses"]).all()
    out = make_golden.run_oracle(tuple(d["intr"]), d["poses"], d["depth"], d["rgb"])
    for k, v in out.items():
        assert np.array_equal(np.asarray(v).view(np.uint8), np.asarray(d[k]).view(np.uint8)), k
