"""Frame ingest formats on the CPU (superscanner_amd/frames.py):

  * getPicNumberString pads to six digits (coreutil.cpp:449-457);
  * intrinsic.txt: the text saveImagestoDisc writes (Eigen's matrix print,
    every coefficient right-aligned to the widest) and its parse
    (coreutil.cpp:381-430) round-trip;
  * depth resizing follows OpenCV's nearest rule, src = floor(dst * s / d);
  * a sequence directory in either naming scheme decodes to the exact
    arrays written, 16-bit depth included.
"""
import os

import numpy as np
import pytest
from PIL import Image

from superscanner_amd import frames as F


def test_pic_number_string():
    assert F.getPicNumberString(7) == "000007"
    assert F.getPicNumberString(123456) == "123456"
    assert F.getPicNumberString(1234567) == "1234567"


def test_intrinsic_text_and_round_trip(tmp_path):
    K = np.array([[525.0, 0, 319.5], [0, 525.0, 239.5], [0, 0, 1]])
    p = tmp_path / "intrinsic.txt"
    F.write_intrinsic(p, 640, 480, K)
    assert p.read_text() == "640\n480\n  525     0 319.5\n    0   525 239.5\n    0     0     1"
    w, h, K2 = F.read_intrinsic(p)
    assert (w, h) == (640, 480) and np.array_equal(K, K2)
    assert F.intrinsic_params(K2) == (525.0, 525.0, 319.5, 239.5)
    p.write_text("848\n480\n430\t0  423.5\n0 430.25 239.5   \n0 0 1\n\n")
    w, h, K3 = F.read_intrinsic(p)
    assert (w, h) == (848, 480) and K3[1, 1] == 430.25 and K3[0, 2] == 423.5
    p.write_text("640\n480\n1 2\n")
    with pytest.raises(ValueError):
        F.read_intrinsic(p)


def test_resize_nearest_rule():
    a = np.arange(4)[None, :].repeat(2, 0)
    assert F.resize_nearest(a, 2, 6)[0].tolist() == [0, 0, 1, 2, 2, 3]
    b = np.arange(6)[None, :]
    assert F.resize_nearest(b, 1, 4)[0].tolist() == [0, 1, 3, 4]
    img = np.arange(12 * 10).reshape(12, 10)
    r = F.resize_nearest(img, 5, 7)
    assert r.shape == (5, 7) and r[4, 6] == img[int(4 * 12 / 5), int(6 * 10 / 7)]
    assert F.resize_nearest(img, 12, 10) is img


def write_sequence(root, depths, rgbs, padded=True):
    os.makedirs(os.path.join(root, "depth"), exist_ok=True)
    os.makedirs(os.path.join(root, "color"), exist_ok=True)
    for k, (d, c) in enumerate(zip(depths, rgbs)):
        name = F.getPicNumberString(k) if padded else str(k)
        Image.fromarray(np.asarray(d, np.uint16)).save(os.path.join(root, "depth", name + ".png"))
        Image.fromarray(np.asarray(c, np.uint8)).save(os.path.join(root, "color", name + (".png" if padded else ".jpg")),
                                                       **({} if padded else {"quality": 100}))


@pytest.mark.parametrize("padded", [True, False])
def test_sequence_files_decode_exactly(tmp_path, padded):
    rng = np.random.default_rng(0)
    depths = [rng.integers(0, 65536, size=(30, 40)).astype(np.uint16) for _ in range(3)]
    rgbs = [rng.integers(0, 256, size=(30, 40, 3)).astype(np.uint8) for _ in range(3)]
    write_sequence(str(tmp_path), depths, rgbs, padded)
    for k in range(3):
        d, c = F.read_frame_files(str(tmp_path), k, 30, 40)
        assert d.dtype == np.uint16 and np.array_equal(d, depths[k])
        if padded:  # lossless png; the unpadded scheme stores jpg
            assert np.array_equal(c, rgbs[k])
        else:
            assert c.shape == (30, 40, 3)
    d, _ = F.read_frame_files(str(tmp_path), 1, 15, 80, with_color=False)
    assert np.array_equal(d, F.resize_nearest(depths[1], 15, 80))
    with pytest.raises(FileNotFoundError):
        F.read_frame_files(str(tmp_path), 3, 30, 40)
