"""Frame ingest: recorded sequences on disk (the reference's data camera,
DataCam.cpp:9-100, with intrinsic.txt, coreutil.cpp:381-430 and :522-526) and
the host <-> device moves of a frame (Frame::MovetoGPU / MovetoCPU,
Frame.cpp:29-60), staged through pinned host memory and copied asynchronously
on a copy stream, so that decoding frame k + 1 overlaps the upload of frame k
and the integration of frame k - 1.

Images are row-major on both sides: depth uint16 [H, W] in mm, colour uint8
[H, W, 3] RGB (the reference's Kokkos views change layout on the host before
the copy, DataCam.cpp:80-92; here there is nothing to change).

Layout of a sequence directory (DataCam.cpp:16-51):
  intrinsic.txt                  width, height, then the 3x3 K, one row a line
  depth/NNNNNN.png  (or N.png)   16-bit grey, mm
  color/NNNNNN.jpg, .png (or N.jpg)
A frame whose size differs from the target (H, W) is resized: depth by
OpenCV's nearest rule (DataCam.cpp:66-69, src = floor(dst * src_size /
dst_size)), colour bilinearly (OpenCV's fixed-point INTER_LINEAR is not
restated: colour resizing is not bit-exact with the reference).
"""
from __future__ import annotations

import os
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch


def getPicNumberString(a: int) -> str:
    """coreutil.cpp:449-457: zero-padded to 6 digits."""
    return str(int(a)).zfill(6)


def read_intrinsic(path):
    """setFromIntrinsicFile / assignIntrinsic, coreutil.cpp:381-430: line 0
    width, line 1 height, lines 2-4 the rows of K (whitespace-separated).
    Returns (width, height, K [3, 3] float64); lines after the fifth are
    ignored."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    if len(lines) < 5:
        raise ValueError(f"{path}: intrinsic file needs 5 lines, has {len(lines)}")
    width, height = int(lines[0].split()[0]), int(lines[1].split()[0])
    K = np.array([[float(v) for v in lines[2 + r].split()[:3]] for r in range(3)], np.float64)
    if K.shape != (3, 3):
        raise ValueError(f"{path}: K rows need 3 values each")
    return width, height, K


def _stream_double(v: float) -> str:
    """std::ostream's default output of a double (precision 6, %g without
    trailing zeros)."""
    s = "%.6g" % v
    return "-0" if s == "-0" else s


def write_intrinsic(path, width, height, K) -> None:
    """saveImagestoDisc's intrinsic.txt, coreutil.cpp:522-526: width and
    height lines, then K as Eigen prints a matrix (every coefficient
    right-aligned to the widest one, one row a line, no final newline)."""
    K = np.asarray(K, np.float64).reshape(3, 3)
    cells = [[_stream_double(v) for v in row] for row in K]
    w = max(len(c) for row in cells for c in row)
    body = "\n".join(" ".join(c.rjust(w) for c in row) for row in cells)
    with open(path, "w") as fh:
        fh.write(f"{int(width)}\n{int(height)}\n{body}")


def intrinsic_params(K):
    """(fx, fy, cx, cy) of K, in the order setIntrinsic takes them."""
    K = np.asarray(K, np.float64)
    return float(K[0, 0]), float(K[1, 1]), float(K[0, 2]), float(K[1, 2])


def resize_nearest(img, H, W):
    """cv::resize(..., INTER_NEAREST): destination pixel x reads source
    min(floor(x * src / dst), src - 1) per axis."""
    img = np.asarray(img)
    h, w = img.shape[:2]
    if (h, w) == (H, W):
        return img
    ys = np.minimum(np.floor(np.arange(H) * (h / H)).astype(np.int64), h - 1)
    xs = np.minimum(np.floor(np.arange(W) * (w / W)).astype(np.int64), w - 1)
    return img[ys][:, xs]


def _read_png16(path):
    from PIL import Image
    with Image.open(path) as im:
        if im.mode not in ("I;16", "I;16B", "I;16L", "I"):
            raise ValueError(f"{path}: depth PNG of mode {im.mode}, expected 16-bit grey")
        return np.array(im, dtype=np.uint16)


def _read_rgb(path):
    from PIL import Image
    with Image.open(path) as im:
        return np.array(im.convert("RGB"), dtype=np.uint8)


def frame_paths(root, n):
    """The files DataCam.cpp:28-51 tries, in its order: (depth, colour or None)."""
    num = getPicNumberString(n)
    depth = [os.path.join(root, "depth", num + ".png"), os.path.join(root, "depth", f"{n}.png")]
    color = [os.path.join(root, "color", num + ".jpg"), os.path.join(root, "color", num + ".png"),
             os.path.join(root, "color", f"{n}.jpg")]
    d = next((p for p in depth if os.path.exists(p)), None)
    c = next((p for p in color if os.path.exists(p)), None)
    if d is None:
        raise FileNotFoundError(f"no depth image for frame {n} under {root}")
    return d, c


def read_frame_files(root, n, H, W, with_color=True):
    """Decoded frame n of a sequence at (H, W): (depth uint16 [H, W], rgb
    uint8 [H, W, 3] or None)."""
    from PIL import Image
    dp, cp = frame_paths(root, n)
    depth = resize_nearest(_read_png16(dp), H, W)
    rgb = None
    if with_color and cp is not None:
        rgb = _read_rgb(cp)
        if rgb.shape[:2] != (H, W):
            rgb = np.array(Image.fromarray(rgb).resize((W, H), Image.BILINEAR), dtype=np.uint8)
    return np.ascontiguousarray(depth), (np.ascontiguousarray(rgb) if rgb is not None else None)


class HostFrame:
    """A frame whose images live in pinned host memory, with an optional
    device copy (Frame.h: depth_h / rgb_h and depth / rgb)."""

    n_gpu_frames = 0  # Frame::nGPUFrames
    _lock = threading.Lock()

    def __init__(self, depth, rgb=None, unique_id=-1):
        d = np.ascontiguousarray(depth, dtype=np.uint16)
        self.H, self.W = d.shape
        self.depth_h = torch.empty((self.H, self.W), dtype=torch.int16, pin_memory=True)
        self.depth_h.numpy()[...] = d.view(np.int16)
        self.rgb_h = None
        if rgb is not None:
            c = np.ascontiguousarray(rgb, dtype=np.uint8)
            self.rgb_h = torch.empty(c.shape, dtype=torch.uint8, pin_memory=True)
            self.rgb_h.numpy()[...] = c
        self.depth = None
        self.rgb = None
        self.ready = None  # event after the upload
        self.unique_id = unique_id

    @property
    def dataOnGPU(self) -> bool:
        return self.depth is not None

    def MovetoGPU(self, stream=None, device=None):
        """Frame.cpp:44-60, asynchronous on `stream` (default: the current
        stream); self.ready is recorded after the copies."""
        if self.depth is not None:
            return self
        dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        s = stream or torch.cuda.current_stream(dev)
        with torch.cuda.stream(s):
            self.depth = torch.empty((self.H, self.W), dtype=torch.int16, device=dev)
            self.depth.copy_(self.depth_h, non_blocking=True)
            if self.rgb_h is not None:
                self.rgb = torch.empty(self.rgb_h.shape, dtype=torch.uint8, device=dev)
                self.rgb.copy_(self.rgb_h, non_blocking=True)
            self.ready = torch.cuda.Event()
            self.ready.record(s)
        with HostFrame._lock:
            HostFrame.n_gpu_frames += 1
        return self

    def MovetoCPU(self, stream=None):
        """Frame.cpp:29-42: copies the device images back (they may have been
        changed there) and releases them."""
        if self.depth is None:
            return self
        s = stream or torch.cuda.current_stream(self.depth.device)
        if self.ready is not None:
            s.wait_event(self.ready)
        with torch.cuda.stream(s):
            self.depth_h.copy_(self.depth, non_blocking=True)
            if self.rgb is not None:
                self.rgb_h.copy_(self.rgb, non_blocking=True)
        s.synchronize()
        self.depth = self.rgb = self.ready = None
        with HostFrame._lock:
            HostFrame.n_gpu_frames -= 1
        return self

    def wait(self, stream=None):
        """Makes `stream` (default: the current one) wait for the upload, and
        keeps the device images alive for the work queued there."""
        if self.ready is None:
            return
        s = stream or torch.cuda.current_stream(self.depth.device)
        s.wait_event(self.ready)
        self.depth.record_stream(s)
        if self.rgb is not None:
            self.rgb.record_stream(s)


class SequenceReader:
    """Frames first..first+count-1 of a recorded sequence, decoded by a
    thread pool `depth` frames ahead and uploaded on a copy stream of their
    own; iterating yields HostFrames already on the GPU, whose uploads the
    consumer's stream waits for (DataCamThreadFunction, DataCam.cpp:94-100,
    without its 30 fps pacing)."""

    def __init__(self, root, first, count, H, W, *, with_color=True, prefetch=4, workers=4, device=None):
        self.root, self.first, self.count = root, int(first), int(count)
        self.H, self.W, self.with_color = int(H), int(W), bool(with_color)
        self.prefetch = max(1, int(prefetch))
        self.pool = ThreadPoolExecutor(max_workers=max(1, int(workers)))
        dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self.copy_stream = torch.cuda.Stream(device=dev)
        self.device = dev
        w, h, K = read_intrinsic(os.path.join(root, "intrinsic.txt"))
        self.width, self.height, self.K = w, h, K

    def _load(self, n):
        d, c = read_frame_files(self.root, n, self.H, self.W, self.with_color)
        return HostFrame(d, c, unique_id=n)

    def __iter__(self):
        pending = []
        nxt = self.first
        end = self.first + self.count
        while nxt < end and len(pending) < self.prefetch:
            pending.append(self.pool.submit(self._load, nxt))
            nxt += 1
        while pending:
            f = pending.pop(0).result()
            if nxt < end:
                pending.append(self.pool.submit(self._load, nxt))
                nxt += 1
            f.MovetoGPU(self.copy_stream, self.device.index)
            f.wait()
            yield f

    def close(self):
        self.pool.shutdown(wait=True)
