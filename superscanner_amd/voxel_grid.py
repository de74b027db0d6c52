"""Host-side mirror of the reference class GPUVoxelGrid
(src/VoxelGrid/GPUVoxelGrid.h:40-120) over the C ABI of include/tsdf.h.

Method names, argument order and meaning follow the reference:
``GPUVoxelGrid(c, trunc, voxlength, resi, resj)``, ``setIntrinsic``,
``integrate(rgb, depth, extr, withColor)``, ``deIntegrate(...)``,
``raycast(extr)`` filling the grid-owned ``rayPcd`` / ``rayColorPcd``.
The reference's exec-space argument becomes the current torch stream.
Differences: errors raise ``TsdfError`` instead of printing
(GPUVoxelGrid.cpp:46-49) or corrupting memory (:497), and calls are
asynchronous until ``sync()`` unless ``fence=True`` (the reference fences at
the end of every call, GPUVoxelGrid.cpp:79,133,146).

torch is used only for device memory and streams; every computation runs in
libtsdf_hip.so.
"""
from __future__ import annotations

import ctypes as C
from contextlib import nullcontext as _nullcontext

import numpy as np
import torch

from . import _lib
from ._lib import TsdfParams, TsdfStats, check


def _pose_ptr(E):
    if isinstance(E, torch.Tensor):
        E = E.detach().cpu().numpy()
    a = np.ascontiguousarray(np.asarray(E, dtype=np.float32).reshape(16))
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def _stream_ptr(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _dev_ptr(t: torch.Tensor | None, dtypes, shape=None, name="tensor"):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.dtype not in dtypes:
        raise ValueError(f"{name} has dtype {t.dtype}, expected one of {dtypes}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise _lib.TsdfError(_lib.TSDF_BAD_DIMS, name,
                             f"shape {tuple(t.shape)} does not match the grid {tuple(shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return C.c_void_p(t.data_ptr())


class GPUVoxelGrid:
    """Voxel-hashed TSDF on one MI355X (see module docstring)."""

    def __init__(self, c: int, trunc: float, voxlength: float, resi: int, resj: int, *,
                 n_buckets: int | None = None, n_bricks: int | None = None,
                 storage: str = "f32", withColor: bool = False, max_weight: float = 0.0,
                 rmin: float = 0.2, d_max: float = 10.0, device: int | None = None):
        L = _lib.load()
        p = TsdfParams()
        L.tsdf_default_params(C.byref(p))
        p.voxel_length = voxlength
        p.truncation = trunc
        p.height = resi
        p.width = resj
        # reference capacity: c voxels -> c / 512 bricks (GPUVoxelGrid.cpp:11,27)
        bricks = int(n_bricks) if n_bricks is not None else max(1, int(c) // 512)
        if n_buckets is None:
            n_buckets = 1 << max(10, (2 * bricks - 1).bit_length())
        p.n_buckets = int(n_buckets)
        p.n_bricks = bricks
        p.storage = {"f32": _lib.STORAGE_F32, "f16": _lib.STORAGE_F16}[storage]
        p.with_color = int(bool(withColor))
        p.max_weight = max_weight
        p.min_depth = rmin
        p.max_ray_length = d_max
        dev = torch.cuda.current_device() if device is None else int(device)
        p.device = dev
        self.params = p
        self.device = torch.device("cuda", dev)
        h = C.c_void_p()
        check(L.tsdf_create(C.byref(p), C.byref(h)), "tsdf_create")
        self._h = h
        self._L = L
        self.resi, self.resj = int(resi), int(resj)
        self.d_trunc, self.vl = float(trunc), float(voxlength)
        self.vbl = 8 * self.vl
        self.d_max, self.rmin = float(d_max), float(rmin)
        self.maxVoxelBlocks = bricks
        self.withColor = bool(withColor)
        self.intrinsicSet = False
        self.mesh = None  # the last marchingCubesHost mesh (GPUVoxelGrid.h:60)
        # grid-owned raycast outputs (rayPcd, rayColorPcd, GPUVoxelGrid.h:67-69)
        self.rayPcd = torch.zeros((resi, resj, 3), dtype=torch.float32, device=self.device)
        self.rayColorPcd = torch.zeros((resi, resj, 3), dtype=torch.float32, device=self.device)
        self.rayDepth = torch.zeros((resi, resj), dtype=torch.float32, device=self.device)
        self.rayUB = torch.zeros((resi, resj), dtype=torch.uint8, device=self.device)

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._L.tsdf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> dict:
        """The parameters the grid runs with (tsdf_get_params)."""
        p = TsdfParams()
        check(self._L.tsdf_get_params(self._h, C.byref(p)), "tsdf_get_params")
        return {n: getattr(p, n) for n, _ in TsdfParams._fields_}

    def reset(self, stream=None):
        check(self._L.tsdf_reset(self._h, _stream_ptr(stream)), "tsdf_reset")

    # -- reference surface -------------------------------------------------
    def setIntrinsic(self, fx, fy=None, cx=None, cy=None):
        """setIntrinsic (GPUVoxelGrid.cpp:253-266); accepts (fx, fy, cx, cy)
        or a 3x3 intrinsic matrix."""
        if fy is None:
            K = np.asarray(fx, np.float64)
            fx, fy, cx, cy = K[0, 0], K[1, 1], K[0, 2], K[1, 2]
        check(self._L.tsdf_set_intrinsics(self._h, float(fx), float(fy), float(cx), float(cy)),
              "tsdf_set_intrinsics")
        self.fx, self.fy, self.cx, self.cy = float(fx), float(fy), float(cx), float(cy)
        self.intrinsicSet = True

    def _integrate(self, rgb, depth, extr, withColor, sign, stream, fence):
        dptr = _dev_ptr(depth, (torch.uint16, torch.int16), None, "depth")
        if depth.dim() != 2 or depth.shape[0] != self.resi or depth.shape[1] < self.resj:
            raise _lib.TsdfError(_lib.TSDF_BAD_DIMS, "integrate",
                                 "The input picture dimensions are not correct, they need to "
                                 "match the constructor dimensions!")
        cptr = None
        if withColor and rgb is not None:
            cptr = _dev_ptr(rgb, (torch.uint8,), (self.resi, self.resj, 3), "rgb")
        keep, E = _pose_ptr(extr)
        check(self._L.tsdf_integrate(self._h, dptr, int(depth.stride(0)), cptr, E, sign,
                                     _stream_ptr(stream)), "tsdf_integrate")
        if fence:
            self.sync(stream)

    def integrate(self, rgb, depth, extr, withColor=True, stream=None, fence=False):
        """GPUVoxelGrid::integrate (GPUVoxelGrid.cpp:41-83)."""
        self._integrate(rgb, depth, extr, withColor, 1, stream, fence)

    def deIntegrate(self, rgb, depth, extr, withColor=True, stream=None, fence=False):
        """GPUVoxelGrid::deIntegrate (GPUVoxelGrid.cpp:95-136)."""
        self._integrate(rgb, depth, extr, withColor, -1, stream, fence)

    def raycast(self, extr, stream=None, fence=False):
        """GPUVoxelGrid::raycast (GPUVoxelGrid.cpp:139-147): fills rayPcd,
        rayColorPcd (and rayDepth, rayUB)."""
        keep, E = _pose_ptr(extr)
        check(self._L.tsdf_raycast(self._h, E, C.c_void_p(self.rayPcd.data_ptr()),
                                   C.c_void_p(self.rayColorPcd.data_ptr()),
                                   C.c_void_p(self.rayDepth.data_ptr()),
                                   C.c_void_p(self.rayUB.data_ptr()), _stream_ptr(stream)),
              "tsdf_raycast")
        if fence:
            self.sync(stream)

    # -- batched reintegration ----------------------------------------------
    def integrate_batch(self, depths, poses, signs, rgbs=None, stream=None, fence=False):
        """Ordered batch of integrate (+1) / deintegrate (-1) ops in one pass;
        bit-identical to the same sequence of plain calls."""
        n = len(depths)
        if not (len(poses) == n and len(signs) == n):
            raise ValueError("depths, poses and signs must have equal length")
        pitch = None
        dp = (C.c_void_p * max(1, n))()
        for k, d in enumerate(depths):
            dp[k] = _dev_ptr(d, (torch.uint16, torch.int16), (self.resi, d.shape[1]), "depth").value
            pitch = int(d.stride(0)) if pitch is None else pitch
            if int(d.stride(0)) != pitch:
                raise ValueError("all depth images of a batch must share one pitch")
        cp = None
        if rgbs is not None:
            cp = (C.c_void_p * max(1, n))()
            for k, c in enumerate(rgbs):
                cp[k] = _dev_ptr(c, (torch.uint8,), (self.resi, self.resj, 3), "rgb").value
        E = np.ascontiguousarray(np.stack([np.asarray(p, np.float32).reshape(16) for p in poses]))
        S = np.ascontiguousarray(np.asarray(signs, np.int32))
        check(self._L.tsdf_integrate_batch(self._h, n, dp, pitch or self.resj, cp,
                                           E.ctypes.data_as(C.POINTER(C.c_float)),
                                           S.ctypes.data_as(C.POINTER(C.c_int32)),
                                           _stream_ptr(stream)), "tsdf_integrate_batch")
        if fence:
            self.sync(stream)

    def reintegrate_batch(self, depths, old_poses, new_poses, rgbs=None, stream=None, fence=False):
        """Keyframe k: deIntegrate(old_k) then integrate(new_k), k = 0..n-1
        (the order of integrate.cpp:108-125), as one batch."""
        n = len(depths)
        dp = (C.c_void_p * max(1, n))()
        pitch = None
        for k, d in enumerate(depths):
            dp[k] = _dev_ptr(d, (torch.uint16, torch.int16), (self.resi, d.shape[1]), "depth").value
            pitch = int(d.stride(0)) if pitch is None else pitch
        cp = None
        if rgbs is not None:
            cp = (C.c_void_p * max(1, n))()
            for k, c in enumerate(rgbs):
                cp[k] = _dev_ptr(c, (torch.uint8,), (self.resi, self.resj, 3), "rgb").value
        Eo = np.ascontiguousarray(np.stack([np.asarray(p, np.float32).reshape(16) for p in old_poses]))
        En = np.ascontiguousarray(np.stack([np.asarray(p, np.float32).reshape(16) for p in new_poses]))
        check(self._L.tsdf_reintegrate_batch(self._h, n, dp, pitch or self.resj, cp,
                                             Eo.ctypes.data_as(C.POINTER(C.c_float)),
                                             En.ctypes.data_as(C.POINTER(C.c_float)),
                                             _stream_ptr(stream)), "tsdf_reintegrate_batch")
        if fence:
            self.sync(stream)

    def reserve_batch(self, n_ops):
        """Allocates the device buffers a batch of n_ops ops needs, so that
        later batches of at most n_ops ops allocate nothing."""
        check(self._L.tsdf_reserve_batch(self._h, int(n_ops)), "tsdf_reserve_batch")

    # -- state -------------------------------------------------------------
    def sync(self, stream=None):
        check(self._L.tsdf_sync(self._h, _stream_ptr(stream)), "tsdf_sync")

    def stats(self, stream=None) -> dict:
        s = TsdfStats()
        check(self._L.tsdf_stats(self._h, C.byref(s), _stream_ptr(stream)), "tsdf_stats")
        return s.as_dict()

    def raycast_footprint(self, extr, stream=None) -> dict:
        """Distinct voxels read and distinct keys probed by the raycast from
        `extr` (diagnostic, synchronous)."""
        keep, E = _pose_ptr(extr)
        nv, nk = C.c_uint64(0), C.c_uint64(0)
        check(self._L.tsdf_raycast_footprint(self._h, E, C.byref(nv), C.byref(nk), _stream_ptr(stream)),
              "tsdf_raycast_footprint")
        return {"u_vox": int(nv.value), "u_blk": int(nk.value)}

    def tuning(self) -> dict:
        t = _lib.TsdfTuning()
        check(self._L.tsdf_get_tuning(self._h, C.byref(t)), "tsdf_get_tuning")
        return t.as_dict()

    def get_tuning(self) -> dict:
        """The grid's launch shapes (tsdf_tuning) as a dict."""
        t = _lib.TsdfTuning()
        check(self._L.tsdf_get_tuning(self._h, C.byref(t)), "tsdf_get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw):
        t = _lib.TsdfTuning()
        check(self._L.tsdf_get_tuning(self._h, C.byref(t)), "tsdf_get_tuning")
        names = {n for n, _ in t._fields_}
        for k, v in kw.items():
            if k not in names:
                raise ValueError(f"unknown tuning field {k!r}; fields: {sorted(names)}")
            setattr(t, k, int(v))
        check(self._L.tsdf_set_tuning(self._h, C.byref(t)), "tsdf_set_tuning")

    def profile(self, enable=True):
        check(self._L.tsdf_profile(self._h, int(bool(enable))), "tsdf_profile")

    def profile_read(self, reset=True) -> dict:
        t = _lib.TsdfKernelTimes()
        check(self._L.tsdf_profile_read(self._h, C.byref(t), int(bool(reset))), "tsdf_profile_read")
        return t.as_dict()

    def size(self) -> int:
        """voxelMap2.size()"""
        return self.stats()["n_bricks"]

    def cleanMemory(self, stream=None) -> int:
        """GPUVoxelGrid::cleanMemory (GPUVoxelGrid.cpp:185-196): erases the
        bricks whose weights are all zero; returns how many."""
        n = C.c_uint64(0)
        check(self._L.tsdf_clean(self._h, C.byref(n), _stream_ptr(stream)), "tsdf_clean")
        return int(n.value)

    def extract_mesh(self, stream=None, method: str = "cubes") -> dict:
        """Mesh of the zero level set.

        method "cubes" (the reference's: marchingCubesHost,
        GPUVoxelGrid.cpp:200-251, Open3D's ExtractTriangleMesh and
        ComputeVertexNormals), tsdf_marching_cubes: {"vertices": [k, 3]
        float64, "faces": [m, 3] int64, "normals": [k, 3] float64 unit
        area-weighted vertex normals (zero where the sum is zero), "colors":
        [k, 3] float64 in 0..1 (rgb / 255) or None, "vertex_keys": [k] int64
        lattice edge of each vertex, "color_scale": 1.0} as host numpy arrays.

        method "tetrahedra": marching tetrahedra (tsdf_extract_mesh), welded on
        the GPU by vertex key: {"triangles": [m, 3, 3] device float32 soup,
        "triangle_colors": [m, 3, 3] device float32 or None, "vertices":
        [k, 3] float32, "faces": [m, 3] int64, "normals": [k, 3] float32,
        "colors": [k, 3] float32 in 0..255 or None, "color_scale": 255.0} --
        the last four as host numpy arrays."""
        if method == "cubes":
            return self._marching_cubes(stream)
        if method != "tetrahedra":
            raise ValueError(f"unknown mesh method {method!r} (cubes or tetrahedra)")
        cap = max(1024, 64 * self.size())
        while True:
            tris = torch.empty((cap, 3, 3), dtype=torch.float32, device=self.device)
            cols = torch.empty((cap, 3, 3), dtype=torch.float32, device=self.device) if self.withColor else None
            keys = torch.empty((cap, 3), dtype=torch.int64, device=self.device)
            n = C.c_uint64(0)
            check(self._L.tsdf_extract_mesh(self._h, C.c_void_p(tris.data_ptr()),
                                            C.c_void_p(cols.data_ptr()) if cols is not None else None,
                                            C.c_void_p(keys.data_ptr()), cap, C.byref(n), _stream_ptr(stream)),
                  "tsdf_extract_mesh")
            if n.value <= cap:
                break
            cap = int(n.value)
        m = int(n.value)
        with torch.cuda.stream(stream) if stream is not None else _nullcontext():
            tris = tris[:m]
            cols = cols[:m] if cols is not None else None
            uk, inv = torch.unique(keys[:m].reshape(-1), return_inverse=True)
            k = uk.numel()
            verts = torch.zeros((k, 3), dtype=torch.float32, device=self.device)
            verts.index_copy_(0, inv, tris.reshape(-1, 3))  # equal keys carry equal positions
            faces = inv.reshape(-1, 3)
            t = tris.double()
            fn = torch.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0], dim=1)
            vn = torch.zeros((k, 3), dtype=torch.float64, device=self.device)
            for q in range(3):
                vn.index_add_(0, faces[:, q], fn)
            norm = vn.norm(dim=1, keepdim=True)
            vn = torch.where(norm > 0, vn / torch.where(norm > 0, norm, torch.ones_like(norm)), vn).float()
            vcol = None
            if cols is not None:
                vcol = torch.zeros((k, 3), dtype=torch.float32, device=self.device)
                vcol.index_copy_(0, inv, cols.reshape(-1, 3))
        return {"triangles": tris, "triangle_colors": cols, "vertices": verts.cpu().numpy(),
                "faces": faces.cpu().numpy(), "normals": vn.cpu().numpy(),
                "colors": vcol.cpu().numpy() if vcol is not None else None, "color_scale": 255.0}

    def marchingCubesHost(self, stream=None) -> dict:
        """GPUVoxelGrid::marchingCubesHost (GPUVoxelGrid.cpp:200-251): the
        marching-cubes mesh, kept as ``self.mesh`` and returned."""
        self.mesh = self._marching_cubes(stream)
        return self.mesh

    def _marching_cubes(self, stream=None) -> dict:
        L, dev = self._L, self.device
        n_b = max(1, self.size())
        vcap, tcap = 64 * n_b, 128 * n_b
        while True:
            verts = torch.empty((max(vcap, 1), 3), dtype=torch.float64, device=dev)
            cols = torch.empty((max(vcap, 1), 3), dtype=torch.float64, device=dev) if self.withColor else None
            nrm = torch.empty((max(vcap, 1), 3), dtype=torch.float64, device=dev)
            keys = torch.empty((max(vcap, 1),), dtype=torch.int64, device=dev)
            faces = torch.empty((max(tcap, 1), 3), dtype=torch.int32, device=dev)
            nv, nt = C.c_uint64(0), C.c_uint64(0)
            check(L.tsdf_marching_cubes(self._h, C.c_void_p(verts.data_ptr()),
                                        C.c_void_p(cols.data_ptr()) if cols is not None else None,
                                        C.c_void_p(nrm.data_ptr()), C.c_void_p(faces.data_ptr()),
                                        C.c_void_p(keys.data_ptr()), vcap, tcap, C.byref(nv), C.byref(nt),
                                        _stream_ptr(stream)), "tsdf_marching_cubes")
            if nv.value <= vcap and nt.value <= tcap:
                break
            vcap, tcap = int(nv.value), int(nt.value)
        k, m = int(nv.value), int(nt.value)
        return {"vertices": verts[:k].cpu().numpy(), "faces": faces[:m].long().cpu().numpy(),
                "normals": nrm[:k].cpu().numpy(), "colors": cols[:k].cpu().numpy() if cols is not None else None,
                "vertex_keys": keys[:k].cpu().numpy(), "color_scale": 1.0}

    def sdf_tensor(self) -> torch.Tensor:
        """The stored sdf values as a device float32 tensor [n_bricks, 512]
        (row = brick id, voxel x + 8y + 64z): the voxel parameters of the
        differentiable render (superscanner_amd.diff_render)."""
        return self.export_bricks()["sdf"]

    def export_bricks(self, stream=None, layout: str = "storage") -> dict:
        """The map as device tensors: coords [n, 3] int32 block indices, sdf and
        w [n, 512], rgb [n, 512, 3] (colour grids).  layout "storage": voxel
        x + 8y + 64z (GPUVoxelGrid.h:28-31); "open3d": voxel 64x + 8y + z, the
        index Open3D's UniformTSDFVolume::IndexOf gives the same voxel in the
        volume marchingCubesHost builds (GPUVoxelGrid.cpp:226-241), i.e. the
        wire format of a unit."""
        if layout not in ("storage", "open3d"):
            raise ValueError(f"unknown layout {layout!r} (storage or open3d)")
        n = self.size()
        dev = self.device
        coords = torch.empty((max(n, 1), 3), dtype=torch.int32, device=dev)
        sdf = torch.empty((max(n, 1), 512), dtype=torch.float32, device=dev)
        w = torch.empty((max(n, 1), 512), dtype=torch.float32, device=dev)
        rgb = torch.empty((max(n, 1), 512, 3), dtype=torch.float32, device=dev) if self.withColor else None
        nout = C.c_uint64(0)
        check(self._L.tsdf_export_bricks(self._h, C.c_void_p(coords.data_ptr()), C.c_void_p(sdf.data_ptr()),
                                         C.c_void_p(w.data_ptr()),
                                         C.c_void_p(rgb.data_ptr()) if rgb is not None else None,
                                         n, C.byref(nout), _stream_ptr(stream)), "tsdf_export_bricks")
        m = min(int(nout.value), n)
        out = {"coords": coords[:m], "sdf": sdf[:m], "w": w[:m]}
        if rgb is not None:
            out["rgb"] = rgb[:m]
        if layout == "open3d":
            out = {k: (v if k == "coords" else storage_to_open3d(v)) for k, v in out.items()}
        return out


def storage_to_open3d(a):
    """Per-brick voxel arrays [n, 512, ...] from the storage index
    x + 8y + 64z to Open3D's 64x + 8y + z (a torch tensor or numpy array;
    trailing dimensions, e.g. the three colour channels, are kept)."""
    n = a.shape[0]
    rest = tuple(a.shape[2:])
    v = a.reshape((n, 8, 8, 8) + rest)  # [n, z, y, x, ...]
    if isinstance(v, torch.Tensor):
        v = v.transpose(1, 3).contiguous()
    else:
        v = np.ascontiguousarray(np.swapaxes(v, 1, 3))
    return v.reshape((n, 512) + rest)


def pose_prepare(E):
    """(E^-1, Ez) exactly as every kernel receives them (host helper)."""
    L = _lib.load()
    a = np.ascontiguousarray(np.asarray(E, np.float32).reshape(16))
    Ei = np.zeros(16, np.float32)
    Ez = np.zeros(3, np.float32)
    check(L.tsdf_pose_prepare(a.ctypes.data_as(C.POINTER(C.c_float)),
                              Ei.ctypes.data_as(C.POINTER(C.c_float)),
                              Ez.ctypes.data_as(C.POINTER(C.c_float))), "tsdf_pose_prepare")
    return Ei.reshape(4, 4), Ez
