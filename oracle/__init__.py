"""Python binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.

The C restatement (tsdf_oracle.c) follows src/VoxelGrid/GPUVoxelGrid.{h,cpp}
of the reference line by line; see the citations there.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

UB_VOXEL_INDEX = 1
UB_MISSING_OUT = 2


class OrcCounts(C.Structure):
    _fields_ = [("n_touched", C.c_int64), ("n_new", C.c_int64), ("n_upd", C.c_int64),
                ("n_valid", C.c_int64)]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class OrcRayCounts(C.Structure):
    _fields_ = [("n_hit", C.c_int64), ("n_ub", C.c_int64), ("samples", C.c_int64),
                ("probes", C.c_int64), ("u_vox", C.c_int64), ("u_blk", C.c_int64)]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


_lib = None


def build() -> None:
    import subprocess
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    fp, ip, vp = P(C.c_float), P(C.c_int), C.c_void_p
    L.orc_prime_hash.argtypes = [C.c_int, C.c_int, C.c_int]
    L.orc_prime_hash.restype = C.c_uint64
    L.orc_f2i.argtypes = [C.c_float]
    L.orc_f2i.restype = C.c_int
    L.orc_vb_index.argtypes = [C.c_float, C.c_float]
    L.orc_vb_index.restype = C.c_int
    L.orc_vin_vb.argtypes = [fp, ip, C.c_float, C.c_float, ip]
    L.orc_vin_vb.restype = None
    L.orc_pose_prepare.argtypes = [fp, fp, fp]
    L.orc_pose_prepare.restype = None
    L.orc_pose_is_rigid.argtypes = [fp]
    L.orc_pose_is_rigid.restype = C.c_int
    L.orc_create.argtypes = [C.c_float, C.c_float, C.c_int, C.c_int, C.c_int64, C.c_int,
                             C.c_float, C.c_float, C.c_float]
    L.orc_create.restype = vp
    L.orc_destroy.argtypes = [vp]
    L.orc_destroy.restype = None
    L.orc_set_intrinsics.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float]
    L.orc_set_intrinsics.restype = None
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set_threads.restype = None
    L.orc_integrate.argtypes = [vp, vp, C.c_int, vp, fp, C.c_int, P(OrcCounts)]
    L.orc_integrate.restype = C.c_int
    L.orc_raycast.argtypes = [vp, fp, vp, vp, vp, vp, C.c_int, P(OrcRayCounts)]
    L.orc_raycast.restype = None
    L.orc_n_bricks.argtypes = [vp]
    L.orc_n_bricks.restype = C.c_int64
    L.orc_export.argtypes = [vp, vp, vp, vp, vp, vp]
    L.orc_export.restype = None
    L.orc_clean.argtypes = [vp]
    L.orc_clean.restype = C.c_int64
    L.orc_import_rgb.argtypes = [vp, vp]
    L.orc_import_rgb.restype = None
    L.orc_set_mutation.argtypes = [C.c_int]
    L.orc_set_mutation.restype = None
    dp = P(C.c_double)
    d = C.c_double
    L.orc_field_create.argtypes = [C.c_int64, vp, vp, vp, C.c_float, C.c_float, C.c_float]
    L.orc_field_create.restype = vp
    L.orc_field_destroy.argtypes = [vp]
    L.orc_field_destroy.restype = None
    L.orc_field_set_sdf.argtypes = [vp, vp]
    L.orc_field_set_sdf.restype = None
    L.orc_render.argtypes = [vp, dp, C.c_int, C.c_int, d, d, d, d, vp, vp]
    L.orc_render.restype = None
    L.orc_render_grad.argtypes = [vp, dp, C.c_int, C.c_int, d, d, d, d, vp, vp, vp, vp, vp]
    L.orc_render_grad.restype = None
    L.orc64_create.argtypes = [d, d, C.c_int, C.c_int, d, d, d, d, dp, dp, d, d]
    L.orc64_create.restype = vp
    L.orc64_destroy.argtypes = [vp]
    L.orc64_destroy.restype = None
    L.orc64_integrate.argtypes = [vp, vp, C.c_int, dp, C.c_int]
    L.orc64_integrate.restype = None
    L.orc64_n_bricks.argtypes = [vp]
    L.orc64_n_bricks.restype = C.c_int64
    L.orc64_export.argtypes = [vp, vp, vp, vp]
    L.orc64_export.restype = None
    L.orc_depth_to_points.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, vp]
    L.orc_depth_to_points.restype = None
    L.orc_normals.argtypes = [vp, C.c_int, C.c_int, vp]
    L.orc_normals.restype = None
    L.orc_icp_associate.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                    C.c_int, C.c_int, C.c_float, vp]
    L.orc_icp_associate.restype = None
    L.orc_icp_sums.argtypes = [vp, vp, vp, C.c_int, C.c_int, vp]
    L.orc_icp_sums.restype = None
    _lib = L
    return L


def _f32(a, n):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))
    assert a.size == n
    return a


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def prime_hash(x: int, y: int, z: int) -> int:
    return int(lib().orc_prime_hash(x, y, z))


def f2i(f: float) -> int:
    return int(lib().orc_f2i(f))


def vb_index(a: float, vbl: float) -> int:
    return int(lib().orc_vb_index(a, vbl))


def vin_vb(p, vb, vbl: float, vl: float):
    pa = _f32(p, 3)
    ba = np.ascontiguousarray(np.asarray(vb, dtype=np.int32))
    out = np.zeros(3, dtype=np.int32)
    lib().orc_vin_vb(_fptr(pa), ba.ctypes.data_as(C.POINTER(C.c_int)), vbl, vl,
                     out.ctypes.data_as(C.POINTER(C.c_int)))
    return out


def pose_prepare(E):
    Ea = _f32(E, 16)
    Ei = np.zeros(16, np.float32)
    Ez = np.zeros(3, np.float32)
    lib().orc_pose_prepare(_fptr(Ea), _fptr(Ei), _fptr(Ez))
    return Ei.reshape(4, 4), Ez


def pose_is_rigid(E) -> bool:
    return bool(lib().orc_pose_is_rigid(_fptr(_f32(E, 16))))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def set_mutation(m: int) -> None:
    """Plants a deliberate mistake in the oracle (0 = none); used only by the
    tests that show the closed-form checks catch such mistakes."""
    lib().orc_set_mutation(int(m))


class OracleGrid:
    """CPU restatement of GPUVoxelGrid (GPUVoxelGrid.h:40-120)."""

    def __init__(self, voxel_length, truncation, height, width, max_bricks=1 << 22,
                 with_color=False, max_weight=0.0, min_depth=0.2, max_ray_length=10.0):
        self.H, self.W = int(height), int(width)
        self.with_color = bool(with_color)
        self.vl = float(voxel_length)
        self._h = lib().orc_create(voxel_length, truncation, self.H, self.W, int(max_bricks),
                                   int(with_color), max_weight, min_depth, max_ray_length)

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_intrinsics(self, fx, fy, cx, cy):
        lib().orc_set_intrinsics(self._h, fx, fy, cx, cy)

    def integrate(self, depth, E, sign=1, rgb=None):
        d = np.ascontiguousarray(depth, dtype=np.uint16)
        assert d.shape == (self.H, self.W)
        c = None
        if rgb is not None:
            c = np.ascontiguousarray(rgb, dtype=np.uint8)
            assert c.shape == (self.H, self.W, 3)
        cnt = OrcCounts()
        st = lib().orc_integrate(self._h, d.ctypes.data, self.W,
                                 c.ctypes.data if c is not None else None,
                                 _fptr(_f32(E, 16)), int(sign), C.byref(cnt))
        return st, cnt.as_dict()

    def deintegrate(self, depth, E, rgb=None):
        return self.integrate(depth, E, -1, rgb)

    def raycast(self, E, stats=False):
        H, W = self.H, self.W
        pts = np.zeros((H, W, 3), np.float32)
        cols = np.zeros((H, W, 3), np.float32)
        dep = np.zeros((H, W), np.float32)
        ub = np.zeros((H, W), np.uint8)
        cnt = OrcRayCounts()
        lib().orc_raycast(self._h, _fptr(_f32(E, 16)), pts.ctypes.data, cols.ctypes.data,
                          dep.ctypes.data, ub.ctypes.data, int(bool(stats)), C.byref(cnt))
        return {"points": pts, "colors": cols, "depth": dep, "ub": ub, "counts": cnt.as_dict()}

    def n_bricks(self) -> int:
        return int(lib().orc_n_bricks(self._h))

    def clean(self) -> int:
        """cleanMemory: erases all-zero-weight bricks; returns the count."""
        return int(lib().orc_clean(self._h))

    def import_rgb(self, rgb):
        """Overwrites every brick's colours ([n, 512, 3] in export order);
        a test hook that gives voxels known colours."""
        assert self.with_color
        a = np.ascontiguousarray(rgb, dtype=np.float32)
        assert a.shape == (self.n_bricks(), 512, 3)
        lib().orc_import_rgb(self._h, a.ctypes.data)

    def export(self):
        n = self.n_bricks()
        coords = np.zeros((n, 3), np.int32)
        keys = np.zeros(n, np.int64)
        sdf = np.zeros((n, 512), np.float32)
        w = np.zeros((n, 512), np.float32)
        rgb = np.zeros((n, 512, 3), np.float32) if self.with_color else None
        lib().orc_export(self._h, coords.ctypes.data, keys.ctypes.data, sdf.ctypes.data,
                         w.ctypes.data, rgb.ctypes.data if rgb is not None else None)
        return {"coords": coords, "keys": keys, "sdf": sdf, "w": w, "rgb": rgb}


class OracleField:
    """fp64 differentiable trilinear render of exported bricks (definition in
    oracle/render_oracle.c; no reference counterpart)."""

    def __init__(self, coords, sdf, w, voxel_length, min_depth=0.2, max_ray_length=10.0):
        self.coords = np.ascontiguousarray(coords, dtype=np.int32).reshape(-1, 3)
        self.n = self.coords.shape[0]
        sdf = np.ascontiguousarray(sdf, dtype=np.float32).reshape(self.n, 512)
        w = np.ascontiguousarray(w, dtype=np.float32).reshape(self.n, 512)
        self._h = lib().orc_field_create(self.n, self.coords.ctypes.data, sdf.ctypes.data, w.ctypes.data,
                                         voxel_length, min_depth, max_ray_length)

    def close(self):
        if self._h:
            lib().orc_field_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sdf(self, sdf):
        a = np.ascontiguousarray(sdf, dtype=np.float32).reshape(self.n, 512)
        lib().orc_field_set_sdf(self._h, a.ctypes.data)

    @staticmethod
    def _pose(E):
        return np.ascontiguousarray(np.asarray(E, dtype=np.float64).reshape(16))

    def render(self, E, h, w, intr, normals=False):
        """(depth [h, w] float64, normals [h, w, 3] or None); 0 = no surface."""
        Ed = self._pose(E)
        dep = np.zeros((h, w), np.float64)
        nrm = np.zeros((h, w, 3), np.float64) if normals else None
        lib().orc_render(self._h, Ed.ctypes.data_as(C.POINTER(C.c_double)), h, w, *map(float, intr),
                         dep.ctypes.data, nrm.ctypes.data if nrm is not None else None)
        return dep, nrm

    def render_grad(self, E, h, w, intr, depth, grad_depth=None, want_sdf=False):
        """(jac [h, w, 6] = dD/dxi, grad_pose [6], grad_sdf [n, 512] or None)."""
        Ed = self._pose(E)
        dep = np.ascontiguousarray(depth, dtype=np.float64).reshape(h, w)
        gd = None if grad_depth is None else np.ascontiguousarray(grad_depth, dtype=np.float64).reshape(h, w)
        jac = np.zeros((h, w, 6), np.float64)
        gp = np.zeros(6, np.float64)
        gs = np.zeros((self.n, 512), np.float64) if want_sdf else None
        lib().orc_render_grad(self._h, Ed.ctypes.data_as(C.POINTER(C.c_double)), h, w, *map(float, intr),
                              dep.ctypes.data, gd.ctypes.data if gd is not None else None, jac.ctypes.data,
                              gp.ctypes.data, gs.ctypes.data if gs is not None else None)
        return jac, gp, gs


class Oracle64Grid:
    """The integrate in double precision on the blocks meeting a crop box
    (oracle/fp64_oracle.c): the fp64 reference for the error of the fp32 /
    fp16 GPU grids."""

    def __init__(self, voxel_length, truncation, height, width, intr, crop_min, crop_max,
                 max_weight=0.0, min_depth=0.2):
        self.H, self.W = int(height), int(width)
        lo = np.ascontiguousarray(crop_min, np.float64)
        hi = np.ascontiguousarray(crop_max, np.float64)
        dp = C.POINTER(C.c_double)
        self._h = lib().orc64_create(float(voxel_length), float(truncation), self.H, self.W,
                                     *map(float, intr), lo.ctypes.data_as(dp), hi.ctypes.data_as(dp),
                                     float(max_weight), float(min_depth))

    def close(self):
        if self._h:
            lib().orc64_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def integrate(self, depth, E, sign=1):
        d = np.ascontiguousarray(depth, dtype=np.uint16)
        assert d.shape == (self.H, self.W)
        Ed = np.ascontiguousarray(np.asarray(E, np.float64).reshape(16))
        lib().orc64_integrate(self._h, d.ctypes.data, self.W, Ed.ctypes.data_as(C.POINTER(C.c_double)), int(sign))

    def export(self):
        n = int(lib().orc64_n_bricks(self._h))
        coords = np.zeros((n, 3), np.int32)
        sdf = np.zeros((n, 512), np.float64)
        w = np.zeros((n, 512), np.float64)
        if n:
            lib().orc64_export(self._h, coords.ctypes.data, sdf.ctypes.data, w.ctypes.data)
        return {"coords": coords, "sdf": sdf, "w": w}
