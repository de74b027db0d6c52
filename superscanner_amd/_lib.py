"""ctypes binding of libtsdf_hip.so (the C ABI declared in include/tsdf.h).

The library is built in-tree by ``make -C superscanner_amd``.  There is no
fallback: if the shared object is missing, loading raises instead of running
anything on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# TSDF_HIP_LIB names another build of the same sources (``make -C
# superscanner_amd variant``) for timing a compile-time alternative; unset,
# the in-tree library is loaded.
LIB_PATH = os.environ.get("TSDF_HIP_LIB") or os.path.join(_HERE, "libtsdf_hip.so")

TSDF_OK = 0
TSDF_BAD_ARG = 1
TSDF_BAD_DIMS = 2
TSDF_MAP_FULL = 3
TSDF_HIP_ERROR = 4
TSDF_NO_INTRINSICS = 5
TSDF_BAD_POSE = 6

STORAGE_F32 = 0
STORAGE_F16 = 1

RAY_UB_VOXEL_INDEX = 1
RAY_UB_MISSING_OUT = 2

# Every symbol include/tsdf.h declares (checked by tests/test_capi_exports.py).
EXPORTS = (
    "tsdf_default_params", "tsdf_create", "tsdf_get_params", "tsdf_destroy", "tsdf_reset",
    "tsdf_set_intrinsics", "tsdf_integrate", "tsdf_integrate_batch",
    "tsdf_reintegrate_batch", "tsdf_reserve_batch", "tsdf_raycast", "tsdf_sync", "tsdf_stats",
    "tsdf_export_bricks", "tsdf_pose_prepare", "tsdf_status_string",
    "tsdf_last_error", "tsdf_raycast_footprint", "tsdf_profile", "tsdf_profile_read",
    "tsdf_default_tuning", "tsdf_get_tuning", "tsdf_set_tuning",
    "tsdf_render", "tsdf_render_backward", "tsdf_track_normal_eq", "tsdf_se3_gauss_newton_step",
    "tsdf_clean", "tsdf_extract_mesh", "tsdf_marching_cubes",
    "tsdf_icp_default_params", "tsdf_depth_to_points", "tsdf_points_to_normals", "tsdf_icp_associate",
    "tsdf_icp_moments", "tsdf_icp_moments_noalloc", "tsdf_icp", "tsdf_icp_release",
    "tsdf_mat_to_dof", "tsdf_dof_to_mat", "tsdf_reint_diff", "tsdf_reint_select",
)

ICP_POINT_TO_POINT = 0
ICP_POINT_TO_PLANE = 1
ICP_MOMENTS = 18

KERNEL_KINDS = ("reset", "collect", "integrate", "collect_batch", "integrate_batch", "raycast")


class TsdfIcpParams(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("n_iterations", C.c_int32),
        ("n_search", C.c_int32),
        ("max_dist", C.c_float),
    ]


class TsdfParams(C.Structure):
    _fields_ = [
        ("voxel_length", C.c_float),
        ("truncation", C.c_float),
        ("height", C.c_int32),
        ("width", C.c_int32),
        ("n_buckets", C.c_uint32),
        ("n_bricks", C.c_uint32),
        ("storage", C.c_int32),
        ("with_color", C.c_int32),
        ("max_weight", C.c_float),
        ("min_depth", C.c_float),
        ("max_ray_length", C.c_float),
        ("device", C.c_int32),
    ]


class TsdfStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "n_bricks", "capacity", "overflow", "last_touched", "total_touched",
        "total_new", "total_updated", "total_valid_px", "ray_hits", "ray_ub",
        "ray_samples", "ray_probes")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class TsdfTuning(C.Structure):
    _fields_ = [("raycast_xcd_map", C.c_int32), ("raycast_skip", C.c_int32),
                ("integrate_wg_per_cu", C.c_int32), ("batch_ops_per_pass", C.c_int32),
                ("batch_graph", C.c_int32), ("raycast_order", C.c_int32), ("raycast_budget", C.c_int32)]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class TsdfKernelTimes(C.Structure):
    _fields_ = [("ms", C.c_double * 6), ("launches", C.c_uint64 * 6)]

    def as_dict(self):
        return {k: {"ms": float(self.ms[i]), "launches": int(self.launches[i])}
                for i, k in enumerate(KERNEL_KINDS)}


class TsdfView(C.Structure):
    _fields_ = [("height", C.c_int32), ("width", C.c_int32), ("fx", C.c_float), ("fy", C.c_float),
                ("cx", C.c_float), ("cy", C.c_float)]


class TsdfError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str):
        self.status = status
        super().__init__(f"{where}: status {status}: {detail}")


_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Loads and types the library once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `make -C {_HERE}` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    # A process must hold one HIP runtime.  torch ships its own; loaded
    # first, it is the one the library's HIP dependency resolves to, whereas
    # loading the library first binds the system runtime and a later torch
    # brings a second (hipSetDevice then fails in one of them).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(path)
    P = C.POINTER
    vp = C.c_void_p
    fp = P(C.c_float)
    L.tsdf_default_params.argtypes = [P(TsdfParams)]
    L.tsdf_default_params.restype = None
    L.tsdf_create.argtypes = [P(TsdfParams), P(vp)]
    L.tsdf_get_params.argtypes = [vp, P(TsdfParams)]
    L.tsdf_destroy.argtypes = [vp]
    L.tsdf_destroy.restype = None
    L.tsdf_reset.argtypes = [vp, vp]
    L.tsdf_set_intrinsics.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float]
    L.tsdf_integrate.argtypes = [vp, vp, C.c_int32, vp, fp, C.c_int32, vp]
    L.tsdf_integrate_batch.argtypes = [vp, C.c_int32, P(vp), C.c_int32, P(vp), fp, P(C.c_int32), vp]
    L.tsdf_reintegrate_batch.argtypes = [vp, C.c_int32, P(vp), C.c_int32, P(vp), fp, fp, vp]
    L.tsdf_reserve_batch.argtypes = [vp, C.c_int32]
    dp = P(C.c_double)
    L.tsdf_render.argtypes = [vp, C.c_int32, fp, P(TsdfView), vp, vp, vp, vp]
    L.tsdf_render_backward.argtypes = [vp, C.c_int32, fp, P(TsdfView), vp, vp, vp, vp, vp, vp]
    L.tsdf_track_normal_eq.argtypes = [vp, C.c_int32, fp, P(TsdfView), vp, C.c_float, vp, vp, vp]
    L.tsdf_se3_gauss_newton_step.argtypes = [dp, dp, dp, dp, dp]
    L.tsdf_clean.argtypes = [vp, P(C.c_uint64), vp]
    L.tsdf_extract_mesh.argtypes = [vp, vp, vp, vp, C.c_uint64, P(C.c_uint64), vp]
    L.tsdf_marching_cubes.argtypes = [vp, vp, vp, vp, vp, vp, C.c_uint64, C.c_uint64, P(C.c_uint64), P(C.c_uint64),
                                      vp]
    L.tsdf_icp_default_params.argtypes = [P(TsdfIcpParams)]
    L.tsdf_icp_default_params.restype = None
    L.tsdf_depth_to_points.argtypes = [vp, P(TsdfView), vp, vp]
    L.tsdf_points_to_normals.argtypes = [vp, C.c_int32, C.c_int32, vp, vp]
    L.tsdf_icp_associate.argtypes = [vp, vp, vp, fp, P(TsdfView), P(TsdfIcpParams), vp, vp]
    L.tsdf_icp_moments.argtypes = [vp, vp, dp, C.c_int32, C.c_int32, dp, vp]
    L.tsdf_icp_moments_noalloc.argtypes = [vp, vp, dp, C.c_int32, C.c_int32, dp, vp]
    L.tsdf_icp.argtypes = [vp, vp, dp, dp, P(TsdfView), P(TsdfIcpParams), dp, fp, fp, vp]
    L.tsdf_icp_release.argtypes = []
    L.tsdf_mat_to_dof.argtypes = [dp, dp]
    L.tsdf_dof_to_mat.argtypes = [dp, dp]
    L.tsdf_reint_diff.argtypes = [dp, dp]
    L.tsdf_reint_diff.restype = C.c_double
    L.tsdf_reint_select.argtypes = [C.c_int32, dp, dp, C.c_double, P(C.c_int32), P(C.c_int32)]
    L.tsdf_raycast.argtypes = [vp, fp, vp, vp, vp, vp, vp]
    L.tsdf_sync.argtypes = [vp, vp]
    L.tsdf_stats.argtypes = [vp, P(TsdfStats), vp]
    L.tsdf_export_bricks.argtypes = [vp, vp, vp, vp, vp, C.c_uint64, P(C.c_uint64), vp]
    L.tsdf_pose_prepare.argtypes = [fp, fp, fp]
    L.tsdf_raycast_footprint.argtypes = [vp, fp, P(C.c_uint64), P(C.c_uint64), vp]
    L.tsdf_profile.argtypes = [vp, C.c_int32]
    L.tsdf_default_tuning.argtypes = [P(TsdfTuning)]
    L.tsdf_default_tuning.restype = None
    L.tsdf_get_tuning.argtypes = [vp, P(TsdfTuning)]
    L.tsdf_set_tuning.argtypes = [vp, P(TsdfTuning)]
    L.tsdf_profile_read.argtypes = [vp, P(TsdfKernelTimes), C.c_int32]
    L.tsdf_status_string.argtypes = [C.c_int]
    L.tsdf_status_string.restype = C.c_char_p
    L.tsdf_last_error.argtypes = []
    L.tsdf_last_error.restype = C.c_char_p
    for name in ("tsdf_create", "tsdf_get_params", "tsdf_reset", "tsdf_set_intrinsics", "tsdf_integrate",
                 "tsdf_integrate_batch", "tsdf_reintegrate_batch", "tsdf_reserve_batch", "tsdf_raycast",
                 "tsdf_sync", "tsdf_stats", "tsdf_export_bricks", "tsdf_pose_prepare",
                 "tsdf_raycast_footprint", "tsdf_profile", "tsdf_profile_read",
                 "tsdf_get_tuning", "tsdf_set_tuning"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(status: int, where: str) -> None:
    if status != TSDF_OK:
        L = load()
        raise TsdfError(status, where, L.tsdf_last_error().decode(errors="replace"))
