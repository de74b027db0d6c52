"""The C ABI library loads and exports every entry point include/tsdf.h
declares; its host-only helpers agree with the oracle.  No device work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from superscanner_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "tsdf.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tsdf_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for name in declared_functions():
        assert hasattr(L, name), name
    syms = os.popen(f"nm -D --defined-only {_lib.LIB_PATH}").read()
    for name in declared_functions():
        assert re.search(rf"\bT {name}\b", syms), name


def test_library_is_built_for_gfx950():
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_default_params_match_reference_defaults():
    L = _lib.load()
    p = _lib.TsdfParams()
    L.tsdf_default_params(C.byref(p))
    assert p.max_ray_length == 10.0           # d_max, GPUVoxelGrid.h:55
    assert abs(p.min_depth - 0.2) < 1e-7      # rmin, GPUVoxelGrid.h:56
    assert p.max_weight == 0.0                # no cap in the reference
    assert p.storage == _lib.STORAGE_F32


@pytest.mark.parametrize("field,value", [
    ("voxel_length", 0.0), ("truncation", 0.0), ("truncation", 10.0), ("n_buckets", 1000),
    ("n_bricks", 0), ("storage", 7), ("height", 0), ("max_ray_length", -1.0),
])
def test_create_rejects_bad_params_without_touching_the_device(field, value):
    L = _lib.load()
    p = _lib.TsdfParams()
    L.tsdf_default_params(C.byref(p))
    setattr(p, field, value)
    h = C.c_void_p()
    st = L.tsdf_create(C.byref(p), C.byref(h))
    assert st in (_lib.TSDF_BAD_ARG, _lib.TSDF_BAD_DIMS)
    assert not h.value
    assert L.tsdf_last_error()


def test_null_handles_are_errors():
    L = _lib.load()
    E = (C.c_float * 16)()
    assert L.tsdf_integrate(None, None, 0, None, E, 1, None) == _lib.TSDF_BAD_ARG
    assert L.tsdf_raycast(None, E, None, None, None, None, None) == _lib.TSDF_BAD_ARG
    assert L.tsdf_status_string(_lib.TSDF_MAP_FULL)


def test_pose_prepare_bit_identical_to_oracle():
    from test_oracle_kat import random_rigid
    from superscanner_amd.voxel_grid import pose_prepare
    rng = np.random.default_rng(11)
    for _ in range(300):
        E = random_rigid(rng, t_scale=20.0)
        Ei_g, Ez_g = pose_prepare(E)
        Ei_o, Ez_o = oracle.pose_prepare(E)
        assert (Ei_g.view(np.uint32) == Ei_o.view(np.uint32)).all()
        assert (Ez_g.view(np.uint32) == Ez_o.view(np.uint32)).all()
