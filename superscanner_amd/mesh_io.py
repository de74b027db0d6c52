"""Mesh output: binary little-endian PLY of the dict that
``GPUVoxelGrid.extract_mesh`` returns (the reference writes its Open3D mesh
with open3d::io::WriteTriangleMesh; PLY is the format it is saved in)."""
from __future__ import annotations

import numpy as np


def write_ply(path, mesh: dict) -> None:
    """Vertices (float32 x y z), optional unit normals (nx ny nz), optional
    colours (uint8 red green blue: colours in 0..color_scale, 255 unless the
    mesh says otherwise, scaled to 0..255, rounded and clamped) and triangular
    faces (uchar count, int32 indices)."""
    v = np.ascontiguousarray(mesh["vertices"], np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(mesh["faces"]).reshape(-1, 3)
    nrm = mesh.get("normals")
    col = mesh.get("colors")
    fields = [("x", "<f4"), ("y", "<f4"), ("z", "<f4")]
    if nrm is not None:
        fields += [("nx", "<f4"), ("ny", "<f4"), ("nz", "<f4")]
    if col is not None:
        fields += [("red", "u1"), ("green", "u1"), ("blue", "u1")]
    rec = np.zeros(len(v), dtype=fields)
    rec["x"], rec["y"], rec["z"] = v[:, 0], v[:, 1], v[:, 2]
    if nrm is not None:
        nrm = np.asarray(nrm, np.float32).reshape(-1, 3)
        rec["nx"], rec["ny"], rec["nz"] = nrm[:, 0], nrm[:, 1], nrm[:, 2]
    if col is not None:
        scale = 255.0 / float(mesh.get("color_scale", 255.0))
        c = np.clip(np.rint(np.asarray(col, np.float64).reshape(-1, 3) * scale), 0, 255).astype(np.uint8)
        rec["red"], rec["green"], rec["blue"] = c[:, 0], c[:, 1], c[:, 2]
    frec = np.zeros(len(f), dtype=[("n", "u1"), ("i", "<i4", (3,))])
    frec["n"] = 3
    frec["i"] = f.astype(np.int32)
    head = ["ply", "format binary_little_endian 1.0", f"element vertex {len(v)}"]
    names = {"<f4": "float", "u1": "uchar"}
    head += [f"property {names[t]} {n}" for n, t in fields]
    head += [f"element face {len(f)}", "property list uchar int vertex_indices", "end_header"]
    with open(path, "wb") as fh:
        fh.write(("\n".join(head) + "\n").encode("ascii"))
        fh.write(rec.tobytes())
        fh.write(frec.tobytes())


FLIP = np.diag([1.0, -1.0, -1.0, 1.0])  # getflip, matrixutil.cpp:27-35


def save_cpu_mesh(path, mesh: dict, flip: bool = True) -> None:
    """Model::saveCPUMesh, Model.cpp:205-228: a copy of the mesh whose colour
    channels outside the valid range [0, color_scale] are set to 0 (the
    reference tests its [0, 1] Open3D colours, the marching-cubes colours;
    marching-tetrahedra colours are 0..255), transformed by getflip
    (vertices and normals; a rotation, so the winding is kept), written as
    binary PLY with normals and colours."""
    m = dict(mesh)
    v = np.asarray(m["vertices"], np.float64).reshape(-1, 3)
    nrm = m.get("normals")
    if flip:
        v = v @ FLIP[:3, :3].T
        if nrm is not None:
            nrm = np.asarray(nrm, np.float64).reshape(-1, 3) @ FLIP[:3, :3].T
    col = m.get("colors")
    if col is not None:
        col = np.asarray(col, np.float64).reshape(-1, 3).copy()
        col[(col > float(m.get("color_scale", 255.0))) | (col < 0.0)] = 0.0
    m["vertices"] = v.astype(np.float32)
    m["normals"] = nrm.astype(np.float32) if nrm is not None else None
    m["colors"] = col
    write_ply(path, m)


def read_ply(path) -> dict:
    """Reads what write_ply wrote (binary little-endian, the same properties)."""
    with open(path, "rb") as fh:
        data = fh.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    lines = data[:end].decode("ascii").splitlines()
    nv = nf = 0
    fields = []
    types = {"float": "<f4", "uchar": "u1"}
    for ln in lines:
        w = ln.split()
        if w[:2] == ["element", "vertex"]:
            nv = int(w[2])
        elif w[:2] == ["element", "face"]:
            nf = int(w[2])
        elif w[0] == "property" and w[1] in types:
            fields.append((w[2], types[w[1]]))
    rec = np.frombuffer(data, dtype=fields, count=nv, offset=end)
    frec = np.frombuffer(data, dtype=[("n", "u1"), ("i", "<i4", (3,))], count=nf,
                         offset=end + rec.nbytes)
    out = {"vertices": np.stack([rec["x"], rec["y"], rec["z"]], 1), "faces": frec["i"].astype(np.int64)}
    names = [n for n, _ in fields]
    out["normals"] = np.stack([rec["nx"], rec["ny"], rec["nz"]], 1) if "nx" in names else None
    out["colors"] = np.stack([rec["red"], rec["green"], rec["blue"]], 1) if "red" in names else None
    return out
