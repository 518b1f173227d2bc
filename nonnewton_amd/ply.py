"""PLY point-cloud I/O for scene ingestion and per-step export.

Replaces the reference's plyfile use: ``read_ply_particles`` (SPH/main.py:121-124),
``read_ply_particles_with_user_data`` (:127-130) and ``export_ply_particles``
(:184-190, binary vertex x/y/z as f4).  Reads ascii and binary little/big endian
files with a ``vertex`` element; other elements are skipped.
"""
from __future__ import annotations

import os
from typing import Dict, List, Tuple

import numpy as np

_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1",
    "short": "i2", "int16": "i2", "ushort": "u2", "uint16": "u2",
    "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


def _parse_header(f) -> Tuple[str, List[Tuple[str, int, List[Tuple[str, str, str]]]]]:
    magic = f.readline().strip()
    if magic != b"ply":
        raise ValueError("not a PLY file")
    fmt = None
    elements = []
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header has no end_header")
        tok = line.decode("ascii", "replace").split()
        if not tok:
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            elements.append((tok[1], int(tok[2]), []))
        elif tok[0] == "property":
            if tok[1] == "list":
                elements[-1][2].append((tok[4], "list", tok[2] + ":" + tok[3]))
            else:
                elements[-1][2].append((tok[2], _TYPES[tok[1]], ""))
        elif tok[0] == "end_header":
            break
    return fmt, elements


def read_vertices(path: str) -> Dict[str, np.ndarray]:
    """All scalar vertex properties as arrays (name -> array)."""
    with open(path, "rb") as f:
        fmt, elements = _parse_header(f)
        out = {}
        if fmt == "ascii":
            rest = f.read().decode("ascii", "replace").split("\n")
            pos = 0
            for name, count, props in elements:
                rows = []
                for _ in range(count):
                    while pos < len(rest) and not rest[pos].strip():
                        pos += 1
                    rows.append(rest[pos].split())
                    pos += 1
                if name == "vertex":
                    for k, (pname, ptype, _) in enumerate(props):
                        if ptype == "list":
                            raise ValueError("list properties on vertices are not supported")
                        out[pname] = np.array([float(r[k]) for r in rows], dtype=ptype)
                    return out
            raise ValueError("no vertex element")
        endian = "<" if fmt == "binary_little_endian" else ">"
        for name, count, props in elements:
            if any(ptype == "list" for _, ptype, _ in props):
                if name == "vertex":
                    raise ValueError("list properties on vertices are not supported")
                if count:
                    raise ValueError("binary list elements before the vertices are not supported")
                continue
            dt = np.dtype([(pname, endian + ptype) for pname, ptype, _ in props])
            data = np.frombuffer(f.read(dt.itemsize * count), dtype=dt, count=count)
            if name == "vertex":
                return {k: np.array(data[k]) for k in data.dtype.names}
        raise ValueError("no vertex element")


def read_ply_particles(path: str) -> np.ndarray:
    """(N, 3) float32 vertex positions (read_ply_particles, SPH/main.py:121-124)."""
    v = read_vertices(path)
    return np.stack([v["x"], v["y"], v["z"]], axis=1).astype(np.float32)


def export_ply_particles(particles: np.ndarray, filename: str) -> None:
    """Binary little-endian PLY with vertex x, y, z as f4 (export_ply_particles, :184-190)."""
    p = np.ascontiguousarray(particles, dtype="<f4").reshape(-1, 3)
    header = (
        "ply\nformat binary_little_endian 1.0\n"
        f"element vertex {p.shape[0]}\nproperty float x\nproperty float y\nproperty float z\nend_header\n"
    )
    d = os.path.dirname(filename)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(filename, "wb") as f:
        f.write(header.encode("ascii"))
        f.write(p.tobytes())


def export_ply_mesh(vertices: np.ndarray, faces: np.ndarray, filename: str, normals=None, velocity=None) -> None:
    """Binary little-endian PLY mesh: vertex x, y, z (and nx, ny, nz; vx, vy, vz) as f4, faces
    as `list uchar int vertex_indices` (SRTool.export_mesh, SPH/surf_recon/srtool.py:1802-1838)."""
    v = np.ascontiguousarray(vertices, dtype="<f4").reshape(-1, 3)
    f = np.ascontiguousarray(faces, dtype="<i4").reshape(-1, 3)
    props = "property float x\nproperty float y\nproperty float z\n"
    if normals is not None:
        n = np.ascontiguousarray(normals, dtype="<f4").reshape(-1, 3)
        if n.shape != v.shape:
            raise ValueError("normals must have one entry per vertex")
        v = np.concatenate([v, n], axis=1)
        props += "property float nx\nproperty float ny\nproperty float nz\n"
    if velocity is not None:
        u = np.ascontiguousarray(velocity, dtype="<f4").reshape(-1, 3)
        if u.shape[0] != v.shape[0]:
            raise ValueError("velocity must have one entry per vertex")
        v = np.concatenate([v, u], axis=1)
        props += "property float vx\nproperty float vy\nproperty float vz\n"
    header = (
        f"ply\nformat binary_little_endian 1.0\nelement vertex {v.shape[0]}\n{props}"
        f"element face {f.shape[0]}\nproperty list uchar int vertex_indices\nend_header\n"
    )
    rec = np.empty(f.shape[0], dtype=[("n", "u1"), ("i", "<i4", (3,))])
    rec["n"] = 3
    rec["i"] = f
    d = os.path.dirname(filename)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(filename, "wb") as out:
        out.write(header.encode("ascii"))
        out.write(np.ascontiguousarray(v, dtype="<f4").tobytes())
        out.write(rec.tobytes())
