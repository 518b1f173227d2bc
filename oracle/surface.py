"""ctypes wrapper of the surface-reconstruction oracle (oracle/surface_oracle.c),
the CPU restatement of SRTool.dual_contouring (SPH/surf_recon/srtool.py:1264-1304).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/ use it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

E_ARG, E_OOM, E_VOXEL_OVERFLOW, E_OUT_OF_GRID, E_RANGE = -1, -3, -5, -6, -7


def _lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle_surface.so")
    if not os.path.exists(path):
        import subprocess

        subprocess.check_call(["sh", os.path.join(_HERE, "build.sh")])
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    f32p, i32p, i64p = P(ctypes.c_float), P(ctypes.c_int32), P(ctypes.c_int64)
    lib.sro_create.restype = ctypes.c_void_p
    lib.sro_create.argtypes = [P(ctypes.c_double), i32p]
    lib.sro_destroy.argtypes = [ctypes.c_void_p]
    lib.sro_run.restype = ctypes.c_int32
    lib.sro_run.argtypes = [ctypes.c_void_p, f32p, f32p, ctypes.c_int32]
    lib.sro_get_velocity.argtypes = [ctypes.c_void_p, f32p, f32p]
    lib.sro_info.argtypes = [ctypes.c_void_p, i64p]
    lib.sro_sd.restype = ctypes.c_double
    lib.sro_sd.argtypes = [ctypes.c_void_p]
    lib.sro_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int32, f32p]
    lib.sro_get_active.argtypes = [ctypes.c_void_p, P(ctypes.c_uint8)]
    lib.sro_get_aniso.argtypes = [ctypes.c_void_p, f32p, f32p]
    lib.sro_get_mesh.argtypes = [ctypes.c_void_p, f32p, i32p, f32p]
    lib.sro_qef.argtypes = [f32p, ctypes.c_int32, f32p]
    lib.sro_expf.restype = ctypes.c_float
    lib.sro_expf.argtypes = [ctypes.c_float]
    _LIB = lib
    return lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class SurfaceError(RuntimeError):
    def __init__(self, code):
        super().__init__(f"surface oracle returned {code}")
        self.code = code


def reconstruct(pos, origin=(0.0, 0.0, 0.0), voxel_size=0.01, particle_radius=0.01, smooth_radius=0.04,
                isolevel=0.0, anisotropic=True, smooth_iterations=1, normals=False, velocity=None) -> dict:
    """Run the oracle on positions pos (n x 3) and return every stage's output:
    grid placement, active blocks, the field before and after smoothing, the
    smoothing's sd, G and x_bar per particle, and the mesh."""
    lib = _lib()
    dp = (ctypes.c_double * 7)(*origin, voxel_size, particle_radius, smooth_radius, isolevel)
    ip = (ctypes.c_int32 * 4)(int(bool(anisotropic)), int(smooth_iterations), int(bool(normals)),
                              int(velocity is not None))
    h = lib.sro_create(dp, ip)
    if not h:
        raise SurfaceError(E_ARG)
    try:
        x = np.ascontiguousarray(np.asarray(pos, dtype=np.float32).reshape(-1, 3))
        vel = None if velocity is None else np.ascontiguousarray(np.asarray(velocity, np.float32).reshape(-1, 3))
        rc = lib.sro_run(h, _fp(x), _fp(vel) if vel is not None else None, len(x))
        if rc != 0:
            raise SurfaceError(rc)
        info = (ctypes.c_int64 * 15)()
        lib.sro_info(h, info)
        lo, dims = tuple(info[0:3]), tuple(info[3:6])
        nv, nt, nact = info[6], info[7], info[8]
        field0 = np.empty(dims, np.float32)
        field = np.empty(dims, np.float32)
        lib.sro_get_field(h, 0, _fp(field0))
        lib.sro_get_field(h, 1, _fp(field))
        bd = tuple(d // 32 for d in dims)
        active = np.empty(bd, np.uint8)
        lib.sro_get_active(h, active.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        G = np.empty((len(x), 3, 3), np.float32)
        xbar = np.empty((len(x), 3), np.float32)
        lib.sro_get_aniso(h, _fp(G), _fp(xbar))
        verts = np.empty((nv, 3), np.float32)
        faces = np.empty((nt, 3), np.int32)
        nrm = np.empty((nv, 3), np.float32) if normals else None
        lib.sro_get_mesh(h, _fp(verts), faces.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                         _fp(nrm) if nrm is not None else None)
        vfield = mvel = None
        if velocity is not None:
            vfield = np.empty(tuple(dims) + (3,), np.float32)
            mvel = np.empty((nv, 3), np.float32)
            lib.sro_get_velocity(h, _fp(vfield), _fp(mvel))
        return dict(lo=lo, dims=dims, active=active, n_active=nact, field0=field0, field=field,
                    velocity_field=vfield, mesh_velocity=mvel,
                    sd=lib.sro_sd(h), G=G, xbar=xbar, vertices=verts, faces=faces, normals=nrm,
                    voxel_lo=tuple(info[9:12]), voxel_dims=tuple(info[12:15]))
    finally:
        lib.sro_destroy(h)


def qef(rows) -> np.ndarray:
    """The oracle's QEF solve on rows (k x 4): (n, n . (p - mean))."""
    r = np.ascontiguousarray(np.asarray(rows, np.float32).reshape(-1, 4))
    out = np.zeros(3, np.float32)
    _lib().sro_qef(_fp(r), len(r), _fp(out))
    return out


def expf(x: float) -> float:
    return float(_lib().sro_expf(float(x)))
