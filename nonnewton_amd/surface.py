"""Particle surface reconstruction on the GPU: the host mirror of SRTool
(SPH/surf_recon/srtool.py:869-1838) over the nn_surface_* entry points of
libnnsph.so.

    sr = SRTool(bounding_box, voxel_size=0.01, particle_radius=0.02)
    sr.dual_contouring(isolevel=0, smooth_radius=0.06, num_particles=n, pos=points)
    sr.export_mesh("torus.ply")

keeps the reference's names: `mesh_vertex`, `mesh_index` (flat, 3 per
triangle), `num_vertices[None]`, `num_indices[None]`, `mesh_normal`.  `pos`
is an (n, 3) array (or anything with `to_numpy()`); `Surface.run_device`
takes positions already in device memory.  There is no CPU path: a missing
library raises.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _lib as L
from .ply import export_ply_mesh


class _Scalar:
    """num_vertices[None] / num_indices[None] of a 0-d Taichi field."""

    def __init__(self, v: int = 0):
        self.v = int(v)

    def __getitem__(self, key):
        return self.v

    def __int__(self):
        return self.v


class Surface:
    """One nn_surface object: fixed parameters, reusable device buffers."""

    def __init__(self, origin=(0.0, 0.0, 0.0), voxel_size=0.01, particle_radius=0.01, smooth_radius=0.04,
                 isolevel=0.0, anisotropic=True, smooth_iterations=1, normals=False, velocity=False):
        lib = L.load()
        p = L.NnSurfaceParams()
        p.origin[:] = [float(v) for v in origin]
        p.voxel_size = float(voxel_size)
        p.particle_radius = float(particle_radius)
        p.smooth_radius = float(smooth_radius)
        p.isolevel = float(isolevel)
        p.anisotropic = int(bool(anisotropic))
        p.smooth_iterations = int(smooth_iterations)
        p.normals = int(bool(normals))
        p.velocity = int(bool(velocity))
        self.params = p
        self._h = ctypes.c_void_p()
        L.check(lib.nn_surface_create(ctypes.byref(p), ctypes.byref(self._h)), "nn_surface_create")
        self.n = 0

    def close(self):
        if self._h:
            L.load().nn_surface_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_budget(self, nbytes: int):
        L.check(L.load().nn_surface_set_budget(self._h, int(nbytes)), "nn_surface_set_budget")

    def run(self, pos, vel=None) -> dict:
        """Reconstruct from (n, 3) host positions (and velocities, for an object
        made with velocity=True)."""
        x = np.ascontiguousarray(np.asarray(pos, dtype=np.float32).reshape(-1, 3))
        self.n = len(x)
        if vel is None:
            L.check(L.load().nn_surface_run(self._h, L.fptr(x), len(x)), "nn_surface_run")
        else:
            u = np.ascontiguousarray(np.asarray(vel, dtype=np.float32).reshape(-1, 3))
            if u.shape != x.shape:
                raise ValueError("velocities must have one (vx, vy, vz) per particle")
            L.check(L.load().nn_surface_run_velocity(self._h, L.fptr(x), L.fptr(u), len(x)), "nn_surface_run_velocity")
        return self.info()

    def run_context(self, ctx) -> dict:
        """Reconstruct from the fluid particles of a nonnewton_amd._lib.Context
        as its last step left them on the device."""
        L.check(L.load().nn_surface_run_context(self._h, ctx.h), "nn_surface_run_context")
        info = self.info()
        self.n = int(info["particles"])
        return info

    def run_device(self, ptr: int, n: int) -> dict:
        """Reconstruct from n positions (x, y, z float triples, caller order)
        already in device memory at address ptr; the caller's writes to them
        must be complete."""
        self.n = int(n)
        L.check(L.load().nn_surface_run_device(self._h, ctypes.c_void_p(int(ptr)), self.n), "nn_surface_run_device")
        return self.info()

    def info(self) -> dict:
        out = (ctypes.c_int64 * 16)()
        sd = ctypes.c_double()
        L.check(L.load().nn_surface_info(self._h, out, ctypes.byref(sd)), "nn_surface_info")
        return dict(lo=tuple(out[0:3]), dims=tuple(out[3:6]), num_vertices=out[6], num_triangles=out[7],
                    active_blocks=out[8], voxel_lo=tuple(out[9:12]), voxel_dims=tuple(out[12:15]), particles=out[15],
                    sd=sd.value)

    def mesh(self):
        """(vertices (V, 3), faces (T, 3), normals (V, 3) or None)."""
        inf = self.info()
        v = np.empty((inf["num_vertices"], 3), np.float32)
        f = np.empty((inf["num_triangles"], 3), np.int32)
        n = np.empty((inf["num_vertices"], 3), np.float32) if self.params.normals else None
        L.check(L.load().nn_surface_mesh(self._h, L.fptr(v), L.iptr(f), L.fptr(n) if n is not None else None),
                "nn_surface_mesh")
        return v, f, n

    def mesh_velocity(self) -> np.ndarray:
        """(V, 3) velocities at the mesh vertices of the last run_velocity."""
        out = np.empty((self.info()["num_vertices"], 3), np.float32)
        L.check(L.load().nn_surface_mesh_velocity(self._h, L.fptr(out)), "nn_surface_mesh_velocity")
        return out

    def field(self, which: int = 1) -> np.ndarray:
        """The vertex field on the grid (dims from info()): 0 before smoothing, 1 after."""
        dims = self.info()["dims"]
        out = np.empty(dims, np.float32)
        L.check(L.load().nn_surface_field(self._h, int(which), L.fptr(out), out.size), "nn_surface_field")
        return out

    def aniso(self):
        """(G (n, 3, 3), x_bar (n, 3)) of the last anisotropic run."""
        G = np.empty((self.n, 3, 3), np.float32)
        xb = np.empty((self.n, 3), np.float32)
        L.check(L.load().nn_surface_aniso(self._h, L.fptr(G), L.fptr(xb), self.n), "nn_surface_aniso")
        return G, xb


class SRTool:
    """SRTool (srtool.py:869-1838): same constructor and entry points."""

    def __init__(self, bounding_box_extent, voxel_size=0.01, particle_radius=0.01, record_normals=False,
                 record_velocity=False, max_num_vertices=8000000, max_num_indices=100000000,
                 max_num_particles=9000000, hash_level_dimensions=None):
        self.bounding_box_extent = [list(map(float, bounding_box_extent[0])), list(map(float, bounding_box_extent[1]))]
        self.voxel_size = float(voxel_size)
        self.radius = float(particle_radius)
        self.record_normals = bool(record_normals)
        self.record_velocity = bool(record_velocity)
        self.max_num_vertices = int(max_num_vertices)
        self.max_num_indices = int(max_num_indices)
        self.max_num_particles = int(max_num_particles)
        self._surf: Optional[Surface] = None
        self._key = None
        self.reset()

    def reset(self):
        self.num_vertices = _Scalar(0)
        self.num_indices = _Scalar(0)
        self.mesh_vertex = np.zeros((0, 3), np.float32)
        self.mesh_index = np.zeros(0, np.int32)
        self.mesh_normal = np.zeros((0, 3), np.float32) if self.record_normals else None
        self.mesh_velocity = np.zeros((0, 3), np.float32) if self.record_velocity else None

    def _surface(self, isolevel, smooth_radius, anisotropic) -> Surface:
        key = (float(isolevel), float(smooth_radius), bool(anisotropic))
        if self._surf is None or self._key != key:
            if self._surf is not None:
                self._surf.close()
            self._surf = Surface(origin=self.bounding_box_extent[0], voxel_size=self.voxel_size,
                                 particle_radius=self.radius, smooth_radius=smooth_radius, isolevel=isolevel,
                                 anisotropic=anisotropic, smooth_iterations=1, normals=self.record_normals,
                                 velocity=self.record_velocity)
            self._key = key
        return self._surf

    def dual_contouring(self, isolevel, smooth_radius, num_particles, pos, velocity=None,
                        use_smooth_fluid_field=True, record_normals=True):
        """srtool.py:1264-1304.  `record_normals` is accepted and, as in the
        reference, the constructor's record_normals decides."""
        n = int(num_particles)
        if n > self.max_num_particles:
            raise ValueError(f"num_particles {n} exceeds max_num_particles {self.max_num_particles}")
        self.reset()
        s = self._surface(isolevel, smooth_radius, use_smooth_fluid_field)
        p = pos.to_numpy() if hasattr(pos, "to_numpy") else pos
        if self.record_velocity:
            if velocity is None:
                raise ValueError("SRTool(record_velocity=True) needs the particles' velocity")
            u = velocity.to_numpy() if hasattr(velocity, "to_numpy") else velocity
            s.run(p[:n], u[:n])
        else:
            s.run(p[:n])
        v, f, nrm = s.mesh()
        if len(v) > self.max_num_vertices or 3 * len(f) > self.max_num_indices:
            raise ValueError(f"mesh of {len(v)} vertices / {3 * len(f)} indices exceeds max_num_vertices "
                             f"{self.max_num_vertices} / max_num_indices {self.max_num_indices}")
        self.mesh_vertex = v
        self.mesh_index = f.reshape(-1)
        self.mesh_normal = nrm
        if self.record_velocity:
            self.mesh_velocity = s.mesh_velocity()
        self.num_vertices = _Scalar(len(v))
        self.num_indices = _Scalar(3 * len(f))

    def export_mesh(self, file_path: str, output_normal=False, output_velocity=False):
        """srtool.py:1802-1838: binary PLY of the last mesh."""
        assert file_path[-3:] == "ply", "Output file must be in ply format!"
        if output_velocity and self.mesh_velocity is None:
            raise ValueError("output_velocity needs SRTool(record_velocity=True)")
        if output_normal and self.mesh_normal is None:
            raise ValueError("output_normal needs SRTool(record_normals=True)")
        export_ply_mesh(self.mesh_vertex, self.mesh_index.reshape(-1, 3), file_path,
                        normals=self.mesh_normal if output_normal else None,
                        velocity=self.mesh_velocity if output_velocity else None)
