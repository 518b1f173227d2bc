"""ctypes wrapper of the CPU oracle (oracle/sph_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Importing it from the
product package is a bug; only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it.
"""
from __future__ import annotations

import ctypes
import os
import sys
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
if _ROOT not in sys.path:
    sys.path.insert(0, _ROOT)

from nonnewton_amd.params import SceneParams, fill_struct, make_ctypes_struct  # noqa: E402

OrParams = make_ctypes_struct("OrParams")

MAX_NB = 60

F_X, F_V, F_ACC, F_VOLUME, F_DENSITY, F_DFSPH_FACTOR, F_DENSITY_ADV = 0, 1, 2, 3, 4, 5, 6
F_STRAIN_NORM, F_VISCOSITY, F_ROTATION, F_STRESS, F_F, F_REST_VOLUME, F_X0 = 7, 8, 9, 10, 11, 12, 13
F_VISC_DIFF = 16
_WIDTH = {F_X: 3, F_V: 3, F_ACC: 3, F_X0: 3, F_ROTATION: 9, F_F: 9, F_STRESS: 6, F_VISC_DIFF: 3}

PROBE_VISCOSITY, PROBE_SURFACE_TENSION, PROBE_ELASTIC = 0, 1, 2

ST_NEIGHBOR_OVERFLOW, ST_CELL_OVERFLOW, ST_OUT_OF_GRID, ST_NAN, ST_REST_OVERFLOW = 1, 2, 4, 8, 16


class OrStats(ctypes.Structure):
    _fields_ = [
        ("iterations_v", ctypes.c_int32),
        ("iterations_p", ctypes.c_int32),
        ("iterations_cg", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("avg_density_err_v", ctypes.c_double),
        ("avg_density_err_p", ctypes.c_double),
        ("cg_rel_residual", ctypes.c_double),
        ("max_density", ctypes.c_double),
        ("num_pairs", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_LIBS = {}


def build():
    import subprocess

    subprocess.check_call(["sh", os.path.join(_HERE, "build.sh")])


def _lib(precision: str):
    if precision in _LIBS:
        return _LIBS[precision]
    # ORACLE_LIB_DIR: a directory of deliberately broken builds, used by the
    # test that checks that the oracle's pinning tests catch planted mistakes
    path = os.path.join(os.environ.get("ORACLE_LIB_DIR", _HERE), f"liboracle_{precision}.so")
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    lib.or_create.restype = ctypes.c_void_p
    lib.or_create.argtypes = [P(OrParams), ctypes.c_int32, P(ctypes.c_float), P(ctypes.c_float), P(ctypes.c_int32)]
    lib.or_destroy.argtypes = [ctypes.c_void_p]
    lib.or_init.argtypes = [ctypes.c_void_p]
    lib.or_init.restype = ctypes.c_int32
    lib.or_search.argtypes = [ctypes.c_void_p]
    lib.or_search.restype = ctypes.c_int32
    lib.or_step.argtypes = [ctypes.c_void_p, P(OrStats)]
    lib.or_step.restype = ctypes.c_int32
    lib.or_get.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_double)]
    lib.or_get.restype = ctypes.c_int32
    lib.or_neighbors.argtypes = [ctypes.c_void_p, P(ctypes.c_int32), P(ctypes.c_int32)]
    lib.or_set.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_double)]
    lib.or_set.restype = ctypes.c_int32
    lib.or_kernel_w.argtypes = [P(OrParams), ctypes.c_double]
    lib.or_kernel_w.restype = ctypes.c_double
    lib.or_kernel_grad.argtypes = [P(OrParams), P(ctypes.c_double), P(ctypes.c_double)]
    lib.or_polar.argtypes = [P(ctypes.c_double), P(ctypes.c_double)]
    lib.or_viscosity_model.argtypes = [P(OrParams), ctypes.c_double]
    lib.or_viscosity_model.restype = ctypes.c_double
    lib.or_weiler_probe.argtypes = [ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_double),
                                    ctypes.c_int32, P(OrStats)]
    lib.or_weiler_probe.restype = ctypes.c_int32
    lib.or_force_probe.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_double)]
    lib.or_force_probe.restype = ctypes.c_int32
    lib.or_rigid_frame.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_double), P(ctypes.c_double),
                                   P(ctypes.c_double)]
    lib.or_rigid_frame.restype = ctypes.c_int32
    lib.or_num_threads.restype = ctypes.c_int32
    lib.or_set_threads.argtypes = [ctypes.c_int32]
    _LIBS[precision] = lib
    return lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def params_struct(p: SceneParams) -> OrParams:
    return fill_struct(OrParams(), p)


class OracleSim:
    """One scene in the oracle.  precision 'f32' (the reference's arithmetic) or 'f64'."""

    def __init__(self, params: SceneParams, x, kind, v=None, precision: str = "f32", compensated: Optional[bool] = None):
        self.lib = _lib(precision)
        self.precision = precision
        if compensated is not None:
            params = params.replace(compensated=bool(compensated))
        self.params = params
        self._p = params_struct(params)
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
        kind = np.ascontiguousarray(kind, dtype=np.int32)
        vv = np.zeros_like(x) if v is None else np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
        self.n = x.shape[0]
        self.h = self.lib.or_create(ctypes.byref(self._p), self.n, _ptr(x, ctypes.c_float), _ptr(vv, ctypes.c_float),
                                    _ptr(kind, ctypes.c_int32))
        if not self.h:
            raise ValueError("or_create failed (bad sizes or domain)")
        self.kind = kind.copy()

    @classmethod
    def from_scene(cls, scene, precision="f32", compensated=None):
        return cls(scene.params, scene.x, scene.kind, scene.v, precision=precision, compensated=compensated)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.or_destroy(h)
            self.h = None

    def init(self) -> int:
        return int(self.lib.or_init(self.h))

    def search(self) -> int:
        return int(self.lib.or_search(self.h))

    def step(self) -> dict:
        st = OrStats()
        self.lib.or_step(self.h, ctypes.byref(st))
        return st.as_dict()

    def get(self, field: int) -> np.ndarray:
        w = _WIDTH.get(field, 1)
        out = np.zeros(self.n * w, dtype=np.float64)
        if self.lib.or_get(self.h, field, _ptr(out, ctypes.c_double)) != 0:
            raise ValueError(f"bad field {field}")
        return out.reshape(self.n, w) if w > 1 else out

    def set(self, field: int, values) -> None:
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(self.n * 3)
        if self.lib.or_set(self.h, field, _ptr(a, ctypes.c_double)) != 0:
            raise ValueError("only x, v and the Weiler warm-start direction can be set")

    def neighbors(self):
        nbr = np.zeros((self.n, MAX_NB), dtype=np.int32)
        cnt = np.zeros(self.n, dtype=np.int32)
        self.lib.or_neighbors(self.h, _ptr(nbr, ctypes.c_int32), _ptr(cnt, ctypes.c_int32))
        return nbr, cnt

    def rigid_frame(self, body: int):
        """(rest centre of mass, centre of mass, rotation) of rigid body `body`
        after its last shape-matching projection."""
        c0 = np.zeros(3)
        cm = np.zeros(3)
        R = np.zeros(9)
        if self.lib.or_rigid_frame(self.h, int(body), _ptr(c0, ctypes.c_double), _ptr(cm, ctypes.c_double),
                                   _ptr(R, ctypes.c_double)) != 0:
            raise ValueError(f"no rigid body {body}")
        return c0, cm, R.reshape(3, 3)

    def force_probe(self, which: int) -> np.ndarray:
        """Accelerations of one force term on the current state (n x 3):
        PROBE_VISCOSITY, PROBE_SURFACE_TENSION or PROBE_ELASTIC."""
        out = np.zeros(self.n * 3, dtype=np.float64)
        if self.lib.or_force_probe(self.h, int(which), _ptr(out, ctypes.c_double)) != 0:
            raise ValueError(f"force probe {which} failed")
        return out.reshape(self.n, 3)

    def weiler_probe(self):
        nf = int(np.count_nonzero(self.kind == 0))
        m = 3 * nf
        mat = np.zeros((m, m), dtype=np.float64)
        b = np.zeros(m, dtype=np.float64)
        xx = np.zeros(m, dtype=np.float64)
        st = OrStats()
        rc = self.lib.or_weiler_probe(self.h, _ptr(mat, ctypes.c_double), _ptr(b, ctypes.c_double),
                                      _ptr(xx, ctypes.c_double), nf, ctypes.byref(st))
        if rc != 0:
            raise ValueError("weiler probe failed")
        return mat, b, xx, st.as_dict()


def kernel_w(params: SceneParams, r: float, precision="f32") -> float:
    p = params_struct(params)
    return float(_lib(precision).or_kernel_w(ctypes.byref(p), float(r)))


def kernel_grad(params: SceneParams, r, precision="f32") -> np.ndarray:
    p = params_struct(params)
    rr = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(3, dtype=np.float64)
    _lib(precision).or_kernel_grad(ctypes.byref(p), _ptr(rr, ctypes.c_double), _ptr(out, ctypes.c_double))
    return out


def polar(a, precision="f64") -> np.ndarray:
    aa = np.ascontiguousarray(a, dtype=np.float64).reshape(9)
    out = np.zeros(9, dtype=np.float64)
    _lib(precision).or_polar(_ptr(aa, ctypes.c_double), _ptr(out, ctypes.c_double))
    return out.reshape(3, 3)


def viscosity_model(params: SceneParams, gamma: float, precision="f64") -> float:
    p = params_struct(params)
    return float(_lib(precision).or_viscosity_model(ctypes.byref(p), float(gamma)))


def set_threads(n: int, precision="f32"):
    _lib(precision).or_set_threads(int(n))


def num_threads(precision="f32") -> int:
    return int(_lib(precision).or_num_threads())
