"""ctypes binding of libnnsph.so (include/nnsph.h).

The library is loaded from this package directory; a missing or unloadable
library raises immediately -- there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
import re
from typing import Optional

import numpy as np

from .params import SceneParams, fill_struct, make_ctypes_struct

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnnsph.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "nnsph.h")

NnParams = make_ctypes_struct("NnParams")

NN_OK = 0
NN_E_ARG = -1
NN_E_HIP = -2
NN_E_OOM = -3
NN_E_NEIGHBOR_OVERFLOW = -4
NN_E_CELL_OVERFLOW = -5
NN_E_OUT_OF_GRID = -6
NN_E_NAN = -7
NN_E_UNSUPPORTED = -8
NN_E_BUSY = -9
NN_E_STATE = -10
NN_E_REST_OVERFLOW = -11
NN_W_NOCONV = 1
POOL_OWN_QUEUE = 1  # NN_POOL_OWN_QUEUE

ERROR_NAMES = {
    NN_E_ARG: "NN_E_ARG",
    NN_E_HIP: "NN_E_HIP",
    NN_E_OOM: "NN_E_OOM",
    NN_E_NEIGHBOR_OVERFLOW: "NN_E_NEIGHBOR_OVERFLOW",
    NN_E_CELL_OVERFLOW: "NN_E_CELL_OVERFLOW",
    NN_E_OUT_OF_GRID: "NN_E_OUT_OF_GRID",
    NN_E_NAN: "NN_E_NAN",
    NN_E_UNSUPPORTED: "NN_E_UNSUPPORTED",
    NN_E_BUSY: "NN_E_BUSY",
    NN_E_STATE: "NN_E_STATE",
    NN_E_REST_OVERFLOW: "NN_E_REST_OVERFLOW",
}

FIELD_X = 0
FIELD_V = 1
FIELD_ACCELERATION = 2
FIELD_VOLUME = 3
FIELD_DENSITY = 4
FIELD_DFSPH_FACTOR = 5
FIELD_DENSITY_ADV = 6
FIELD_STRAIN_NORM = 7
FIELD_VISCOSITY = 8
FIELD_ROTATION = 9
FIELD_STRESS = 10
FIELD_REST_VOLUME = 12
FIELD_X0 = 13
FIELD_KIND = 14
FIELD_NUM_NEIGHBORS = 15
FIELD_VISC_DIFF = 16

FIELD_WIDTH = {FIELD_X: 3, FIELD_V: 3, FIELD_ACCELERATION: 3, FIELD_X0: 3, FIELD_ROTATION: 9, FIELD_STRESS: 6,
               FIELD_VISC_DIFF: 3}
INT_FIELDS = {FIELD_KIND, FIELD_NUM_NEIGHBORS}

MAX_NEIGHBORS = 60


class NnStepStats(ctypes.Structure):
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
        ("num_pairs_fluid", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class NnSurfaceParams(ctypes.Structure):
    """nn_surface_params of include/nnsph.h."""

    _fields_ = [
        ("origin", ctypes.c_double * 3),
        ("voxel_size", ctypes.c_double),
        ("particle_radius", ctypes.c_double),
        ("smooth_radius", ctypes.c_double),
        ("isolevel", ctypes.c_double),
        ("anisotropic", ctypes.c_int32),
        ("smooth_iterations", ctypes.c_int32),
        ("normals", ctypes.c_int32),
        ("velocity", ctypes.c_int32),
    ]


class NNError(RuntimeError):
    def __init__(self, code: int, where: str, text: str = ""):
        self.code = code
        super().__init__(f"{where}: {ERROR_NAMES.get(code, code)} {text}".strip())


_lib = None


def declared_symbols():
    """Every function name declared in include/nnsph.h."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(nn_[a-z0-9_]+)\s*\(", text)))


def load():
    """Load libnnsph.so (raises if it is missing: there is no other path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    i32 = ctypes.c_int32
    i64 = ctypes.c_int64
    sz = ctypes.c_size_t
    sig = {
        "nn_version": (i32, []),
        "nn_last_error": (ctypes.c_char_p, []),
        "nn_device_info": (i32, [ctypes.c_char_p, i32, P(i64), P(i32)]),
        "nn_create": (i32, [P(NnParams), i64, P(vp)]),
        "nn_destroy": (i32, [vp]),
        "nn_upload": (i32, [vp, P(ctypes.c_float), P(ctypes.c_float), P(i32), i64]),
        "nn_init": (i32, [vp]),
        "nn_step": (i32, [vp, i32, P(NnStepStats)]),
        "nn_search": (i32, [vp]),
        "nn_download": (i32, [vp, i32, vp, sz]),
        "nn_upload_field": (i32, [vp, i32, vp, sz]),
        "nn_neighbors": (i32, [vp, P(i32), P(i32), i32]),
        "nn_neighbors_noalloc": (i32, [P(NnParams), vp, i64, vp, vp, i32, vp, vp]),
        "nn_set_param": (i32, [vp, ctypes.c_char_p, ctypes.c_double]),
        "nn_counts": (i32, [vp, P(i64), P(i64)]),
        "nn_set_block": (i32, [vp, i32]),
        "nn_pool_create": (i32, [i64, P(vp)]),
        "nn_pool_create_ex": (i32, [i64, i32, P(vp)]),
        "nn_pool_destroy": (i32, [vp]),
        "nn_pool_footprint": (i64, [i64, i64, P(NnParams)]),
        "nn_pool_admit": (i32, [vp, P(NnParams), P(ctypes.c_float), P(ctypes.c_float), P(i32), i64, P(i32)]),
        "nn_pool_tick": (i32, [vp, P(NnStepStats), i32, P(i32)]),
        "nn_pool_tick_async": (i32, [vp]),
        "nn_pool_sync": (i32, [vp, P(NnStepStats), i32, P(i32)]),
        "nn_pool_retire": (i32, [vp, i32]),
        "nn_pool_steps": (i64, [vp, i32]),
        "nn_pool_read": (i32, [vp, i32, i32, vp, sz]),
        "nn_pool_info": (i32, [vp, P(i32), P(i32), P(i64)]),
        "nn_pool_capacity": (i32, []),
        "nn_pool_set_block": (i32, [vp, i32]),
        "nn_kernel_name": (ctypes.c_char_p, [i32]),
        "nn_profile": (i32, [vp, i32]),
        "nn_set_kernel_block": (i32, [vp, i32, i32]),
        "nn_set_kernel_depth": (i32, [vp, i32, i32]),
        "nn_pool_set_kernel_depth": (i32, [vp, i32, i32]),
        "nn_pool_set_kernel_block": (i32, [vp, i32, i32]),
        "nn_profile_mask": (i32, [vp, ctypes.c_uint32, i32]),
        "nn_kernel_launches": (i32, [vp, P(i64), i32]),
        "nn_kernel_times": (i32, [vp, P(ctypes.c_double), P(i64), i32]),
        "nn_pool_profile": (i32, [vp, i32]),
        "nn_pool_profile_mask": (i32, [vp, ctypes.c_uint32, i32]),
        "nn_pool_kernel_launches": (i32, [vp, P(i64), i32]),
        "nn_pool_kernel_times": (i32, [vp, P(ctypes.c_double), P(i64), i32]),
        "nn_event_overhead": (i32, [i32, P(ctypes.c_double)]),
        "nn_surface_create": (i32, [P(NnSurfaceParams), P(vp)]),
        "nn_surface_destroy": (i32, [vp]),
        "nn_surface_set_budget": (i32, [vp, i64]),
        "nn_surface_run": (i32, [vp, P(ctypes.c_float), i64]),
        "nn_surface_run_device": (i32, [vp, vp, i64]),
        "nn_surface_info": (i32, [vp, P(i64), P(ctypes.c_double)]),
        "nn_surface_mesh": (i32, [vp, P(ctypes.c_float), P(i32), P(ctypes.c_float)]),
        "nn_surface_field": (i32, [vp, i32, P(ctypes.c_float), sz]),
        "nn_surface_aniso": (i32, [vp, P(ctypes.c_float), P(ctypes.c_float), i64]),
        "nn_surface_run_velocity": (i32, [vp, P(ctypes.c_float), P(ctypes.c_float), i64]),
        "nn_surface_run_context": (i32, [vp, vp]),
        "nn_surface_mesh_velocity": (i32, [vp, P(ctypes.c_float)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


KERNEL_COUNT = 21


def event_overhead_ms(reps: int = 200) -> float:
    """Launch and event latency inside a timed launch's interval (nn_event_overhead)."""
    ms = ctypes.c_double()
    check(load().nn_event_overhead(int(reps), ctypes.byref(ms)), "nn_event_overhead")
    return float(ms.value)


def device_name() -> str:
    """Name of the device the library runs on (nn_device_info)."""
    lib = load()
    buf = ctypes.create_string_buffer(256)
    hbm = ctypes.c_int64()
    cus = ctypes.c_int32()
    check(lib.nn_device_info(buf, 256, ctypes.byref(hbm), ctypes.byref(cus)), "nn_device_info")
    return buf.value.decode(errors="replace")


def kernel_mask(kernels) -> int:
    """Bit mask of the named kernels for the _mask profiling calls."""
    names = kernel_names()
    m = 0
    for k in kernels:
        if k not in names:
            raise ValueError(f"unknown kernel {k!r}")
        m |= 1 << names.index(k)
    return m


def kernel_names():
    lib = load()
    return [lib.nn_kernel_name(k).decode() for k in range(KERNEL_COUNT)]


def _seen(fn, h):
    n = (ctypes.c_int64 * KERNEL_COUNT)()
    check(fn(h, n, KERNEL_COUNT), "kernel_launches")
    return {name: int(n[k]) for k, name in enumerate(kernel_names())}


def _times(fn, h):
    ms = (ctypes.c_double * KERNEL_COUNT)()
    cnt = (ctypes.c_int64 * KERNEL_COUNT)()
    check(fn(h, ms, cnt, KERNEL_COUNT), "kernel_times")
    return {name: (float(ms[k]), int(cnt[k])) for k, name in enumerate(kernel_names())}


def last_error() -> str:
    t = load().nn_last_error()
    return t.decode() if t else ""


def check(rc: int, where: str, allow_noconv: bool = True) -> int:
    if rc < 0:
        raise NNError(rc, where, last_error())
    if rc == NN_W_NOCONV and not allow_noconv:
        raise NNError(rc, where, "solver did not converge")
    return rc


def params_struct(p: SceneParams) -> NnParams:
    return fill_struct(NnParams(), p)


def fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def iptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


class Context:
    """One scene on the GPU (nn_ctx)."""

    def __init__(self, params: SceneParams, max_particles: int):
        self.lib = load()
        self.params = params
        self._p = params_struct(params)
        h = ctypes.c_void_p()
        check(self.lib.nn_create(ctypes.byref(self._p), int(max_particles), ctypes.byref(h)), "nn_create")
        self.h = h
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.nn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, x, kind, v=None):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
        kind = np.ascontiguousarray(kind, dtype=np.int32)
        vv = np.zeros_like(x) if v is None else np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
        self.n = x.shape[0]
        check(self.lib.nn_upload(self.h, fptr(x), fptr(vv), iptr(kind), self.n), "nn_upload")

    def init(self):
        check(self.lib.nn_init(self.h), "nn_init")

    def step(self, nsteps: int = 1, allow_noconv: bool = True) -> dict:
        st = NnStepStats()
        rc = self.lib.nn_step(self.h, int(nsteps), ctypes.byref(st))
        check(rc, "nn_step", allow_noconv)
        d = st.as_dict()
        d["rc"] = rc
        return d

    def search(self):
        check(self.lib.nn_search(self.h), "nn_search")

    def download(self, field: int) -> np.ndarray:
        w = FIELD_WIDTH.get(field, 1)
        dt = np.int32 if field in INT_FIELDS else np.float32
        out = np.zeros((self.n, w) if w > 1 else (self.n,), dtype=dt)
        check(self.lib.nn_download(self.h, field, out.ctypes.data_as(ctypes.c_void_p), out.nbytes), "nn_download")
        return out

    def upload_field(self, field: int, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(self.n, 3)
        check(self.lib.nn_upload_field(self.h, field, a.ctypes.data_as(ctypes.c_void_p), a.nbytes), "nn_upload_field")

    def neighbors(self, cap: int = MAX_NEIGHBORS):
        nbr = np.zeros((self.n, cap), dtype=np.int32)
        cnt = np.zeros(self.n, dtype=np.int32)
        check(self.lib.nn_neighbors(self.h, iptr(nbr), iptr(cnt), cap), "nn_neighbors")
        return nbr, cnt

    def set_param(self, key: str, value: float):
        check(self.lib.nn_set_param(self.h, key.encode(), float(value)), "nn_set_param")

    def set_block(self, block: int):
        check(self.lib.nn_set_block(self.h, int(block)), "nn_set_block")

    def set_kernel_block(self, kernel: str, block: int):
        check(self.lib.nn_set_kernel_block(self.h, kernel_names().index(kernel), int(block)), "nn_set_kernel_block")

    def set_kernel_depth(self, kernel: str, depth: int):
        check(self.lib.nn_set_kernel_depth(self.h, kernel_names().index(kernel), int(depth)), "nn_set_kernel_depth")

    def profile(self, enable: bool = True, kernels=None, stride: int = 1):
        """Time every launch of every kernel (kernels None), or every stride-th
        launch of the named ones."""
        if kernels is None and stride == 1:
            check(self.lib.nn_profile(self.h, int(enable)), "nn_profile")
        else:
            m = (kernel_mask(kernels) if kernels is not None else 0xFFFFFFFF) if enable else 0
            check(self.lib.nn_profile_mask(self.h, m, int(stride)), "nn_profile_mask")

    def kernel_launches(self):
        """{kernel name: launches made while timing was on (timed or not)}."""
        return _seen(self.lib.nn_kernel_launches, self.h)

    def kernel_times(self):
        """{kernel name: (total ms, launches)} since profiling was enabled."""
        return _times(self.lib.nn_kernel_times, self.h)


class Pool:
    """Many resident scenes stepped together (nn_pool)."""

    def __init__(self, budget_bytes: int, own_queue: bool = False):
        """own_queue: the pool's stream gets a hardware queue of its own
        (NN_POOL_OWN_QUEUE), for pools stepped concurrently (lanes)."""
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.nn_pool_create_ex(int(budget_bytes), POOL_OWN_QUEUE if own_queue else 0, ctypes.byref(h)),
              "nn_pool_create_ex")
        self.h = h
        self.capacity = int(self.lib.nn_pool_capacity())
        self._stats = (NnStepStats * self.capacity)()
        self.sizes = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.nn_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def footprint(self, params: SceneParams, n: int, n_elastic: int = 0) -> int:
        p = params_struct(params)
        return int(self.lib.nn_pool_footprint(int(n), int(n_elastic), ctypes.byref(p)))

    def admit(self, params: SceneParams, x, kind, v=None) -> Optional[int]:
        """Slot id, or None when the budget is exhausted (NN_E_BUSY)."""
        p = params_struct(params)
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
        kind = np.ascontiguousarray(kind, dtype=np.int32)
        vv = np.zeros_like(x) if v is None else np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
        slot = ctypes.c_int32(-1)
        rc = self.lib.nn_pool_admit(self.h, ctypes.byref(p), fptr(x), fptr(vv), iptr(kind), x.shape[0], ctypes.byref(slot))
        if rc == NN_E_BUSY:
            return None
        check(rc, "nn_pool_admit")
        self.sizes[slot.value] = x.shape[0]
        return int(slot.value)

    # statuses a scene can report from a step; the other scenes of the tick are unaffected
    SCENE_ERRORS = (NN_E_NAN, NN_E_OUT_OF_GRID, NN_E_CELL_OVERFLOW, NN_E_NEIGHBOR_OVERFLOW)
    STATUS_ERROR_BITS = 1 | 2 | 4 | 8

    def tick(self):
        """One step of every resident scene: (rc, scenes stepped).  A scene-level
        error is returned, not raised; find the scene through stats()['status']."""
        stepped = ctypes.c_int32(0)
        rc = self.lib.nn_pool_tick(self.h, self._stats, self.capacity, ctypes.byref(stepped))
        if rc not in self.SCENE_ERRORS:
            check(rc, "nn_pool_tick")
        return rc, int(stepped.value)

    def tick_async(self):
        """Start a tick on a library thread and return at once (sync() waits)."""
        check(self.lib.nn_pool_tick_async(self.h), "nn_pool_tick_async")

    def sync(self):
        """Wait for the tick tick_async started: (rc, scenes stepped), as tick()."""
        stepped = ctypes.c_int32(0)
        rc = self.lib.nn_pool_sync(self.h, self._stats, self.capacity, ctypes.byref(stepped))
        if rc not in self.SCENE_ERRORS:
            check(rc, "nn_pool_sync")
        return rc, int(stepped.value)

    def stats(self, slot: int) -> dict:
        return self._stats[slot].as_dict()

    def stats_array(self, slots) -> np.ndarray:
        """Copy of the last tick's statistics of `slots` as a numpy structured
        array (fields of nn_step_stats), without per-scene Python objects."""
        view = np.ctypeslib.as_array(self._stats)
        return view[np.asarray(slots, dtype=np.int64)].copy()

    def retire(self, slot: int):
        check(self.lib.nn_pool_retire(self.h, int(slot)), "nn_pool_retire")
        self.sizes.pop(slot, None)

    def steps(self, slot: int) -> int:
        """Ticks the scene in `slot` has been stepped (0 while still admitted)."""
        return int(self.lib.nn_pool_steps(self.h, int(slot)))

    def read(self, slot: int, field: int) -> np.ndarray:
        n = self.sizes[slot]
        w = FIELD_WIDTH.get(field, 1)
        dt = np.int32 if field in INT_FIELDS else np.float32
        out = np.zeros((n, w) if w > 1 else (n,), dtype=dt)
        check(self.lib.nn_pool_read(self.h, int(slot), field, out.ctypes.data_as(ctypes.c_void_p), out.nbytes), "nn_pool_read")
        return out

    def info(self):
        r, q, b = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        check(self.lib.nn_pool_info(self.h, ctypes.byref(r), ctypes.byref(q), ctypes.byref(b)), "nn_pool_info")
        return int(r.value), int(q.value), int(b.value)

    def set_block(self, block: int):
        check(self.lib.nn_pool_set_block(self.h, int(block)), "nn_pool_set_block")

    def set_kernel_block(self, kernel: str, block: int):
        check(self.lib.nn_pool_set_kernel_block(self.h, kernel_names().index(kernel), int(block)),
              "nn_pool_set_kernel_block")

    def set_kernel_depth(self, kernel: str, depth: int):
        check(self.lib.nn_pool_set_kernel_depth(self.h, kernel_names().index(kernel), int(depth)),
              "nn_pool_set_kernel_depth")

    def profile(self, enable: bool = True, kernels=None, stride: int = 1):
        """Time every launch of every kernel (kernels None), or every stride-th
        launch of the named ones."""
        if kernels is None and stride == 1:
            check(self.lib.nn_pool_profile(self.h, int(enable)), "nn_pool_profile")
        else:
            m = (kernel_mask(kernels) if kernels is not None else 0xFFFFFFFF) if enable else 0
            check(self.lib.nn_pool_profile_mask(self.h, m, int(stride)), "nn_pool_profile_mask")

    def kernel_launches(self):
        """{kernel name: launches made while timing was on (timed or not)}."""
        return _seen(self.lib.nn_pool_kernel_launches, self.h)

    def kernel_times(self):
        return _times(self.lib.nn_pool_kernel_times, self.h)
