"""The C-ABI library loads and exports every function include/nnsph.h declares;
the ctypes structs have the C layout (checked with the system C compiler);
argument errors are reported without touching a GPU."""
import ctypes
import os
import subprocess

import pytest

from nonnewton_amd import _lib as L
from nonnewton_amd.params import C_FIELDS, SceneParams

from conftest import ROOT


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = L.declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing


def test_version_and_kernel_names():
    lib = L.load()
    assert lib.nn_version() == 1
    names = L.kernel_names()
    assert names[0] == "cell_count" and names[-1] == "cg_guess" and len(set(names)) == L.KERNEL_COUNT
    assert lib.nn_kernel_name(L.KERNEL_COUNT) == b""


def test_struct_layout_matches_header(tmp_path):
    fields = [f for f, _ in C_FIELDS]
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "nnsph.h"', "int main(void){",
           'printf("%zu %zu\\n", sizeof(nn_params), sizeof(nn_step_stats));']
    src += [f'printf("%zu\\n", offsetof(nn_params, {f}));' for f in fields]
    stats = [f for f, _ in L.NnStepStats._fields_]
    src += [f'printf("%zu\\n", offsetof(nn_step_stats, {f}));' for f in stats]
    src += ["return 0;}"]
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)])
    out = subprocess.check_output([str(exe)]).decode().split()
    sp, ss = int(out[0]), int(out[1])
    assert sp == ctypes.sizeof(L.NnParams) and ss == ctypes.sizeof(L.NnStepStats)
    offs = [int(v) for v in out[2:]]
    assert offs[: len(fields)] == [getattr(L.NnParams, f).offset for f in fields]
    assert offs[len(fields):] == [getattr(L.NnStepStats, f).offset for f in stats]


def test_argument_errors_without_gpu():
    lib = L.load()
    h = ctypes.c_void_p()
    assert lib.nn_create(None, 10, ctypes.byref(h)) == L.NN_E_ARG
    bad = L.params_struct(SceneParams(particle_radius=0.0))
    assert lib.nn_create(ctypes.byref(bad), 10, ctypes.byref(h)) == L.NN_E_ARG
    assert "positive" in L.last_error()
    bad = L.params_struct(SceneParams(viscosity_method=7))
    assert lib.nn_create(ctypes.byref(bad), 10, ctypes.byref(h)) == L.NN_E_ARG
    bad = L.params_struct(SceneParams(warm_start=2))
    assert lib.nn_create(ctypes.byref(bad), 10, ctypes.byref(h)) == L.NN_E_UNSUPPORTED
    assert lib.nn_pool_footprint(1000, 0, ctypes.byref(L.params_struct(SceneParams()))) > 1000 * 500
    # pool flags: anything but NN_POOL_OWN_QUEUE, a zero budget or no output is refused before any HIP call
    assert lib.nn_pool_create_ex(1 << 30, 2, ctypes.byref(h)) == L.NN_E_ARG
    assert lib.nn_pool_create_ex(1 << 30, L.POOL_OWN_QUEUE | 4, ctypes.byref(h)) == L.NN_E_ARG
    assert lib.nn_pool_create_ex(0, L.POOL_OWN_QUEUE, ctypes.byref(h)) == L.NN_E_ARG
    assert lib.nn_pool_create_ex(1 << 30, 0, None) == L.NN_E_ARG


def test_product_does_not_import_the_oracle():
    import glob
    for path in glob.glob(os.path.join(ROOT, "nonnewton_amd", "**", "*.py"), recursive=True):
        text = open(path).read()
        assert "oracle" not in text.replace("the oracle", ""), path


def test_surface_argument_errors_without_gpu():
    """nn_surface_create rejects bad parameters before it touches a device."""
    import ctypes

    lib = L.load()
    h = ctypes.c_void_p()
    assert lib.nn_surface_create(None, ctypes.byref(h)) == L.NN_E_ARG
    for field, value in (("voxel_size", 0.0), ("particle_radius", -1.0), ("smooth_radius", 0.0),
                         ("smooth_iterations", -1), ("smooth_iterations", 65)):
        p = L.NnSurfaceParams()
        p.voxel_size, p.particle_radius, p.smooth_radius, p.smooth_iterations = 0.01, 0.02, 0.06, 1
        setattr(p, field, value)
        assert lib.nn_surface_create(ctypes.byref(p), ctypes.byref(h)) == L.NN_E_ARG, field
    p = L.NnSurfaceParams()
    p.voxel_size, p.particle_radius, p.smooth_radius, p.anisotropic = 0.01, 0.02, 0.2, 1  # box of 82 vertices
    assert lib.nn_surface_create(ctypes.byref(p), ctypes.byref(h)) == L.NN_E_UNSUPPORTED
    assert "32" in L.last_error()
    assert lib.nn_surface_run(None, None, 0) == L.NN_E_ARG


def test_surface_params_layout_matches_header(tmp_path):
    fields = [f for f, _ in L.NnSurfaceParams._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "nnsph.h"', "int main(void){",
           'printf("%zu\\n", sizeof(nn_surface_params));']
    src += [f'printf("%zu\\n", offsetof(nn_surface_params, {f}));' for f in fields]
    src += ["return 0;}"]
    c = tmp_path / "layout_sr.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout_sr"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)])
    out = [int(v) for v in subprocess.check_output([str(exe)]).decode().split()]
    assert out[0] == ctypes.sizeof(L.NnSurfaceParams)
    assert out[1:] == [getattr(L.NnSurfaceParams, f).offset for f in fields]


def test_async_tick_argument_errors_without_gpu():
    lib = L.load()
    assert lib.nn_pool_tick_async(None) == L.NN_E_ARG
    assert lib.nn_pool_sync(None, None, 0, None) == L.NN_E_ARG
