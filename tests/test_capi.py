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


def test_product_does_not_import_the_oracle():
    import glob
    for path in glob.glob(os.path.join(ROOT, "nonnewton_amd", "**", "*.py"), recursive=True):
        text = open(path).read()
        assert "oracle" not in text.replace("the oracle", ""), path
