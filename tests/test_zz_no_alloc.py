"""nn_neighbors_noalloc, the entry point that allocates no device memory,
against nn_neighbors of a context (the same lists: sets and the reference's
serial order, SPH/main.py:815-847) on two calls in a row, the second with a
smaller input; free device memory does not change across calls; sizes beyond
the fixed scratch are refused on the host; an out-of-grid particle is
reported in the status word."""
import ctypes

import numpy as np
import pytest

from nonnewton_amd import _lib as L
from nonnewton_amd import scenes

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
    return _hip


def check(rc, what):
    assert rc == 0, f"{what} failed ({rc})"


class DevBuf:
    def __init__(self, nbytes):
        self.p = ctypes.c_void_p()
        self.n = nbytes
        check(hip().hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(nbytes)), "hipMalloc")

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        check(hip().hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1), "hipMemcpy")

    def get(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        check(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(out.nbytes), 2), "hipMemcpy")
        return out

    def free(self):
        hip().hipFree(self.p)


def free_bytes():
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    check(hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value


def context_lists(sc):
    ctx = L.Context(sc.params, sc.n)
    ctx.upload(sc.x, sc.kind, sc.v)
    ctx.init()
    nbr, cnt = ctx.neighbors()
    ctx.close()
    return nbr, cnt


def noalloc(lib, params, bufs, x):
    n = len(x)
    bufs["x"].put(np.ascontiguousarray(x, dtype=np.float32))
    p = L.params_struct(params)
    rc = lib.nn_neighbors_noalloc(ctypes.byref(p), bufs["x"].p, n, bufs["nbr"].p, bufs["cnt"].p, 60,
                                  bufs["status"].p, None)
    assert rc == 0, L.last_error()
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
    return (bufs["nbr"].get((n, 60), np.int32), bufs["cnt"].get((n,), np.int32),
            int(bufs["status"].get((1,), np.int32)[0]))


def test_noalloc_lists_equal_context_lists_twice():
    lib = L.load()
    big = scenes.dam_break()                       # 22,530 particles
    small = scenes.tiny_box(fluid_counts=(7, 5, 6))  # fewer particles, another grid
    assert small.n < big.n
    bufs = {"x": DevBuf(big.n * 12), "nbr": DevBuf(big.n * 240), "cnt": DevBuf(big.n * 4), "status": DevBuf(4)}
    try:
        ref_big = context_lists(big)
        ref_small = context_lists(small)
        nb, cnt, st = noalloc(lib, big.params, bufs, big.x)
        free0 = free_bytes()
        assert st == 0
        assert np.array_equal(cnt, ref_big[1]) and np.array_equal(nb, ref_big[0])
        nb, cnt, st = noalloc(lib, small.params, bufs, small.x)
        assert st == 0
        assert np.array_equal(cnt, ref_small[1]) and np.array_equal(nb, ref_small[0])
        # the first call's scratch holds the big scene's cells: a third call
        # with the big scene again still equals the context (counts cleared)
        nb, cnt, st = noalloc(lib, big.params, bufs, big.x)
        assert st == 0 and np.array_equal(cnt, ref_big[1]) and np.array_equal(nb, ref_big[0])
        assert free_bytes() == free0  # nothing allocated by the second and third calls
    finally:
        for b in bufs.values():
            b.free()


def test_noalloc_reports_out_of_grid_and_refuses_oversize():
    lib = L.load()
    sc = scenes.tiny_box()
    x = sc.x.copy()
    x[5] = (-0.5, 0.1, 0.1)
    bufs = {"x": DevBuf(sc.n * 12), "nbr": DevBuf(sc.n * 240), "cnt": DevBuf(sc.n * 4), "status": DevBuf(4)}
    try:
        nb, cnt, st = noalloc(lib, sc.params, bufs, x)
        assert st & 4
        assert cnt[5] == 0 and np.all(nb[5] == -1)
        assert not np.any(nb == 5)
        p = L.params_struct(sc.params)
        rc = lib.nn_neighbors_noalloc(ctypes.byref(p), bufs["x"].p, (1 << 20) + 1, bufs["nbr"].p, bufs["cnt"].p, 60,
                                      bufs["status"].p, None)
        assert rc == L.NN_E_ARG and "1048576" in L.last_error()
        wide = sc.params.replace(domain_end=(400.0, 400.0, 400.0))  # 10^4 cells per axis
        rc = lib.nn_neighbors_noalloc(ctypes.byref(L.params_struct(wide)), bufs["x"].p, sc.n, bufs["nbr"].p,
                                      bufs["cnt"].p, 60, bufs["status"].p, None)
        assert rc == L.NN_E_ARG and "cells" in L.last_error()
    finally:
        for b in bufs.values():
            b.free()
