"""Lanes (nonnewton_amd.lanes.LanePool): scenes spread over pools stepped
concurrently give the same bits as in one pool, whatever their lane."""
import numpy as np
import pytest

from nonnewton_amd import _lib as L
from nonnewton_amd import scenes
from nonnewton_amd.lanes import LanePool

pytestmark = pytest.mark.gpu


def test_lanes_equal_one_pool():
    scs = scenes.mixed_batch(seed=11, count=7, lo=2000, hi=6000)
    ticks = 4
    one = L.Pool(8 << 30)
    slots = [one.admit(s.params, s.x, s.kind, s.v) for s in scs]
    for _ in range(ticks):
        one.tick()
    ref = [(one.read(q, L.FIELD_X), one.read(q, L.FIELD_V), one.stats(q)["iterations_cg"]) for q in slots]
    one.close()
    lp = LanePool(8 << 30, lanes=3)
    keys = [lp.admit(s.params, s.x, s.kind, s.v, n_elastic=s.n_elastic) for s in scs]
    assert {k[0] for k in keys} == {0, 1, 2}
    per_lane = [[k[1] for k in keys if k[0] == l] for l in range(3)]
    elapsed, out = lp.run(ticks, per_lane)
    assert elapsed > 0 and all(len(o[0]) == ticks and len(o[1]) == ticks for o in out)
    for k, (x, v, it) in zip(keys, ref):
        assert np.array_equal(lp.read(k, L.FIELD_X), x) and np.array_equal(lp.read(k, L.FIELD_V), v)
        assert lp.stats(k)["iterations_cg"] == it
        assert lp.steps(k) == ticks
    # retirement frees the lane's load and budget
    before = lp.bytes_in_use()
    lp.retire(keys[0])
    assert lp.bytes_in_use() < before
    lp.tick_all()
    lp.close()


def test_lane_budget_holds_over_all_lanes():
    sc = scenes.mixed_batch(seed=12, count=1, lo=3000, hi=3000)[0]
    probe = L.Pool(1 << 34)
    fp = probe.footprint(sc.params, sc.n, sc.n_elastic)
    probe.close()
    lp = LanePool(int(3.5 * fp), lanes=3)
    a = lp.admit(sc.params, sc.x, sc.kind, sc.v, n_elastic=sc.n_elastic)
    b = lp.admit(sc.params, sc.x, sc.kind, sc.v, n_elastic=sc.n_elastic)
    c = lp.admit(sc.params, sc.x, sc.kind, sc.v, n_elastic=sc.n_elastic)
    assert a is not None and b is not None and a[0] != b[0]
    assert c is None  # 2 resident + 2 for the joining copy > 3.5 footprints
    lp.retire(a)
    assert lp.admit(sc.params, sc.x, sc.kind, sc.v, n_elastic=sc.n_elastic) is not None
    lp.close()


def test_lanes_with_every_scene_kind_equal_alone():
    """Rigid bodies, an iterating pressure solve, elastic blocks and
    non-Newtonian Weiler scenes spread over three lanes: every scene's state
    after 6 ticks equals its run alone in a context, bit for bit (the lanes'
    streams each have a hardware queue of their own and tick concurrently)."""
    todo = scenes.mixed_batch(seed=13, count=3, lo=2000, hi=5000)
    todo += [scenes.rigid_drop(fluid_counts=(12, 8, 12), cube=5), scenes.impact_block(), scenes.tiny_box()]
    assert any(s.n_elastic for s in todo)
    ticks = 6
    lp = LanePool(8 << 30, lanes=3)
    keys = [lp.admit(s.params, s.x, s.kind, s.v, n_elastic=s.n_elastic) for s in todo]
    assert {k[0] for k in keys} == {0, 1, 2}
    lp.run(ticks, [[] for _ in range(3)])
    iters_p = []
    for sc, key in zip(todo, keys):
        ctx = L.Context(sc.params, sc.n)
        ctx.upload(sc.x, sc.kind, sc.v)
        ctx.init()
        iters_p.append(max(ctx.step()["iterations_p"] for _ in range(ticks)))
        for f in (L.FIELD_X, L.FIELD_V, L.FIELD_DENSITY):
            assert np.array_equal(lp.read(key, f), ctx.download(f)), (sc.name, f)
        assert lp.steps(key) == ticks
        ctx.close()
    lp.close()
    assert iters_p[4] >= 3  # the impact block's pressure solve iterates
