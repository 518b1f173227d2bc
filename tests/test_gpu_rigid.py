"""GPU parity of rigid bodies (RigidBodySolver, SPH/main.py:1099-1158, and their
two-way coupling with the fluid) through the C-ABI against the FP32 oracle, bit
for bit: a cube dropped into a fluid pool, two bodies of one scene (the
reference projects them one after the other, with a ground collision before
each), coupling intervals 0 and 2, a pool mixing rigid and other scenes, and the
kinds read back."""
import numpy as np
import pytest

from nonnewton_amd import _lib as L
from nonnewton_amd import scenes
from nonnewton_amd.params import KIND_RIGID, SceneParams
from oracle import oracle as O

from test_gpu_parity import FIELDS, assert_fields_equal, assert_lists_equal, pair

pytestmark = pytest.mark.gpu

SP = scenes.SPACING


def run_pair(sc, steps):
    ctx, orc = pair(sc)
    for step in range(steps):
        sg = ctx.step()
        so = orc.step()
        assert (sg["iterations_v"], sg["iterations_p"]) == (so["iterations_v"], so["iterations_p"]), step
        assert sg["num_pairs"] == so["num_pairs"], step
    assert_lists_equal(ctx, orc)
    assert_fields_equal(ctx, orc, FIELDS)
    return ctx, orc


def test_rigid_drop_bit_exact():
    """An 8^3 rigid cube released above a jittered 24x12x24 fluid pool: 40 steps,
    every field equal to the oracle's; the cube has moved and is still rigid."""
    sc = scenes.rigid_drop()
    ctx, orc = run_pair(sc, 40)
    x = ctx.download(L.FIELD_X).reshape(-1, 3).astype(np.float64)
    body = sc.kind >= KIND_RIGID
    assert np.mean(x[body, 1]) < np.mean(sc.x[body, 1]) - 1e-3
    ctx.close()


def two_bodies(coupling):
    """A small fluid block, one cube sitting below the padding height (the
    ground collision lifts it and the walls) and one falling onto the fluid."""
    rng = np.random.Generator(np.random.PCG64(21))
    cells = (30, 30, 30)
    p = SceneParams(domain_end=tuple(c * SP for c in cells), gravity=(0.0, -9.81, 0.0), dt=5e-4,
                    coupling_interval=coupling)
    fluid = scenes.jitter(rng, scenes.lattice((6 * SP, 3 * SP, 6 * SP), (12, 8, 12)))
    low = scenes.lattice((20 * SP, 1.2 * SP, 20 * SP), (4, 4, 4))
    high = scenes.lattice((9 * SP, 14 * SP, 9 * SP), (5, 3, 4))
    walls = scenes.shell(cells)
    return scenes.assemble("two_bodies", p, fluid, walls, None, 1, rigid=[low, high])


@pytest.mark.parametrize("coupling", [1, 2, 0])
def test_two_bodies_and_coupling_interval_bit_exact(coupling):
    sc = two_bodies(coupling)
    ctx, orc = run_pair(sc, 12)
    kinds = ctx.download(L.FIELD_KIND)
    assert np.array_equal(kinds, sc.kind)  # rigid particles read back as KIND_RIGID + body
    ctx.close()


def test_pool_with_rigid_and_other_scenes_equals_alone():
    """A rigid scene between two scenes without bodies, and a second rigid scene
    appended after a tick: every scene equals itself stepped alone."""
    batch = scenes.mixed_batch(count=2, lo=2000, hi=4000)
    todo = [batch[0], scenes.rigid_drop(fluid_counts=(12, 8, 12), cube=5), batch[1], two_bodies(1)]
    pool = L.Pool(8 << 30)
    slots = [pool.admit(s.params, s.x, s.kind, s.v) for s in todo[:3]]
    pool.tick()
    slots.append(pool.admit(todo[3].params, todo[3].x, todo[3].kind, todo[3].v))
    for _ in range(5):
        pool.tick()
    ticks = [6, 6, 6, 5]
    for sc, slot, n in zip(todo, slots, ticks):
        ctx = L.Context(sc.params, sc.n)
        ctx.upload(sc.x, sc.kind, sc.v)
        ctx.init()
        for _ in range(n):
            ctx.step()
        for f in (L.FIELD_X, L.FIELD_V, L.FIELD_DENSITY):
            assert np.array_equal(pool.read(slot, f), ctx.download(f)), (sc.name, f)
        ctx.close()
    pool.close()


def test_body_numbering_gap_is_an_argument_error():
    p = SceneParams(domain_end=(0.4, 0.4, 0.4))
    pts = np.array([[0.1, 0.1, 0.1], [0.2, 0.2, 0.2]])
    ctx = L.Context(p, len(pts))
    with pytest.raises(L.NNError) as e:
        ctx.upload(pts, np.array([KIND_RIGID + 1, KIND_RIGID + 1], np.int32))  # body 1 without body 0
    assert e.value.code == L.NN_E_ARG
    ctx.close()
