"""Scene generators, PLY I/O and the reference's scene-file keys."""
import json
import os

import numpy as np
import pytest

from nonnewton_amd import ply, scenes
from nonnewton_amd.params import KIND_ELASTIC, KIND_FLUID, KIND_STATIC, VISCOSITY_WEILER, params_from_config

from conftest import GOLDEN


def test_dam_break_counts():
    sc = scenes.dam_break()
    assert (sc.n_fluid, sc.n_static, sc.n_elastic) == (13824, 8706, 0)
    assert sc.params.domain_end == pytest.approx((1.08, 0.68, 0.68))
    assert sc.params.grid_num == (27, 17, 17)
    assert np.array_equal(sc.x, scenes.dam_break().x)  # seeded
    f = sc.x[sc.kind == KIND_FLUID]
    lat = scenes.lattice((0.06, 0.06, 0.06), (24, 24, 24))
    assert np.abs(f - lat).max() <= 0.01 * 0.02 * 1.0001  # +-1% jitter


def test_big_scene_counts():
    sc = scenes.nonnewton_dam()
    assert sc.n_fluid == 131072 and sc.params.viscosity_method == VISCOSITY_WEILER
    sc = scenes.elastic_drop()
    assert sc.n_elastic == 98304 and sc.n_static == 40000 and sc.params.youngs_modulus == 1e5


def test_mixed_batch_ranges():
    b = scenes.mixed_batch()
    assert len(b) == 64
    dyn = np.array([s.n_dynamic for s in b])
    assert dyn.min() >= 2000 and dyn.max() <= 60000
    assert 0.1 <= min(s.meta["eta0"] for s in b) and max(s.meta["eta0"] for s in b) <= 100
    assert {s.meta["model"] for s in b} == {0, 1, 2}
    assert 8 <= sum(s.meta["elastic"] for s in b) <= 24


def test_serving_requests():
    r = scenes.serving_requests(count=8)
    assert all(50 <= q.steps <= 500 for q in r)
    assert all(r[k].arrival < r[k + 1].arrival for k in range(7))


def test_ply_roundtrip_and_reference_fixture(tmp_path):
    pts = np.load(os.path.join(GOLDEN, "test_neighbor_points.npy"))
    assert pts.shape == (9, 3) and pts[4].tolist() == pytest.approx([0.03, 0.03, 0.0])
    p = tmp_path / "a.ply"
    ply.export_ply_particles(pts, str(p))
    assert np.array_equal(ply.read_ply_particles(str(p)), pts.astype(np.float32))


def test_params_from_reference_scene_files():
    with open(os.path.join(GOLDEN, "scene_test_nonNewton.json")) as f:
        p, unknown = params_from_config(json.load(f))
    assert p.use_non_newtonian and p.non_newton_method == 1 and p.viscosity_method == VISCOSITY_WEILER
    assert p.viscosity0 == 10.0 and p.consistency_index == 4.5 and p.weiler_tol == 0.001
    assert unknown == []
    with open(os.path.join(GOLDEN, "scene_test_deformable.json")) as f:
        p, unknown = params_from_config(json.load(f))
    assert p.domain_end == (1.0, 1.0, 1.0)  # "domain_end" is a typo the reference ignores too
    assert p.particle_radius == 0.025 and p.max_error == 0.01 and p.youngs_modulus == 250000.0
    with open(os.path.join(GOLDEN, "scene_dambreak.json")) as f:
        p, _ = params_from_config(json.load(f))
    assert p.dt == 0.004 and p.gravity == (0.0, -9.81, 0.0) and p.domain_end == (2.0, 2.0, 2.0)
