"""The reference's object surface (nonnewton_amd.sph mirrors SPH/main.py) driving
the HIP step: initialize / build_solver / solver.initialize / solver.step and
meta.pd / meta.ns / meta.parm, read like the reference's own usage."""
import json
import os
import shutil

import numpy as np
import pytest

from nonnewton_amd import ply, scenes
from nonnewton_amd import sph
from nonnewton_amd.sph import meta
from oracle import oracle as O

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def test_surface_step_matches_oracle():
    sc = scenes.tiny_box()
    sph.initialize(scene=sc)
    solver = sph.build_solver()
    solver.initialize()
    orc = O.OracleSim.from_scene(sc)
    orc.init()
    nbr, cnt = orc.neighbors()
    assert np.array_equal(meta.ns.neighbors.to_numpy(), nbr)
    assert np.array_equal(meta.ns.num_neighbors.to_numpy(), cnt)
    for _ in range(2):
        solver.step()
        orc.step()
    assert np.array_equal(meta.pd.x.to_numpy().astype(np.float64), orc.get(O.F_X))
    assert np.array_equal(meta.pd.v.to_numpy().astype(np.float64), orc.get(O.F_V))
    assert np.array_equal(meta.pd.density.to_numpy().astype(np.float64), orc.get(O.F_DENSITY))
    assert meta.ns.get_num_neighbors(0) == cnt[0]
    meta.parm.viscosity[None] = 0.05
    assert meta.parm.viscosity[None] == 0.05
    solver.step()
    assert solver.last_stats["num_pairs"] > 0


def test_scene_json_run(tmp_path):
    """main() on a reference-format scene with a PLY geometry file."""
    root = tmp_path / "SPH"
    (root / "data" / "models").mkdir(parents=True)
    (root / "data" / "scenes").mkdir(parents=True)
    pts = np.load(os.path.join(GOLDEN, "test_deformable_points.npy"))
    ply.export_ply_particles(pts, str(root / "data" / "models" / "test_deformable.ply"))
    shutil.copyfile(os.path.join(GOLDEN, "scene_test_deformable.json"), root / "data" / "scenes" / "d.json")
    solver = sph.main(str(root / "data" / "scenes" / "d.json"), max_step=3)
    assert meta.num_elastic_bodies == 1
    assert meta.particle_max_num == 343
    R = meta.elastic_bodies[0].m_rotations.to_numpy()
    assert R.shape == (343, 3, 3)
    assert np.abs(R - np.eye(3)).max() < 1e-3
    x = meta.pd.x.to_numpy()
    assert np.all(x[:, 1] < pts[:, 1].max() + 1e-6)  # falling


def test_rigid_coupling_scene_json_run(tmp_path):
    """main() on the reference's test_oneway_coupling.json (a fluid cube above a
    RIGID cube) with lattice stand-ins for its cube PLY files: one rigid body in
    meta.rbs, kinds in the reference's phase order, and the positions after 5
    steps equal to the oracle's bit for bit."""
    root = tmp_path / "SPH"
    (root / "data" / "models").mkdir(parents=True)
    (root / "data" / "scenes").mkdir(parents=True)
    fluid = scenes.lattice((0.26, 1.26, 0.26), (12, 12, 12)).astype(np.float32)
    body = scenes.lattice((0.26, 0.26, 0.26), (12, 12, 12)).astype(np.float32)
    ply.export_ply_particles(fluid, str(root / "data" / "models" / "cube2.ply"))
    ply.export_ply_particles(body, str(root / "data" / "models" / "cube.ply"))
    cfg_path = root / "data" / "scenes" / "c.json"
    shutil.copyfile(os.path.join(GOLDEN, "scene_test_oneway_coupling.json"), cfg_path)
    sph.main(str(cfg_path), max_step=5)
    assert len(meta.rbs) == 1 and meta.rbs[0].body == 0
    n_f = len(fluid)
    assert np.array_equal(meta.pd.kind, np.r_[np.zeros(n_f), np.full(len(body), 3)].astype(np.int32))
    with open(cfg_path) as f:
        params, _ = sph.params_from_config(json.load(f))
    assert params.coupling_interval == 1
    orc = O.OracleSim(params, np.concatenate([fluid, body]), meta.pd.kind)
    orc.init()
    for _ in range(5):
        orc.step()
    assert np.array_equal(meta.pd.x.to_numpy().astype(np.float64), orc.get(O.F_X))
    assert np.array_equal(meta.pd.v.to_numpy().astype(np.float64), orc.get(O.F_V))
