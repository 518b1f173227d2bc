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


def test_main_exports_ply_every_step(tmp_path, monkeypatch):
    """main() with exportPly writes results/<step>.ply after every step
    (SPH/main.py:939-940, binary little-endian, vertex x y z float): each file
    reads back as the positions of that step, here a tank of fluid filled by
    FluidTanks (fill_tank, :147-181) beside a PLY block."""
    root = tmp_path / "SPH"
    (root / "data" / "models").mkdir(parents=True)
    (root / "data" / "scenes").mkdir(parents=True)
    block = scenes.lattice((0.3, 0.3, 0.3), (5, 5, 5))
    ply.export_ply_particles(block, str(root / "data" / "models" / "block.ply"))
    cfg = {"Configuration": {"domainStart": [0.0, 0.0, 0.0], "domainEnd": [0.6, 0.6, 0.6], "particleRadius": 0.01,
                             "density0": 1000, "simulationMethod": 4, "gravitation": [0.0, -9.81, 0.0],
                             "timeStepSize": 0.001, "exportPly": True, "maxStep": 3},
           "FluidParticles": [{"geometryFile": "./data/models/block.ply", "translation": [0.0, 0.1, 0.0]}],
           "FluidTanks": [{"height": 0.1}]}
    scene = root / "data" / "scenes" / "tank.json"
    scene.write_text(json.dumps(cfg))
    monkeypatch.chdir(tmp_path)
    solver = sph.main(str(scene))
    n = meta.particle_max_num
    assert n == 125 + 27 * 4 * 27  # block + tank: x, z multiples of 0.02 in [0.04, 0.56], y in [0.04, 0.10]
    files = sorted(os.listdir(tmp_path / "results"))
    assert files == ["0.ply", "1.ply", "2.ply"]
    with open(tmp_path / "results" / "2.ply", "rb") as f:
        head = f.read(200)
    assert head.startswith(b"ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % n)
    last = ply.read_ply_particles(str(tmp_path / "results" / "2.ply"))
    assert np.array_equal(last.astype(np.float32), meta.pd.x.to_numpy())
    # the first file: the positions after one step of the same scene
    sph.initialize(str(scene))
    s2 = sph.build_solver()
    s2.initialize()
    meta.export_ply = False
    s2.step()
    first = ply.read_ply_particles(str(tmp_path / "results" / "0.ply"))
    assert np.array_equal(first.astype(np.float32), meta.pd.x.to_numpy())


def test_snapshot_and_resume_equal_the_uninterrupted_run(tmp_path):
    """snapshot after 3 steps, 4 more steps; resume from the snapshot in a new
    context, 4 steps: positions and velocities equal bit for bit (a fluid with
    the Cross model and the warm-started Weiler CG, and an elastic block)."""
    rng = np.random.Generator(np.random.PCG64(17))
    sc = scenes.mixed_scene(rng, 3000, 2, 20.0, 1.0, True, "snap", dt=5e-4)
    sph.initialize(scene=sc)
    solver = sph.build_solver()
    solver.initialize()
    for _ in range(3):
        solver.step()
    path = str(tmp_path / "snap.npz")
    sph.snapshot(path)
    for _ in range(4):
        solver.step()
    x, v = meta.pd.x.to_numpy(), meta.pd.v.to_numpy()
    assert meta.step_num == 7
    solver = sph.resume(path)
    assert meta.step_num == 3
    for _ in range(4):
        solver.step()
    assert np.array_equal(meta.pd.x.to_numpy(), x)
    assert np.array_equal(meta.pd.v.to_numpy(), v)
