"""Pins the oracle's Becker 2009 elasticity (SPH/main.py:1182-1375): the polar
decomposition against SciPy/SVD, and the rest, rigid-rotation and uniform-stretch
states."""
import os

import numpy as np
import pytest
from scipy.linalg import polar as sp_polar

from nonnewton_amd.params import KIND_ELASTIC, SceneParams
from oracle import oracle as O

from conftest import GOLDEN


def proper_rotation(a):
    u, _, vt = np.linalg.svd(a)
    dm = np.diag([1.0, 1.0, np.linalg.det(u @ vt)])
    return u @ dm @ vt


@pytest.mark.parametrize("seed", range(20))
def test_polar_matches_scipy_and_svd(seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(3, 3))
    r64 = O.polar(a, "f64")
    if np.linalg.det(a) > 0:
        assert np.allclose(r64, sp_polar(a)[0], atol=1e-10)
    assert np.allclose(r64, proper_rotation(a), atol=1e-10)
    r32 = O.polar(a, "f32")
    assert np.allclose(r32, proper_rotation(a), atol=2e-5)
    assert np.linalg.det(r64) == pytest.approx(1.0, abs=1e-12)


def test_polar_degenerate_inputs():
    assert np.allclose(O.polar(np.zeros((3, 3)), "f64"), np.eye(3))
    a = np.outer([1.0, 2.0, 0.5], [0.3, -1.0, 2.0])  # rank 1
    r = O.polar(a, "f64")
    assert np.allclose(r.T @ r, np.eye(3), atol=1e-12) and np.linalg.det(r) == pytest.approx(1.0)
    u = a @ np.array([0.3, -1.0, 2.0])
    assert np.allclose(r @ (np.array([0.3, -1.0, 2.0]) / np.linalg.norm([0.3, -1.0, 2.0])), u / np.linalg.norm(u),
                       atol=1e-10)


def block():
    pts = np.load(os.path.join(GOLDEN, "test_deformable_points.npy")).astype(np.float64)
    p = SceneParams(particle_radius=0.025, domain_end=(1.0, 1.0, 1.0), gravity=(0, 0, 0), dt=1e-4,
                    youngs_modulus=1e5, poisson_ratio=0.33)
    return p, pts


def run_posed(pose, precision="f64"):
    p, pts = block()
    s = O.OracleSim(p, pts, np.full(len(pts), KIND_ELASTIC, np.int32), precision=precision)
    s.init()
    if pose is not None:
        s.set(O.F_X, pose(pts))
    s.step()
    return s, pts


def interior(pts):
    c = pts.mean(axis=0)
    return np.all(np.abs(pts - c) < 0.08, axis=1)


def test_rest_state_is_stress_free():
    s, pts = run_posed(None)
    R = s.get(O.F_ROTATION).reshape(-1, 3, 3)
    assert np.abs(R - np.eye(3)).max() < 1e-12
    assert np.abs(s.get(O.F_STRESS)).max() < 1e-6
    assert np.abs(s.get(O.F_ACC)).max() < 1e-6


def test_rigid_rotation_is_stress_free():
    th = 0.4
    Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    c = block()[1].mean(axis=0)
    s, pts = run_posed(lambda x: (x - c) @ Q.T + c)
    R = s.get(O.F_ROTATION).reshape(-1, 3, 3)
    assert np.abs(R - Q).max() < 1e-10
    assert np.abs(s.get(O.F_STRESS)).max() < 1e-6


def test_uniform_stretch_gives_uniform_strain():
    e = 1e-3
    c = block()[1].mean(axis=0)
    s, pts = run_posed(lambda x: (x - c) * (1 + e) + c)
    R = s.get(O.F_ROTATION).reshape(-1, 3, 3)
    assert np.abs(R - np.eye(3)).max() < 1e-10
    inner = interior(pts)
    F = s.get(O.F_F).reshape(-1, 3, 3)[inner]
    # F = grad u + I with grad u = e I up to the kernel's discrete consistency
    assert np.allclose(F, (1 + e) * np.eye(3), atol=0.1 * e)
    p = block()[0]
    lam = p.youngs_modulus * p.poisson_ratio / ((1 + p.poisson_ratio) * (1 - 2 * p.poisson_ratio))
    mu = p.youngs_modulus / (2 * (1 + p.poisson_ratio))
    sxx = s.get(O.F_STRESS)[inner][:, 0]
    assert np.allclose(sxx, (3 * lam + 2 * mu) * e, rtol=0.1)
