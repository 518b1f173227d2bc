"""Scene ingestion pinned by closed forms: fill_tank (SPH/main.py:147-181,
trimesh's voxelization of the tank box restated) and transform (:133-144,
rotation about the point set's centre, then translation)."""
import math

import numpy as np
import pytest

from nonnewton_amd import sph
from nonnewton_amd.params import SceneParams


def with_domain(start, end, r=0.01):
    sph.meta.parm = sph.Parameter(SceneParams(domain_start=start, domain_end=end, particle_radius=r))


def test_fill_tank_counts_and_extents():
    """Box [0, 1]^3, r = 0.01: pad h = 0.04, pitch 0.02; the tank is the
    multiples of 0.02 from 0.04 to 0.96 in x and z and to the height 0.30 in
    y: 47 x 14 x 47 = 30,926 particles, x slowest, z fastest."""
    with_domain((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    pts = sph.fill_tank(0.30)
    assert pts.shape == (47 * 14 * 47, 3)
    assert pts.min(axis=0) == pytest.approx([0.04, 0.04, 0.04])
    assert pts.max(axis=0) == pytest.approx([0.96, 0.30, 0.96])
    k = pts / 0.02
    assert np.abs(k - np.round(k)).max() < 1e-9  # voxel centres on the pitch lattice
    assert np.allclose(pts[1] - pts[0], [0.0, 0.0, 0.02]) and np.allclose(pts[47] - pts[0], [0.0, 0.02, 0.0])
    assert len(np.unique(np.round(k).astype(int), axis=0)) == len(pts)
    # volume: one particle per pitch^3 of the (voxel-closed) box
    vol = (0.96 - 0.04 + 0.02) * (0.30 - 0.04 + 0.02) * (0.96 - 0.04 + 0.02)
    assert len(pts) == pytest.approx(vol / 0.02 ** 3, rel=1e-9)


def test_fill_tank_snaps_to_the_pitch_lattice():
    """A domain start that is not a multiple of the pitch: the voxel centres
    are still multiples of the pitch, the box's faces rounded to the nearest."""
    with_domain((0.013, 0.0, 0.0), (0.5, 0.5, 0.5))
    pts = sph.fill_tank(0.2)
    k = pts / 0.02
    assert np.abs(k - np.round(k)).max() < 1e-9
    assert pts[:, 0].min() == pytest.approx(0.06)  # round(0.053 / 0.02) = 3
    assert pts[:, 0].max() == pytest.approx(0.46)  # round(0.46 / 0.02) = 23


def rot(angle_deg, axis):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    t = math.radians(angle_deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * K @ K


def test_transform_rotates_about_the_centre_then_translates():
    rng = np.random.default_rng(3)
    pts = rng.uniform(0.2, 0.4, (50, 3))
    c = pts.mean(axis=0)
    out = sph.transform(pts, {"rotationAngle": 90.0, "rotationAxis": [0, 0, 2], "translation": [1.0, 0.0, -0.5]})
    # closed form: a quarter turn about z maps (dx, dy, dz) to (-dy, dx, dz)
    d = pts - c
    expect = c + np.stack([-d[:, 1], d[:, 0], d[:, 2]], axis=1) + [1.0, 0.0, -0.5]
    assert np.allclose(out, expect, atol=1e-12)
    # any axis: centre kept, distances kept, R (p - c) + c
    out2 = sph.transform(pts, {"rotationAngle": 33.0, "rotationAxis": [1, 2, 3]})
    assert np.allclose(out2.mean(axis=0), c, atol=1e-12)
    assert np.allclose(np.linalg.norm(out2[1:] - out2[0], axis=1), np.linalg.norm(pts[1:] - pts[0], axis=1))
    assert np.allclose(out2, (pts - c) @ rot(33.0, [1, 2, 3]).T + c, atol=1e-12)


def test_transform_without_rotation_or_translation():
    pts = np.arange(12, dtype=float).reshape(4, 3)
    assert np.array_equal(sph.transform(pts, {}), pts)
    assert np.array_equal(sph.transform(pts, {"rotationAngle": 45.0}), pts)  # no axis: no rotation
    assert np.allclose(sph.transform(pts, {"translation": [1, 2, 3]}), pts + [1, 2, 3])
