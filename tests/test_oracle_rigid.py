"""Rigid bodies in the oracle (RigidBodySolver, SPH/main.py:1099-1158; the
coupling reactions of :1755-1760 and :1835-1841; advect :1086-1093), pinned by
what the mechanics fixes rather than by retyping the formulas:

  * a free-falling body keeps its shape and its centre follows the closed form
    of symplectic Euler, x_n = x_0 + g dt^2 n (n + 1) / 2;
  * shape matching of a rotated, perturbed body recovers the rotation (a
    transposed A would return its inverse) and removes the perturbation;
  * ground collision lifts every solid particle (static included) below
    y = padding and no fluid particle, and only when a rigid body exists;
  * coupling_interval 0 switches shape matching and ground collision off;
  * the pressure reactions on the body conserve the momentum of fluid + body.
"""
import numpy as np
import pytest

from nonnewton_amd import scenes
from nonnewton_amd.params import KIND_FLUID, KIND_RIGID, KIND_STATIC, SceneParams
from oracle import oracle as O

SP = scenes.SPACING


def _rot(axis, angle):
    a = np.asarray(axis, dtype=np.float64)
    a /= np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def _box(counts, origin, params, extra=None, extra_kind=KIND_STATIC):
    body = scenes.lattice(origin, counts)
    parts = [body] if extra is None else [extra, body]
    kinds = [np.full(len(body), KIND_RIGID, np.int32)]
    if extra is not None:
        kinds = [np.full(len(extra), extra_kind, np.int32)] + kinds
    x = np.concatenate(parts).astype(np.float32)
    return x, np.concatenate(kinds), body


def test_free_fall_keeps_shape_and_follows_closed_form():
    dt = 1e-3
    p = SceneParams(domain_end=(1.0, 2.0, 1.0), gravity=(0.0, -9.81, 0.0), dt=dt)
    x, kind, body = _box((4, 3, 5), (0.4, 1.5, 0.4), p)
    o = O.OracleSim(p, x, kind)
    o.init()
    c0, _, _ = o.rigid_frame(0)
    assert np.allclose(c0, body.mean(axis=0), atol=1e-6)
    n = 30
    for _ in range(n):
        o.step()
    xn = o.get(O.F_X)
    g = np.array([0.0, -9.81, 0.0])
    expect = body.mean(axis=0) + g * dt * dt * n * (n + 1) / 2.0
    assert np.allclose(xn.mean(axis=0), expect, atol=2e-5)
    # pairwise distances of the body are those of the rest shape
    d = np.linalg.norm(xn[:, None, :] - xn[None, :, :], axis=-1)
    d0 = np.linalg.norm(body[:, None, :] - body[None, :, :], axis=-1)
    assert np.max(np.abs(d - d0)) < 2e-6


def test_shape_matching_recovers_rotation_and_removes_noise():
    p = SceneParams(domain_end=(1.0, 1.0, 1.0), gravity=(0.0, 0.0, 0.0), dt=1e-3)
    x, kind, body = _box((5, 3, 2), (0.3, 0.4, 0.45), p)  # no symmetry that hides a transpose
    o = O.OracleSim(p, x, kind)
    o.init()
    c0, _, _ = o.rigid_frame(0)
    R0 = _rot((0.3, -0.5, 0.8), 0.7)
    centre = np.array([0.52, 0.47, 0.49])
    q = body - c0
    rng = np.random.Generator(np.random.PCG64(11))
    noise = rng.uniform(-2e-3, 2e-3, size=q.shape)
    posed = centre + q @ R0.T + noise
    o.set(O.F_X, posed)
    o.step()
    _, cm, R = o.rigid_frame(0)
    # the least-squares rotation (Kabsch: SVD of sum (x - cm)(x0 - cm0)^T in FP64)
    P = posed - posed.mean(axis=0)
    U, _, Vt = np.linalg.svd(P.T @ q)
    D = np.diag([1.0, 1.0, np.sign(np.linalg.det(U @ Vt))])
    R_ls = U @ D @ Vt
    assert np.allclose(R, R_ls, atol=2e-5)
    assert np.allclose(R, R0, atol=2e-2)            # the posed rotation, up to the noise ...
    assert not np.allclose(R, R0.T, atol=0.1)       # ... and not its inverse (a transposed A)
    assert np.allclose(np.linalg.det(R), 1.0, atol=1e-5)
    m_noise = noise.mean(axis=0)
    assert np.allclose(cm, centre + m_noise, atol=1e-6)
    xs = o.get(O.F_X)
    rigid_pose = cm + q @ R.T
    assert np.max(np.abs(xs - rigid_pose)) < 1e-6   # the projection lands on cm + R q
    assert np.max(np.abs(xs - posed)) > 1e-3        # and it moved the particles off the noisy pose


def test_ground_collision_lifts_solids_only_and_only_with_a_body():
    p = SceneParams(domain_start=(0.0, -0.5, 0.0), domain_end=(1.0, 1.0, 1.0), gravity=(0.0, 0.0, 0.0), dt=1e-3)
    h = np.float32(p.support_radius)
    low = np.array([[0.2, 0.01, 0.2], [0.8, 0.015, 0.8]], dtype=np.float32)  # fluid, static, both below y = h
    body = scenes.lattice((0.45, 0.5, 0.45), (3, 3, 3))
    x = np.concatenate([low, body]).astype(np.float32)
    kind = np.concatenate([[KIND_FLUID, KIND_STATIC], np.full(len(body), KIND_RIGID)]).astype(np.int32)
    o = O.OracleSim(p, x, kind)
    o.init()
    o.step()
    xs = o.get(O.F_X)
    assert xs[1, 1] == h                  # the static particle is lifted to exactly y = padding
    assert xs[0, 1] < 0.02                # the fluid particle is not (the clamp is at y = -0.46)
    # without a rigid body nothing lifts the static particle
    kind2 = np.concatenate([[KIND_FLUID, KIND_STATIC], np.full(len(body), KIND_STATIC)]).astype(np.int32)
    o2 = O.OracleSim(p, x, kind2)
    o2.init()
    o2.step()
    assert o2.get(O.F_X)[1, 1] == np.float32(0.015)


def test_coupling_interval_zero_switches_shape_matching_off():
    p = SceneParams(domain_end=(1.0, 1.0, 1.0), gravity=(0.0, 0.0, 0.0), dt=1e-3, coupling_interval=0)
    x, kind, body = _box((4, 3, 2), (0.3, 0.4, 0.45), p)
    o = O.OracleSim(p, x, kind)
    o.init()
    rng = np.random.Generator(np.random.PCG64(12))
    posed = body + rng.uniform(-2e-3, 2e-3, size=body.shape)
    o.set(O.F_X, posed)
    o.step()
    assert np.array_equal(o.get(O.F_X), posed.astype(np.float32).astype(np.float64))
    p1 = p.replace(coupling_interval=1)
    o1 = O.OracleSim(p1, x, kind)
    o1.init()
    o1.set(O.F_X, posed)
    o1.step()
    assert np.max(np.abs(o1.get(O.F_X) - posed)) > 1e-4


def test_pressure_reactions_conserve_momentum():
    """Fluid pressed against a body with gravity off and no divergence solve,
    surface tension or viscosity: the body's reactions (:1835-1841) are equal
    and opposite to the fluid's velocity changes from the body, so the total
    momentum of fluid and body stays put."""
    p = SceneParams(domain_end=(1.0, 1.0, 1.0), gravity=(0.0, 0.0, 0.0), dt=1e-3,
                    enable_divergence_solver=0, use_surface_tension=0, use_viscosity=0)
    rng = np.random.Generator(np.random.PCG64(13))
    fluid = scenes.jitter(rng, scenes.lattice((0.3, 0.3, 0.3), (8, 8, 8)), frac=0.2)
    body = scenes.lattice((0.3, 0.3 + 8 * SP, 0.3), (8, 4, 8))  # resting on the fluid, overlapping its top
    x = np.concatenate([fluid, body]).astype(np.float32)
    kind = np.concatenate([np.full(len(fluid), KIND_FLUID), np.full(len(body), KIND_RIGID)]).astype(np.int32)
    v = np.zeros_like(x)
    v[: len(fluid), 1] = 0.5  # the fluid moves into the body
    o = O.OracleSim(p, x, kind, v)
    o.init()
    o.step()
    vs = o.get(O.F_V)
    nf = len(fluid)
    p_fluid = vs[:nf].sum(axis=0)
    p_body = vs[nf:].sum(axis=0)
    assert p_body[1] > 100.0                                # the body was pushed up (sum of 256 velocities) ...
    assert np.allclose(p_fluid + p_body, v.sum(axis=0), atol=1e-4)   # ... by what the fluid lost


def test_bodies_must_be_numbered_without_gaps():
    p = SceneParams(domain_end=(1.0, 1.0, 1.0))
    x = scenes.lattice((0.4, 0.4, 0.4), (2, 2, 2)).astype(np.float32)
    kind = np.full(len(x), KIND_RIGID + 1, np.int32)  # body 1 without a body 0
    with pytest.raises(ValueError):
        O.OracleSim(p, x, kind)
