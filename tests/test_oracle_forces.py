"""Pins the oracle's force terms against mathematics and against brute-force
NumPy statements of the reference's formulas (tests/sph_np.py):

- ViscosityModel (SPH/main.py:1417-1468): the fluid term converges to
  nu * Laplacian(v) (Monaghan's 2(d+2) form, with the 0.01 h^2 regularisation
  factor computed by quadrature) on fine lattices; the oracle equals the
  brute-force statement on jittered scenes, with and without the boundary
  term and its reaction on dynamic solids (:1446-1460); the reaction conserves
  the summed acceleration; the boundary term is Galilean invariant and drags.
- Weiler 2018 implicit viscosity (stub :1471-1513): the oracle's operator
  equals a dense assembly from the pair formula with the mean viscosity, and
  for a constant viscosity (b - A b) / dt equals rho0 / rho_i times the
  reference's explicit viscosity acceleration.
- Becker elasticity (:1334-1375): the forces conserve momentum under any
  deformation, and deep inside a block they equal one half of minus the
  gradient of the discrete strain energy 1/2 sum V_i sigma_i : eps_i (the
  reference's 0.5 factor, :1372), checked by finite differences.
"""
import numpy as np
import pytest
from scipy import integrate

from nonnewton_amd.params import KIND_ELASTIC, KIND_FLUID, KIND_STATIC, VISCOSITY_WEILER, SceneParams
from oracle import oracle as O

import sph_np as S


def quiet(**kw):
    base = dict(domain_end=(1.0, 1.0, 1.0), gravity=(0.0, 0.0, 0.0), use_surface_tension=False)
    base.update(kw)
    return SceneParams(**base)


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


# ---------------------------------------------------------------------------
# explicit viscosity
# ---------------------------------------------------------------------------
def regularisation_factor(h):
    """int r^3 W'(r) r^2 / (r^2 + 0.01 h^2) dr / int r^3 W'(r) dr: what the
    0.01 h^2 term leaves of the continuum Laplacian."""
    eps = 0.01 * h * h
    num = integrate.quad(lambda r: r ** 5 * S.kernel_dw(r, h) / (r * r + eps), 0, h, points=[h / 2], epsabs=1e-16)[0]
    den = integrate.quad(lambda r: r ** 3 * S.kernel_dw(r, h), 0, h, points=[h / 2], epsabs=1e-16)[0]
    return num / den


def visc_at_center(spacing, n, field, r=0.01, nu=0.01):
    """sph_np's explicit viscosity of the lattice's central particle, its
    neighbors found by a KD-tree (the lattice is too fine for the oracle's
    60-neighbor cap; the oracle is tied to sph_np on coarse scenes below)."""
    from scipy.spatial import cKDTree

    x = S.lattice(n, spacing, 0.2)
    h = 4 * r
    V0 = 0.8 * (2 * r) ** 3
    m = V0 * 1000.0
    ctr = x.mean(axis=0)
    v = np.zeros_like(x)
    if field == "shear":
        v[:, 0] = 0.5 * (x[:, 1] - ctr[1]) ** 2  # Laplacian (1, 0, 0), divergence 0
    else:
        v[:, 0] = 0.5 * (x[:, 0] - ctr[0]) ** 2  # Laplacian (1, 0, 0), grad div (1, 0, 0)
    tree = cKDTree(x)

    def rho(i):
        js = np.array([j for j in tree.query_ball_point(x[i], h) if j != i])
        rn = np.linalg.norm(x[i] - x[js], axis=1)
        return 1000.0 * (V0 * S.kernel_w(0.0, h) + np.sum(V0 * S.kernel_w(rn, h)))

    ci = int(np.argmin(np.linalg.norm(x - ctr, axis=1)))
    a = np.zeros(3)
    for j in tree.query_ball_point(x[ci], h):
        if j == ci:
            continue
        rv = x[ci] - x[j]
        rn = np.linalg.norm(rv)
        a += 10 * nu * (m / rho(j)) * np.dot(v[ci] - v[j], rv) / (rn ** 2 + 0.01 * h * h) * S.kernel_grad(rv, h)
    return a / nu


def test_explicit_viscosity_converges_to_the_laplacian():
    """2(d+2) sum_j (m_j/rho_j) (v_ij . r_ij) / (|r|^2 + 0.01 h^2) gradW_ij
    -> c (Lap v + 2 grad div v) with c the regularisation factor (0.963): a
    wrong 2(d+2), a missing m_j / rho_j or a sign slip misses by >= 20%."""
    c = regularisation_factor(0.04)
    assert c == pytest.approx(0.96298, abs=1e-4)
    shear = visc_at_center(0.008, 21, "shear")
    comp = visc_at_center(0.008, 21, "compress")
    assert shear[0] == pytest.approx(c, rel=2e-3) and abs(shear[1]) < 1e-9 and abs(shear[2]) < 1e-9
    # the lattice's fourth angular moments are isotropic to 0.5% at h / spacing = 5
    assert comp[0] == pytest.approx(3 * c, rel=1e-2)


def jittered_block(n=6, seed=0, origin=0.3):
    rng = np.random.default_rng(seed)
    x = S.lattice(n, 0.02, origin) + rng.uniform(-2e-4, 2e-4, (n ** 3, 3))
    v = rng.normal(0, 0.1, (n ** 3, 3))
    return f32(x), f32(v)


def test_explicit_viscosity_equals_bruteforce():
    x, v = jittered_block()
    p = quiet()
    o = O.OracleSim(p, x, np.zeros(len(x), np.int32), v=v, precision="f64")
    a = o.force_probe(O.PROBE_VISCOSITY)
    ref = S.Scene(x, v, np.zeros(len(x), int), nu=p.viscosity).viscosity_acc()
    assert np.abs(a - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(ref).max() > 1.0


def fluid_wall_elastic(seed=1, wall_kind=KIND_STATIC):
    """A jittered fluid block resting on a one-layer wall lattice, with an
    elastic slab on its side, random velocities everywhere."""
    rng = np.random.default_rng(seed)
    fl = S.lattice((5, 4, 5), 0.02, 0.3) + rng.uniform(-2e-4, 2e-4, (100, 3))
    fl[:, 1] += 0.02
    wall = S.lattice((9, 1, 9), 0.02, 0.26)
    wall[:, 1] = 0.3
    el = S.lattice((2, 4, 5), 0.02, 0.3)
    el[:, 0] += 0.1
    el[:, 1] += 0.02
    x = np.concatenate([fl, wall, el])
    kind = np.array([KIND_FLUID] * len(fl) + [wall_kind] * len(wall) + [KIND_ELASTIC] * len(el), np.int32)
    v = rng.normal(0, 0.1, x.shape)
    v[kind == KIND_STATIC] = 0.0
    return f32(x), f32(v), kind


@pytest.mark.parametrize("bnu", [0.0, 0.05])
def test_boundary_viscosity_equals_bruteforce(bnu):
    """Fluid term, boundary term rho0 V_j / rho_i (:1448-1456) and its reaction
    a_j -= f_v on the dynamic (elastic) solid (:1458-1460)."""
    x, v, kind = fluid_wall_elastic()
    p = quiet(boundary_viscosity=bnu, particle_radius=0.01)
    o = O.OracleSim(p, x, kind, v=v, precision="f64")
    o.init()
    a = o.force_probe(O.PROBE_VISCOSITY)
    sk = np.where(kind == KIND_FLUID, 0, np.where(kind == KIND_ELASTIC, 2, 1))
    ref = S.Scene(x, v, sk, nu=p.viscosity, bnu=bnu).viscosity_acc()
    assert np.abs(a - ref).max() <= 1e-12 * np.abs(ref).max()
    elastic = kind == KIND_ELASTIC
    if bnu == 0.0:
        assert np.all(a[elastic] == 0.0)
    else:
        assert np.abs(a[elastic]).max() > 1e-3 * np.abs(a).max()


def test_boundary_viscosity_reaction_balances():
    """With the fluid term off (nu = 0) only fluid-solid pairs act; every pair's
    term is added to the fluid particle and subtracted from a dynamic solid, so
    the accelerations of fluid and elastic particles sum to zero."""
    x, v, kind = fluid_wall_elastic(seed=2)
    p = quiet(boundary_viscosity=0.05, viscosity=0.0)
    o = O.OracleSim(p, x, kind, v=v, precision="f64")
    o.init()
    a = o.force_probe(O.PROBE_VISCOSITY)
    el = kind == KIND_ELASTIC
    fl = kind == KIND_FLUID
    # static wall terms are not balanced (no reaction on static solids)
    xs, vs, ks = x[kind != KIND_STATIC], v[kind != KIND_STATIC], kind[kind != KIND_STATIC]
    o2 = O.OracleSim(p, xs, ks, v=vs, precision="f64")
    o2.init()
    a2 = o2.force_probe(O.PROBE_VISCOSITY)
    tot = a2[ks == KIND_FLUID].sum(axis=0) + a2[ks == KIND_ELASTIC].sum(axis=0)
    assert np.abs(tot).max() <= 1e-12 * np.abs(a2).sum()
    assert np.abs(a2[ks == KIND_ELASTIC]).max() > 0
    assert np.abs(a[fl]).max() > 0 and np.abs(a[el]).max() > 0


def test_boundary_viscosity_drags_and_is_galilean_invariant():
    """Fluid gliding at (u, 0, 0) over a wall: the boundary term opposes the
    motion and has no y or z part over the wall's interior; a wall moving with
    the fluid (an elastic layer at the same velocity) exerts nothing."""
    fl = S.lattice((5, 2, 5), 0.02, 0.4)
    fl[:, 1] = np.where(fl[:, 1] < 0.41, 0.32, 0.34)
    wall = S.lattice((13, 1, 13), 0.02, 0.32)
    wall[:, 1] = 0.3
    x = np.concatenate([fl, wall])
    for wall_kind, wall_v in ((KIND_STATIC, 0.0), (KIND_ELASTIC, 0.2)):
        kind = np.array([KIND_FLUID] * len(fl) + [wall_kind] * len(wall), np.int32)
        v = np.zeros_like(x)
        v[:, 0] = 0.2
        v[len(fl):, 0] = wall_v
        p = quiet(boundary_viscosity=0.05, viscosity=0.0)
        o = O.OracleSim(p, f32(x), kind, v=f32(v), precision="f64")
        o.init()
        a = o.force_probe(O.PROBE_VISCOSITY)[: len(fl)]
        centre = np.flatnonzero((np.abs(fl[:, 0] - 0.44) < 1e-9) & (np.abs(fl[:, 2] - 0.44) < 1e-9))
        if wall_kind == KIND_STATIC:
            near = fl[:, 1] < 0.33  # the top layer is beyond h of the wall
            assert np.all(a[near, 0] < 0) and np.all(a[~near] == 0.0)
            assert np.abs(a[centre, 1:]).max() <= 1e-9 * np.abs(a[centre, 0]).max()
        else:
            assert np.abs(a).max() == 0.0


# ---------------------------------------------------------------------------
# Weiler 2018 operator
# ---------------------------------------------------------------------------
def weiler_params(non_newton, **kw):
    base = dict(dt=5e-4, viscosity_method=VISCOSITY_WEILER, weiler_tol=1e-8, weiler_max_iter=500,
                use_non_newtonian=non_newton, non_newton_method=2, viscosity0=5.0, viscosity_inf=0.1,
                consistency_index=1.3, power_index=0.7, viscosity=0.05)
    base.update(kw)
    return quiet(**base)


def test_weiler_operator_equals_dense_assembly():
    """Cross model (viscosity varies per particle): the oracle's operator equals
    the dense pair-formula assembly with mu_ij = (mu_i + mu_j) / 2, mu = rho0 eta."""
    x, v = jittered_block(n=5, seed=3)
    p = weiler_params(True)
    o = O.OracleSim(p, x, np.zeros(len(x), np.int32), v=v, precision="f64")
    o.init()
    A, b, _, _ = o.weiler_probe()
    eta = o.get(O.F_VISCOSITY)
    assert eta.max() > 1.5 * eta.min()  # the viscosity does vary
    sc = S.Scene(x, v, np.zeros(len(x), int))
    ref = S.Scene.weiler_matrix(sc, p.dt, p.density0 * eta)
    off = np.abs(ref - np.eye(len(ref))).max()
    assert off > 1e-3
    assert np.abs(A - ref).max() <= 1e-11 * off


def test_weiler_operator_matches_explicit_viscosity():
    """Constant viscosity: (v - A v) / dt = rho0 / rho_i * a_visc, the
    reference's explicit viscosity acceleration (:1427-1445) with the same nu."""
    x, v = jittered_block(n=5, seed=4)
    p = weiler_params(False)
    o = O.OracleSim(p, x, np.zeros(len(x), np.int32), v=v, precision="f64")
    o.init()
    A, b, _, _ = o.weiler_probe()
    sc = S.Scene(x, v, np.zeros(len(x), int), nu=p.viscosity)
    acc = sc.viscosity_acc()
    lhs = ((b - A @ b) / p.dt).reshape(-1, 3)
    rhs = (p.density0 / sc.rho)[:, None] * acc
    assert np.abs(lhs - rhs).max() <= 1e-9 * np.abs(rhs).max()
    assert np.abs(rhs).max() > 0.1


# ---------------------------------------------------------------------------
# Becker elasticity
# ---------------------------------------------------------------------------
def elastic_block(n):
    pts = f32(S.lattice(n, 0.05, 0.15))
    p = quiet(particle_radius=0.025, dt=1e-4, youngs_modulus=1e5, poisson_ratio=0.33)
    o = O.OracleSim(p, pts, np.full(len(pts), KIND_ELASTIC, np.int32), precision="f64")
    o.init()
    return o, p, pts


@pytest.mark.parametrize("scale", [1e-3, 3e-3])
def test_elastic_forces_conserve_momentum(scale):
    """Each rest pair adds -V_i V_j (R_i sigma_i + R_j sigma_j) gradW0 / 2 to one
    particle and the opposite to the other, so the forces of a free block sum
    to zero under any deformation (a sign slip in dji or dij breaks this)."""
    o, p, pts = elastic_block(7)
    rng = np.random.default_rng(int(scale * 1e4))
    o.set(O.F_X, pts + rng.normal(0, scale, pts.shape))
    a = o.force_probe(O.PROBE_ELASTIC)
    assert np.abs(a.sum(axis=0)).max() <= 1e-12 * np.abs(a).sum()
    assert np.abs(a).max() > 1.0


def strain_energy(o, V):
    """1/2 sum_i V_i sigma_i : eps_i from the oracle's stress and F (Voigt
    shear entries counted twice in the double contraction)."""
    F = o.get(O.F_F).reshape(-1, 3, 3)
    sg = o.get(O.F_STRESS)
    U = F - np.eye(3)
    e = np.stack([U[:, 0, 0], U[:, 1, 1], U[:, 2, 2], 0.5 * (U[:, 0, 1] + U[:, 1, 0]),
                  0.5 * (U[:, 0, 2] + U[:, 2, 0]), 0.5 * (U[:, 1, 2] + U[:, 2, 1])], 1)
    dens = sg[:, :3] * e[:, :3]
    return 0.5 * np.sum(V * (dens.sum(axis=1) + 2 * np.sum(sg[:, 3:] * e[:, 3:], axis=1)))


def test_elastic_force_is_half_the_energy_gradient():
    """Near rest the corotated forces are the linear ones, f = -1/2 dE/dx
    (the reference halves the pair sum, :1372).  Checked by central
    differences at particles three or more layers inside a 13^3 block, where
    the rotation estimate adds nothing to the symmetric strain."""
    n = 13
    o, p, pts = elastic_block(n)
    V = o.get(O.F_REST_VOLUME)
    rng = np.random.default_rng(7)
    x0 = pts + rng.normal(0, 1e-5, pts.shape)
    o.set(O.F_X, x0)
    f0 = o.force_probe(O.PROBE_ELASTIC) * p.mass
    idx = np.arange(len(pts)).reshape(n, n, n)
    step = 1e-8
    for cell in [(6, 6, 6), (4, 4, 4), (3, 6, 9)]:
        i = idx[cell]
        for k in range(3):
            xe = []
            for sgn in (1, -1):
                xx = x0.copy()
                xx[i, k] += sgn * step
                o.set(O.F_X, xx)
                o.force_probe(O.PROBE_ELASTIC)
                xe.append(strain_energy(o, V))
            grad = (xe[0] - xe[1]) / (2 * step)
            assert f0[i, k] == pytest.approx(-0.5 * grad, rel=1e-3), (cell, k)
