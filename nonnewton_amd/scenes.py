"""Seeded synthetic scenes for the five workloads of the SPH step.

All generators use NumPy's PCG64 and return a :class:`Scene`: parameters plus
FP32 positions, velocities and particle kinds in the reference's particle order
(fluid, then static solids, rigid bodies, elastic solids; SPH/main.py:627-666).  Lattices
are filled with z fastest, then y, then x, the order of the reference's
``cube.ply``.  "Jitter +-1%" is a per-axis uniform offset of +-1% of the spacing.

Scenes (sizes in particles; spacing d = 2r = 0.02 m, h = 4r = 0.04 m):
  dam_break          24^3 = 13,824 fluid + 8,706 wall, dt 1e-3, explicit viscosity 0.01
  nonnewton_dam      64x32x64 = 131,072 fluid + walls, Cross model, Weiler CG, dt 5e-4
  elastic_drop       48x32x64 = 98,304 elastic + 40,000 plate, E 1e5, nu 0.33, dt 2e-4
  rigid_drop         24x12x24 = 6,912 fluid + an 8^3 rigid cube + walls, dt 5e-4
  mixed_batch        64 scenes, dynamic count log-uniform 2k-60k, seed 3
  serving_requests   Poisson request stream, 1k-500k particles, 50-500 steps, seed 4
  impact_block       8^3 fluid block thrown at the floor (the pressure solve iterates)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from .params import KIND_ELASTIC, KIND_FLUID, KIND_RIGID, KIND_STATIC, VISCOSITY_WEILER, SceneParams

SPACING = 0.02  # 2r with r = 0.01


@dataclass
class Scene:
    name: str
    params: SceneParams
    x: np.ndarray  # (N, 3) float32
    v: np.ndarray  # (N, 3) float32
    kind: np.ndarray  # (N,) int32
    steps: int = 1
    meta: Dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_fluid(self) -> int:
        return int(np.count_nonzero(self.kind == KIND_FLUID))

    @property
    def n_elastic(self) -> int:
        return int(np.count_nonzero(self.kind == KIND_ELASTIC))

    @property
    def n_static(self) -> int:
        return int(np.count_nonzero(self.kind == KIND_STATIC))

    @property
    def n_rigid(self) -> int:
        return int(np.count_nonzero(self.kind >= KIND_RIGID))

    @property
    def n_dynamic(self) -> int:
        return self.n_fluid + self.n_elastic + self.n_rigid


def lattice(origin, counts, spacing=SPACING) -> np.ndarray:
    """Regular lattice, z fastest, then y, then x (float64)."""
    nx, ny, nz = counts
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pts = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64) * spacing
    return pts + np.asarray(origin, dtype=np.float64)


def jitter(rng: np.random.Generator, pts: np.ndarray, spacing=SPACING, frac=0.01) -> np.ndarray:
    return pts + rng.uniform(-frac * spacing, frac * spacing, size=pts.shape)


def shell(domain_cells: Tuple[int, int, int], spacing=SPACING) -> np.ndarray:
    """One-layer wall lattice at k*spacing, k = 1..m-1 per axis, boundary layer only.

    With the domain [0, m*spacing]^3 and padding h = 2*spacing, the walls sit one
    spacing outside the clamp planes, so fluid at a clamp plane sees them."""
    mx, my, mz = domain_cells
    pts = lattice((spacing, spacing, spacing), (mx - 1, my - 1, mz - 1), spacing)
    idx = np.rint(pts / spacing).astype(np.int64)
    on = (idx[:, 0] == 1) | (idx[:, 0] == mx - 1) | (idx[:, 1] == 1) | (idx[:, 1] == my - 1) | (idx[:, 2] == 1) | (idx[:, 2] == mz - 1)
    return pts[on]


def assemble(name, params, fluid, static, elastic, steps, meta=None, rigid=None) -> Scene:
    """Particles in the reference's phase order (SPH/main.py:627-666): fluid,
    static, rigid bodies (`rigid`: a list of arrays, body b gets kind
    KIND_RIGID + b), elastic."""
    rigid = [r for r in (rigid or []) if len(r)]
    parts = [p for p in [fluid, static] + rigid + [elastic] if p is not None and len(p)]
    x = np.concatenate(parts, axis=0).astype(np.float32)
    kind = np.concatenate(
        [np.full(len(fluid) if fluid is not None else 0, KIND_FLUID, np.int32),
         np.full(len(static) if static is not None else 0, KIND_STATIC, np.int32)]
        + [np.full(len(r), KIND_RIGID + b, np.int32) for b, r in enumerate(rigid)]
        + [np.full(len(elastic) if elastic is not None else 0, KIND_ELASTIC, np.int32)]
    )
    v = np.zeros_like(x)
    return Scene(name=name, params=params, x=x, v=v, kind=kind, steps=steps, meta=meta or {})


def dam_break(seed: int = 0, fluid_counts=(24, 24, 24), steps: int = 100) -> Scene:
    """Plain DFSPH dam break: 24^3 jittered fluid in a 1.0 x 0.6 x 0.6 m box."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cells = (54, 34, 34)  # domain 1.08 x 0.68 x 0.68; clamp box 1.0 x 0.6 x 0.6
    p = SceneParams(
        domain_start=(0.0, 0.0, 0.0),
        domain_end=tuple(c * SPACING for c in cells),
        particle_radius=0.01,
        density0=1000.0,
        gravity=(0.0, -9.81, 0.0),
        dt=1e-3,
        viscosity=0.01,
        max_error=0.05,
        max_error_v=0.1,
    )
    fluid = jitter(rng, lattice((0.06, 0.06, 0.06), fluid_counts))
    walls = shell(cells)
    return assemble("dam_break", p, fluid, walls, None, steps)


def nonnewton_dam(seed: int = 1, fluid_counts=(64, 32, 64), steps: int = 200, cells=(134, 54, 134)) -> Scene:
    """Non-Newtonian dam break: Cross model (eta0 50, eta_inf 0.1, K 1, n 0.7),
    Weiler implicit viscosity (tol 1e-4, <= 100 iterations), 2.6 x 1.0 x 2.6 m box."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = SceneParams(
        domain_end=tuple(c * SPACING for c in cells),
        gravity=(0.0, -9.81, 0.0),
        dt=5e-4,
        max_error=0.1,
        max_error_v=0.1,
        use_non_newtonian=True,
        non_newton_method=2,
        viscosity0=50.0,
        viscosity_inf=0.1,
        consistency_index=1.0,
        power_index=0.7,
        viscosity_method=VISCOSITY_WEILER,
        weiler_tol=1e-4,
        weiler_max_iter=100,
    )
    fluid = jitter(rng, lattice((0.06, 0.06, 0.06), fluid_counts))
    walls = shell(cells)
    return assemble("nonnewton_dam", p, fluid, walls, None, steps)


def elastic_drop(seed: int = 2, block_counts=(48, 32, 64), plate_counts=(200, 200), steps: int = 500) -> Scene:
    """Becker elastic block dropped on a 200 x 200 static plate (E 1e5, nu 0.33)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    px, pz = plate_counts
    cells = (px + 4, 60, pz + 4)  # clamp box spans the plate; 1.2 m tall domain
    p = SceneParams(
        domain_end=tuple(c * SPACING for c in cells),
        gravity=(0.0, -9.81, 0.0),
        dt=2e-4,
        youngs_modulus=1e5,
        poisson_ratio=0.33,
    )
    plate = lattice((2 * SPACING, SPACING, 2 * SPACING), (px, 1, pz))
    bx, by, bz = block_counts
    ox = 2 * SPACING + 0.5 * (px - bx) * SPACING
    oz = 2 * SPACING + 0.5 * (pz - bz) * SPACING
    # bottom layer 0.05 m above the plate: farther than h, so no rest neighbor is a plate particle
    block = jitter(rng, lattice((ox, SPACING + 0.05, oz), block_counts))
    return assemble("elastic_drop", p, None, plate, block, steps)


def rigid_drop(seed: int = 5, fluid_counts=(24, 12, 24), cube=8, steps: int = 200, dt: float = 5e-4) -> Scene:
    """Two-way coupling (the reference's test_oneway_coupling.json layout): a
    jittered fluid pool on the floor of a walled box and a rigid cube of
    cube^3 particles released above it; explicit viscosity, coupling
    interval 1."""
    rng = np.random.Generator(np.random.PCG64(seed))
    fx, fy, fz = fluid_counts
    cells = (fx + 8, fy + cube + 14, fz + 8)
    p = SceneParams(
        domain_end=tuple(c * SPACING for c in cells),
        gravity=(0.0, -9.81, 0.0),
        dt=dt,
        coupling_interval=1,
    )
    fl = jitter(rng, lattice((4 * SPACING, 3 * SPACING, 4 * SPACING), fluid_counts))
    c0 = ((fx + 8 - cube) / 2.0 * SPACING, (fy + 6) * SPACING, (fz + 8 - cube) / 2.0 * SPACING)
    body = lattice(c0, (cube, cube, cube))
    walls = shell(cells)
    return assemble("rigid_drop", p, fl, walls, None, steps, meta={"cube": cube}, rigid=[body])


def _fluid_block_dims(n: int) -> Tuple[int, int, int]:
    a = max(4, int(round(n ** (1.0 / 3.0))))
    b = int(math.ceil(n / (a * a)))
    return a, b, a


def mixed_scene(rng: np.random.Generator, n_dyn: int, model: int, eta0: float, n_index: float,
                elastic: bool, name: str, dt: float = 5e-4, steps: int = 1) -> Scene:
    """One scene of the mixed batch: a jittered fluid column (dam break) in a walled
    box, optionally with an elastic block beside it, viscosity by the given model
    solved with the Weiler implicit solver."""
    n_el = int(round(0.2 * n_dyn)) if elastic else 0
    n_fl = n_dyn - n_el
    fa, fb, fc = _fluid_block_dims(n_fl)
    ea = eb = ec = 0
    if n_el:
        ea = max(3, int(round(n_el ** (1.0 / 3.0))))
        eb = int(math.ceil(n_el / (ea * ea)))
        ec = ea
    # domain in units of the spacing: 2 cells of padding on each side
    lx = 2 * fa + (ea + 6 if n_el else 0)
    ly = max(2 * fb, fb + 8, eb + 8)
    lz = max(fc, ec) + 4
    cells = (lx + 4, ly + 4, lz + 4)
    params = SceneParams(
        domain_end=tuple(c * SPACING for c in cells),
        gravity=(0.0, -9.81, 0.0),
        dt=dt,
        use_non_newtonian=True,
        non_newton_method=model,
        viscosity0=eta0,
        viscosity_inf=eta0 / 500.0 if model == 2 else 1.0,
        consistency_index=(eta0 if model == 1 else 1.0),
        power_index=(n_index if model == 1 else 0.7),
        viscosity_method=VISCOSITY_WEILER,
        weiler_tol=1e-4,
        weiler_max_iter=100,
        max_viscosity=1.0e4,
        youngs_modulus=1e5,
        poisson_ratio=0.33,
    )
    fl = lattice((3 * SPACING, 3 * SPACING, 3 * SPACING), (fa, fb, fc))[:n_fl]
    fl = jitter(rng, fl)
    el = None
    if n_el:
        ex = (3 + 2 * fa + 3) * SPACING
        el = lattice((ex, 3 * SPACING, 3 * SPACING), (ea, eb, ec))[:n_el]
        el = jitter(rng, el)
    walls = shell(cells)
    meta = {"model": model, "eta0": eta0, "n_index": n_index, "elastic": bool(n_el)}
    return assemble(name, params, fl, walls, el, steps, meta)


def mixed_batch(seed: int = 3, count: int = 64, lo: int = 2000, hi: int = 60000, dt: float = 5e-4) -> List[Scene]:
    """64 independent scenes: dynamic particle count log-uniform in [2k, 60k],
    viscosity model uniform over {Newtonian, Cross, power law n ~ U[0.5, 1.5]},
    eta0 log-uniform in [0.1, 100], 25% with an elastic block (20% of the count)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    scenes = []
    for k in range(count):
        n_dyn = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        model = [0, 2, 1][int(rng.integers(0, 3))]
        eta0 = float(math.exp(rng.uniform(math.log(0.1), math.log(100.0))))
        n_index = float(rng.uniform(0.5, 1.5))
        elastic = bool(rng.uniform() < 0.25)
        scenes.append(mixed_scene(rng, n_dyn, model, eta0, n_index, elastic, f"mixed_{k:02d}", dt=dt))
    return scenes


@dataclass
class Request:
    rid: int
    arrival: float  # seconds since the start of the stream
    scene: Scene
    steps: int


def serving_requests(seed: int = 4, rate: float = 20.0, count: int = 64, lo: int = 1000, hi: int = 500000,
                     steps_lo: int = 50, steps_hi: int = 500, dt: float = 5e-4) -> List[Request]:
    """Poisson request stream: exponential inter-arrival times at `rate` per second,
    sizes log-uniform in [1k, 500k] dynamic particles, 50-500 steps each,
    parameters drawn as in :func:`mixed_batch`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    reqs = []
    for k in range(count):
        t += float(rng.exponential(1.0 / rate))
        n_dyn = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        model = [0, 2, 1][int(rng.integers(0, 3))]
        eta0 = float(math.exp(rng.uniform(math.log(0.1), math.log(100.0))))
        n_index = float(rng.uniform(0.5, 1.5))
        elastic = bool(rng.uniform() < 0.25)
        steps = int(rng.integers(steps_lo, steps_hi + 1))
        sc = mixed_scene(rng, n_dyn, model, eta0, n_index, elastic, f"req_{k:03d}", dt=dt, steps=steps)
        reqs.append(Request(k, t, sc, steps))
    return reqs


def tiny_box(seed: int = 0, fluid_counts=(6, 6, 6), cells=(14, 14, 14), **kw) -> Scene:
    """A small walled fluid block for tests (216 fluid particles by default)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = SceneParams(domain_end=tuple(c * SPACING for c in cells), gravity=(0.0, -9.81, 0.0), dt=1e-3)
    if kw:
        p = p.replace(**kw)
    fluid = jitter(rng, lattice((3 * SPACING, 3 * SPACING, 3 * SPACING), fluid_counts))
    walls = shell(cells)
    return assemble("tiny_box", p, fluid, walls, None, 1)


def impact_block(r: float = 0.011, vy: float = -6.0, divergence: bool = False, max_error: float = 0.02,
                 max_iterations: int = 100, seed: int = 0) -> Scene:
    """An 8^3 jittered fluid block thrown at the floor of a walled box at vy m/s,
    its particle radius a little above half the 0.02 lattice spacing (the rest
    state is over-dense): the DFSPH pressure solve iterates every step (3-6
    iterations with the defaults, dt 1e-3)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cl = (16, 20, 16)
    p = SceneParams(domain_end=tuple(c * SPACING for c in cl), gravity=(0.0, -9.81, 0.0), dt=1e-3,
                    particle_radius=r, enable_divergence_solver=divergence, max_error=max_error,
                    max_iterations=max_iterations)
    fl = jitter(rng, lattice((3 * SPACING, 2.6 * SPACING, 3 * SPACING), (8, 8, 8)))
    sc = assemble("impact_block", p, fl, shell(cl), None, 1)
    sc.v[: len(fl), 1] = vy
    return sc


def by_name(name: str, **kw):
    return {
        "dam_break": dam_break,
        "nonnewton_dam": nonnewton_dam,
        "elastic_drop": elastic_drop,
        "rigid_drop": rigid_drop,
        "mixed_batch": mixed_batch,
        "serving_requests": serving_requests,
        "tiny_box": tiny_box,
        "impact_block": impact_block,
    }[name](**kw)
