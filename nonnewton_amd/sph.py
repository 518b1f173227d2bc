"""The object surface of chunleili/nonNewton SPH/main.py over the HIP step.

A caller of the reference keeps its code: ``initialize()``, ``build_solver()``,
``solver.initialize()``, ``solver.step()``, and the state on the ``meta``
singleton -- ``meta.pd.<field>.to_numpy() / from_numpy()``, ``meta.ns.neighbors``,
``meta.ns.run_search()``, ``meta.parm.dt[None]``, ``meta.parm.viscosity[None]``
(README.md:25-33, SPH/main.py:19-23).  Behind the surface every call goes to
libnnsph.so through ctypes; there is no host or CPU fallback.

Differences from the reference, by design:
  * ``SimConfig`` takes the scene path (no tkinter dialog, SPH/main.py:47-57).
  * Rigid bodies (solidType RIGID) run the reference's RigidBodySolver
    (:1099-1158) inside ``solver.step()``; ``meta.rbs`` lists them.  The unused
    RigidBody class (:498-591) is not mirrored.
  * Per-step ``print`` output (:916, :1548, :1704, :1785) is kept in
    ``solver.last_stats`` instead.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass
from typing import Any, Dict, List, Optional

import numpy as np

from . import _lib
from . import ply as _ply
from .params import (ELASTIC, FLUID, KIND_ELASTIC, KIND_FLUID, KIND_RIGID, KIND_STATIC, RIGID, SOLID, STATIC,
                     SceneParams, params_from_config)

RED = (1, 0, 0)
GREEN = (0, 1, 0)
BLUE = (0, 0, 1)
ORANGE = (1, 0.5, 0)
WHITE = (1, 1, 1)
BLACK = (0, 0, 0)
YELLOW = (1, 1, 0)


class Meta:
    ...


meta = Meta()
meta.sph_root_path = os.getcwd()


# ---------------------------------------------------------------------------
# scene configuration (SimConfig / get_cfg, SPH/main.py:47-96)
# ---------------------------------------------------------------------------
class SimConfig:
    def __init__(self, scene_path: str = "", config: Optional[Dict[str, Any]] = None) -> None:
        if config is not None:
            self.config = config
        else:
            if not scene_path:
                raise ValueError("SimConfig needs a scene path (no file dialog here)")
            with open(scene_path, "r") as f:
                self.config = json.load(f)
            meta.sph_root_path = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(scene_path))))

    def get_cfg(self, name, default=None):
        return self.config.get("Configuration", {}).get(name, default)


def _config():
    if not hasattr(meta, "config"):
        raise RuntimeError("no scene configuration: call initialize(scene_path) first")
    return meta.config


def get_cfg(name, default=None):
    return _config().get_cfg(name, default)


def get_fluid_cfg():
    return _config().config.get("FluidParticles", [])


def get_solid_cfg():
    return _config().config.get("SolidParticles", [])


def get_cfg_list(name, default=None):
    return _config().config.get(name, default)


@dataclass
class PhaseInfo:  # SPH/main.py:99-109
    uid: int = 0
    parnum: int = 0
    startnum: int = 0
    material: int = FLUID
    is_dynamic: bool = False
    cfg: dict = None
    color: tuple = WHITE
    pos: np.ndarray = None
    solid_type: int = STATIC


def read_ply_particles(geometry_file: str) -> np.ndarray:
    return _ply.read_ply_particles(geometry_file)


def _rotation_matrix(angle_rad: float, axis) -> np.ndarray:
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    c, s = math.cos(angle_rad), math.sin(angle_rad)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) * c + s * K + (1 - c) * np.outer(a, a)


def transform(points: np.ndarray, cfg: dict) -> np.ndarray:
    """Rotation by rotationAngle degrees about rotationAxis through the point
    set's centroid, then translation (transform, SPH/main.py:133-144).  The
    reference rotates about trimesh's center_mass of a face-less mesh of the
    points; the centroid is taken here (that point is not pinned by any
    reference output)."""
    pts = np.asarray(points, dtype=np.float64)
    ang, axis = cfg.get("rotationAngle"), cfg.get("rotationAxis")
    if ang is not None and axis is not None:
        c = pts.mean(axis=0)
        pts = (pts - c) @ _rotation_matrix(math.radians(ang), axis).T + c
    t = cfg.get("translation")
    if t is not None:
        pts = pts + np.asarray(t, dtype=np.float64)
    return pts


def fill_tank(height: float) -> np.ndarray:
    """The particles filling the box [domain_start + pad, domain_end - pad] up
    to `height` (fill_tank, SPH/main.py:147-181).  The reference voxelizes the
    box mesh at pitch 2r with trimesh and fills it (points_from_volume,
    :115-118: mesh.voxelized(pitch).fill().points); trimesh's subdivision
    voxelizer rounds surface points to the nearest multiple of the pitch, so
    the filled voxel centres are the multiples k * pitch per axis with k from
    round(lo / pitch) to round(hi / pitch), x slowest and z fastest."""
    p = meta.parm
    lo = np.asarray(p.domain_start, dtype=np.float64) + p.padding
    hi = np.asarray(p.domain_end, dtype=np.float64) - p.padding
    hi = np.array([hi[0], height, hi[2]])
    pitch = p.particle_diameter
    k0 = np.round(lo / pitch).astype(np.int64)
    k1 = np.round(hi / pitch).astype(np.int64)
    axes = [np.arange(k0[k], k1[k] + 1, dtype=np.float64) * pitch for k in range(3)]
    g = np.meshgrid(*axes, indexing="ij")
    return np.stack([a.ravel() for a in g], axis=1)


def color_selector(cfg, default=WHITE):
    return {"RED": RED, "GREEN": GREEN, "BLUE": BLUE, "ORANGE": ORANGE, "WHITE": WHITE, "BLACK": BLACK,
            "YELLOW": YELLOW}.get(cfg.get("color", None), default) if "color" in cfg else default


def solid_type_selector(cfg, default=STATIC):
    return {"STATIC": STATIC, "RIGID": RIGID, "ELASTIC": ELASTIC}.get(cfg.get("solidType"), STATIC) \
        if "solidType" in cfg else default


# ---------------------------------------------------------------------------
# fields backed by the device context
# ---------------------------------------------------------------------------
class DeviceField:
    """A particle field; ``to_numpy`` downloads it in original particle order."""

    def __init__(self, field_id: Optional[int], n: int, width: int = 1, dtype=np.float32, host: Optional[np.ndarray] = None):
        self.field_id = field_id
        self.n = n
        self.width = width
        self.dtype = dtype
        self._host = host
        self.shape = (n,) if width == 1 else (n, width)

    def _ctx(self):
        return getattr(meta, "ctx", None)

    def to_numpy(self) -> np.ndarray:
        ctx = self._ctx()
        if self.field_id is not None and ctx is not None and ctx.n:
            a = ctx.download(self.field_id)
            if self.field_id == _lib.FIELD_ROTATION:
                return a.reshape(-1, 3, 3)
            return a
        if self._host is None:
            return np.zeros(self.shape, dtype=self.dtype)
        return self._host.copy()

    def from_numpy(self, arr) -> None:
        a = np.asarray(arr, dtype=self.dtype).reshape(self.shape)
        ctx = self._ctx()
        if ctx is not None and ctx.n and self.field_id in (_lib.FIELD_X, _lib.FIELD_V):
            ctx.upload_field(self.field_id, a)
        else:
            self._host = a.copy()

    def fill(self, value) -> None:
        self.from_numpy(np.broadcast_to(np.asarray(value, dtype=self.dtype), self.shape))

    def __getitem__(self, idx):
        return self.to_numpy()[idx]


class ScalarField:
    """0-d field written as ``field[None]`` (meta.parm.dt / viscosity, :611-612,
    :620-621); writes reach the device through nn_set_param."""

    def __init__(self, key: str, value: float):
        self.key = key
        self.value = float(value)

    def __getitem__(self, idx):
        return self.value

    def __setitem__(self, idx, value):
        self.value = float(value)
        ctx = getattr(meta, "ctx", None)
        if ctx is not None:
            ctx.set_param(self.key, self.value)
        if hasattr(meta, "scene_params"):
            meta.scene_params = meta.scene_params.replace(**{self.key: self.value})


class Parameter:
    """Scene parameters (Parameter, SPH/main.py:594-621)."""

    def __init__(self, params: Optional[SceneParams] = None):
        if params is None:
            params, unknown = params_from_config(_config().config)
            self.unknown_keys = unknown
        self.params = params
        p = params
        self.domain_start = np.array(p.domain_start)
        self.domain_end = np.array(p.domain_end)
        self.domain_size = self.domain_end - self.domain_start
        self.dim = len(self.domain_size)
        self.simulation_method = 4
        self.particle_radius = p.particle_radius
        self.particle_diameter = p.particle_diameter
        self.support_radius = p.support_radius
        self.volume0 = p.volume0
        self.density0 = p.density0
        self.gravity = np.array(p.gravity)
        self.boundary_viscosity = p.boundary_viscosity
        self.dt = ScalarField("dt", p.dt)
        self.sticky_coeff = p.sticky_coeff
        self.collision_coeff = p.collision_coeff
        self.padding = p.padding
        self.viscosity = ScalarField("viscosity", p.viscosity)


def load_particles():
    """Phases in the reference's order: fluid, tanks, static, rigid, elastic
    (load_particles, SPH/main.py:628-666)."""
    counter = {"n": 0}

    def parse(cfg_i, cnt, phase_id_start, material, solid_type, color, is_dynamic, pos=None):
        if pos is None:
            pos = read_ply_particles(os.path.normpath(os.path.join(meta.sph_root_path, cfg_i["geometryFile"])))
        pos = transform(pos, cfg_i)
        phase_id = cfg_i.get("id", cnt) + phase_id_start
        meta.phase_info[phase_id] = PhaseInfo(uid=phase_id, parnum=pos.shape[0], startnum=counter["n"],
                                              material=material, cfg=cfg_i, is_dynamic=bool(is_dynamic), pos=pos,
                                              color=color_selector(cfg_i, color), solid_type=solid_type)
        counter["n"] += pos.shape[0]
        return cnt + 1

    cnt = 0
    for c in get_fluid_cfg():
        cnt = parse(c, cnt, 0, FLUID, STATIC, BLUE, 1)
    for c in get_cfg_list("FluidTanks", []) or []:
        cnt = parse(c, cnt, 0, FLUID, STATIC, BLUE, 1, fill_tank(c["height"]))
    fluid = counter["n"]
    solids = get_solid_cfg()
    for st, start, color, dyn in ((STATIC, 1000, WHITE, 0), (RIGID, 2000, ORANGE, 1), (ELASTIC, 3000, YELLOW, 1)):
        cnt = 0
        for c in solids:
            if solid_type_selector(c) == st:
                cnt = parse(c, cnt, start, SOLID, st, color, dyn)
    return counter["n"], fluid


class ParticleData:
    """Particle fields (ParticleData, SPH/main.py:697-723)."""

    def __init__(self, particle_max_num: int):
        n = particle_max_num
        self.particle_max_num = n
        self.dim = 3
        F = _lib
        self.x = DeviceField(F.FIELD_X, n, 3)
        self.x_0 = DeviceField(F.FIELD_X0, n, 3)
        self.v = DeviceField(F.FIELD_V, n, 3)
        self.acceleration = DeviceField(F.FIELD_ACCELERATION, n, 3)
        self.volume = DeviceField(F.FIELD_VOLUME, n)
        self.m = DeviceField(None, n)
        self.density = DeviceField(F.FIELD_DENSITY, n)
        self.pressure = DeviceField(None, n)
        self.material = DeviceField(None, n, dtype=np.int32)
        self.color = DeviceField(None, n, 3)
        self.is_dynamic = DeviceField(None, n, dtype=np.int32)
        self.particle_id = DeviceField(None, n, dtype=np.int32)
        self.phase_id = DeviceField(None, n, dtype=np.int32)
        self.dfsph_factor = DeviceField(F.FIELD_DFSPH_FACTOR, n)
        self.density_adv = DeviceField(F.FIELD_DENSITY_ADV, n)
        self.kind = np.zeros(n, dtype=np.int32)


def initialize_particles(pd: ParticleData):
    """initialize_particles, SPH/main.py:735-758."""
    all_par = np.concatenate([ph.pos for ph in meta.phase_info.values()], axis=0).astype(np.float32)
    pd.x.from_numpy(all_par)
    pd.x_0.from_numpy(all_par)
    pd.v.fill(0.0)
    pd.volume.fill(meta.parm.volume0)
    pd.m.fill(meta.parm.volume0 * 1000.0)
    pd.density.fill(meta.parm.density0)
    pd.pressure.fill(0.0)
    n = pd.particle_max_num
    material = np.full(n, FLUID, np.int32)
    dyn = np.ones(n, np.int32)
    phase = np.zeros(n, np.int32)
    color = np.ones((n, 3), np.float32)
    body = 0  # rigid bodies are numbered in phase order, the order of meta.rbs (:1882-1887)
    for ph in meta.phase_info.values():
        s, e = ph.startnum, ph.startnum + ph.parnum
        material[s:e] = ph.material
        dyn[s:e] = int(ph.is_dynamic)
        phase[s:e] = ph.uid
        color[s:e] = ph.color
        if ph.material == FLUID:
            pd.kind[s:e] = KIND_FLUID
        elif ph.solid_type == ELASTIC:
            pd.kind[s:e] = KIND_ELASTIC
        elif ph.solid_type == RIGID:
            pd.kind[s:e] = KIND_RIGID + body
            body += 1
        else:
            pd.kind[s:e] = KIND_STATIC
    pd.material.from_numpy(material)
    pd.is_dynamic.from_numpy(dyn)
    pd.phase_id.from_numpy(phase)
    pd.color.from_numpy(color)
    pd.particle_id.from_numpy(np.arange(n, dtype=np.int32))


class _NeighborField:
    def __init__(self, which: str):
        self.which = which

    def to_numpy(self):
        nbr, cnt = meta.ctx.neighbors(_lib.MAX_NEIGHBORS)
        return nbr if self.which == "neighbors" else cnt

    def __getitem__(self, idx):
        return self.to_numpy()[idx]


class NeighborhoodSearchSparse:
    """Uniform-grid neighbor search (NeighborhoodSearchSparse, SPH/main.py:764-853),
    counting sort + compact lists on the GPU."""

    def __init__(self, positions, num_particle, support_radius, domain_size, use_sparse_grid=False):
        self.num_particle = num_particle
        self.support_radius = support_radius
        self.domain_size = np.asarray(domain_size)
        self.grid_size = support_radius
        self.grid_num = np.ceil(self.domain_size / self.grid_size).astype(int)
        self.max_num_neighbors = _lib.MAX_NEIGHBORS
        self.max_num_particles_in_grid = 50
        self.neighbors = _NeighborField("neighbors")
        self.num_neighbors = _NeighborField("num_neighbors")

    def run_search(self):
        meta.ctx.search()

    def get_num_neighbors(self, i):
        return int(self.num_neighbors.to_numpy()[i])

    def get_neighbor(self, i, j):
        return int(self.neighbors.to_numpy()[i, j])


class NonNewton:
    """Non-Newtonian viscosity (NonNewton, SPH/main.py:332-492) -- computed on the
    GPU inside the step; this object exposes its fields and min/max/avg."""

    def __init__(self, num_particles):
        self.num_particles = num_particles
        self.strainRateNorm = DeviceField(_lib.FIELD_STRAIN_NORM, num_particles)
        self.nonNewtonViscosity = DeviceField(_lib.FIELD_VISCOSITY, num_particles)
        self.maxViscosity = self.minViscosity = self.avgViscosity = 0.0

    def step(self):
        eta = self.nonNewtonViscosity.to_numpy()[meta.pd.kind == KIND_FLUID]
        if eta.size:
            self.maxViscosity, self.minViscosity, self.avgViscosity = float(eta.max()), float(eta.min()), float(eta.mean())


class Elasticity:
    """Becker 2009 elasticity (Elasticity, SPH/main.py:1182-1375), on the GPU."""

    def __init__(self, numParticles):
        p = meta.scene_params
        self.m_youngsModulus = p.youngs_modulus
        self.m_poissonRatio = p.poisson_ratio
        self.numParticles = numParticles
        self.max_num_neighbors = 50
        self.m_rotations = DeviceField(_lib.FIELD_ROTATION, numParticles, 9)
        self.m_stress = DeviceField(_lib.FIELD_STRESS, numParticles, 6)
        self.m_restVolumes = DeviceField(_lib.FIELD_REST_VOLUME, numParticles)


class RigidBodySolver:
    """One rigid body (RigidBodySolver, SPH/main.py:1099-1158).  Its substep --
    ground collision and shape matching, coupling_interval times -- runs on the
    GPU inside ``solver.step()``, after elasticity and before advection."""

    def __init__(self, phase_id, body):
        self.phase_id = phase_id
        self.body = body  # the body's number in the particle kinds (KIND_RIGID + body)


class SPHBase:
    def __init__(self):
        self.gravity = meta.parm.gravity
        self.density_0 = meta.parm.density0
        self.dt = meta.parm.dt
        self.last_stats: Dict[str, Any] = {}

    def initialize(self):
        """run_search + static and moving boundary volumes (SPH/main.py:957-960),
        with the elastic rest state (:1216-1236)."""
        meta.ctx.init()

    def step(self):
        """One step (SPHBase.step, SPH/main.py:915-940)."""
        st = meta.ctx.step(1)
        self.last_stats = st
        meta.step_num = getattr(meta, "step_num", 0) + 1
        if getattr(meta, "export_ply", False):
            _ply.export_ply_particles(meta.pd.x.to_numpy(), f"results/{meta.step_num - 1}.ply")
        return st

    def substep(self):
        raise NotImplementedError("the substep runs fused inside step() on the GPU")


class DFSPHSolver(SPHBase):
    """Divergence-free SPH (DFSPHSolver, SPH/main.py:1519-1848)."""

    def __init__(self):
        super().__init__()
        p = meta.scene_params
        self.enable_divergence_solver = p.enable_divergence_solver
        self.m_max_iterations_v = p.max_iterations_v
        self.m_max_iterations = p.max_iterations
        self.m_eps = p.eps
        self.max_error_V = p.max_error_v
        self.max_error = p.max_error
        self.num_particles = meta.particle_max_num
        self.use_surfaceTensionModel = p.use_surface_tension
        self.use_nonNewtonianModel = p.use_non_newtonian
        self.use_viscosityModel = p.use_viscosity
        if self.use_nonNewtonianModel:
            self.nonNewtonianModel = NonNewton(self.num_particles)


def build_solver():
    """build_solver, SPH/main.py:196-203."""
    solver_type = get_cfg("simulationMethod") if hasattr(meta, "config") else 4
    if solver_type == 4:
        return DFSPHSolver()
    raise NotImplementedError(f"Solver type {solver_type} has not been implemented.")


def initialize(scene_path: str = "", scene=None, params: Optional[SceneParams] = None):
    """initialize, SPH/main.py:1870-1895.  With `scene` (a scenes.Scene) the particle
    arrays come from it instead of geometry files."""
    old = getattr(meta, "ctx", None)
    if old is not None:
        old.close()
    meta.ctx = None
    meta.phase_info = {}
    if scene is not None:
        meta.config = SimConfig(config={"Configuration": {"simulationMethod": 4}})
        meta.scene_params = params or scene.params
        meta.parm = Parameter(meta.scene_params)
        n = scene.n
        meta.particle_max_num, meta.fluid_particle_num = n, scene.n_fluid
        meta.pd = ParticleData(n)
        meta.pd.x.from_numpy(scene.x)
        meta.pd.x_0.from_numpy(scene.x)
        meta.pd.v.from_numpy(scene.v)
        meta.pd.kind[:] = scene.kind
        meta.pd.material.from_numpy(np.where(scene.kind == KIND_FLUID, FLUID, SOLID).astype(np.int32))
        meta.pd.is_dynamic.from_numpy((scene.kind != KIND_STATIC).astype(np.int32))
        meta.pd.particle_id.from_numpy(np.arange(n, dtype=np.int32))
        meta.pd.m.fill(meta.scene_params.mass)
    else:
        meta.config = SimConfig(scene_path)
        meta.parm = Parameter()
        meta.scene_params = meta.parm.params
        meta.particle_max_num, meta.fluid_particle_num = load_particles()
        meta.pd = ParticleData(meta.particle_max_num)
        initialize_particles(meta.pd)
        meta.export_ply = bool(get_cfg("exportPly", False))
    ctx = _lib.Context(meta.scene_params, meta.particle_max_num)
    x0, v0 = meta.pd.x.to_numpy(), meta.pd.v.to_numpy()
    # what init builds the rest state and static volumes from (snapshot/resume)
    meta.initial_state = (np.array(x0, np.float32), np.array(v0, np.float32), np.array(meta.pd.kind, np.int32))
    ctx.upload(x0, meta.pd.kind, v0)
    meta.ctx = ctx
    meta.ns = NeighborhoodSearchSparse(meta.pd.x, meta.particle_max_num, meta.parm.support_radius,
                                      meta.parm.domain_size, use_sparse_grid=True)
    n_body = int(meta.pd.kind.max()) - KIND_RIGID + 1 if meta.pd.kind.size and meta.pd.kind.max() >= KIND_RIGID else 0
    rigid_phases = [ph.uid for ph in getattr(meta, "phase_info", {}).values() if ph.solid_type == RIGID and ph.material == SOLID]
    meta.rbs = [RigidBodySolver(rigid_phases[b] if b < len(rigid_phases) else None, b) for b in range(n_body)]
    n_el = int(np.count_nonzero(meta.pd.kind == KIND_ELASTIC))
    meta.elastic_bodies = [Elasticity(meta.particle_max_num)] if n_el else []
    meta.num_elastic_bodies = len(meta.elastic_bodies)
    meta.step_num = 0


def snapshot(path: str) -> None:
    """Write what a run continues from to `path` (.npz): the scene parameters,
    the initial particle arrays (the rest state, the rest lists and the static
    boundary volumes are rebuilt from them by init, as in the original run), the
    current positions, velocities and Weiler warm-start direction, and the step
    count.  The reference has no snapshot; it writes per-step PLY positions
    (SPH/main.py:939-940), which carry no velocities."""
    import dataclasses

    x0, v0, kind = meta.initial_state
    np.savez(path, params=np.array(json.dumps(dataclasses.asdict(meta.scene_params))), x0=x0, v0=v0, kind=kind,
             x=meta.pd.x.to_numpy(), v=meta.pd.v.to_numpy(), vdiff=meta.ctx.download(_lib.FIELD_VISC_DIFF),
             step_num=np.array(getattr(meta, "step_num", 0)))


def resume(path: str):
    """Continue a run from snapshot(path): a new context is initialized on the
    snapshot's initial arrays, then given its positions, velocities and Weiler
    warm-start direction.  Returns the solver; its steps equal, bit for bit,
    those the original run would have taken."""
    from .scenes import Scene

    with np.load(path, allow_pickle=False) as z:
        fields = {k: z[k] for k in z.files}
    d = json.loads(str(fields["params"]))
    params = SceneParams(**{k: tuple(v) if isinstance(v, list) else v for k, v in d.items()})
    sc = Scene(name="resumed", params=params, x=fields["x0"], v=fields["v0"], kind=fields["kind"])
    initialize(scene=sc)
    solver = build_solver()
    solver.initialize()
    meta.ctx.upload_field(_lib.FIELD_X, fields["x"])
    meta.ctx.upload_field(_lib.FIELD_V, fields["v"])
    meta.ctx.upload_field(_lib.FIELD_VISC_DIFF, fields["vdiff"])
    meta.step_num = int(fields["step_num"])
    return solver


def main(scene_path: str, max_step: Optional[int] = None):
    """Headless run (SPH/main.py:1909-1921; the GUI loop is not built)."""
    initialize(scene_path)
    solver = build_solver()
    solver.initialize()
    steps = max_step if max_step is not None else get_cfg("maxStep", 1)
    for _ in range(int(steps)):
        solver.step()
    return solver
