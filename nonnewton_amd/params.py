"""Scene parameters with the reference's keys and defaults.

Every field mirrors a value the reference reads from the scene JSON or hard-codes:
``Parameter.__init__`` (SPH/main.py:598-621), ``DFSPHSolver.__init__`` (:1520-1544),
``NonNewton.initParameters`` (:358-371), ``SurfaceTension`` (:1380-1382) and
``Elasticity.__init__`` (:1186-1189).  Fields the reference lacks (the Weiler solver
tolerance it declares but never uses, :1480-1481; the viscosity clamp; Young's modulus,
which it hard-codes) are marked where they are defined.
"""
from __future__ import annotations

import dataclasses
import json
import math
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Tuple

FLUID = 0  # SPH/main.py:27
SOLID = 1  # SPH/main.py:28
STATIC = 0  # SPH/main.py:38
RIGID = 1
ELASTIC = 2

# particle kinds understood by the C-ABI (material + solid type folded into one int)
KIND_FLUID = 0
KIND_STATIC = 1
KIND_ELASTIC = 2
KIND_RIGID = 3  # KIND_RIGID + b: a particle of rigid body b

NON_NEWTON_METHODS = {
    "NEWTONIAN": 0,
    "POWER_LAW": 1,
    "CROSS": 2,
    "CASSON": 3,
    "CARREAU": 4,
    "BINGHAM": 5,
    "HERSCHEL_BULKLEY": 6,
}

VISCOSITY_EXPLICIT = 0  # ViscosityModel, SPH/main.py:1417-1468
VISCOSITY_WEILER = 1  # Viscosity_Weiler2018, SPH/main.py:1477-1513 (stub there)


@dataclass
class SceneParams:
    # Parameter, SPH/main.py:598-621
    domain_start: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    domain_end: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    particle_radius: float = 0.01
    density0: float = 1000.0
    gravity: Tuple[float, float, float] = (0.0, -9.8, 0.0)
    boundary_viscosity: float = 0.0
    dt: float = 1e-4
    sticky_coeff: float = 0.999
    collision_coeff: float = 0.5
    viscosity: float = 0.01
    # DFSPHSolver.__init__, SPH/main.py:1524-1535
    max_iterations_v: int = 100
    max_iterations: int = 100
    max_error_v: float = 0.1
    max_error: float = 0.05
    eps: float = 1e-5
    enable_divergence_solver: bool = True
    use_surface_tension: bool = True
    use_viscosity: bool = True
    use_non_newtonian: bool = False
    viscosity_method: int = VISCOSITY_EXPLICIT
    non_newton_method: int = 0
    # SurfaceTension, SPH/main.py:1382 (hard-coded there)
    surface_tension: float = 0.01
    # NonNewton.initParameters, SPH/main.py:359-366
    power_index: float = 0.667
    consistency_index: float = 100.0
    viscosity0: float = 2000.0
    viscosity_inf: float = 1.0
    critical_strain_rate: float = 20.0
    mu_c: float = 10.0
    yield_stress: float = 200.0
    # not in the reference: clamp for models that diverge at zero strain rate
    max_viscosity: float = 1.0e6
    # Viscosity_Weiler2018.__init__, SPH/main.py:1480-1481 (m_maxIter, m_maxError)
    weiler_max_iter: int = 100
    compensated: bool = True
    weiler_tol: float = 0.01
    # Elasticity.__init__, SPH/main.py:1188-1189 (hard-coded 25e4 / 0.33 there)
    youngs_modulus: float = 25e4
    poisson_ratio: float = 0.33
    # <= 0 means volume0 * 1000 as in initialize_particles, SPH/main.py:743
    particle_mass: float = 0.0
    # bit 0: the Weiler CG starts from v + (x - v) of the previous solve (the
    # initial guess of the published solver; the reference declares its
    # m_vDiff fields, SPH/main.py:1486-1487, with an empty solver around them)
    warm_start: int = 1
    # rigid-body substeps (ground collision + shape matching) per step,
    # Parameter.coupling_interval, SPH/main.py:616 ("couplingInterval")
    coupling_interval: int = 1

    # ---- derived values, as the reference computes them in Python double ----
    @property
    def support_radius(self) -> float:
        return self.particle_radius * 4.0  # :606

    @property
    def particle_diameter(self) -> float:
        return 2.0 * self.particle_radius  # :605

    @property
    def volume0(self) -> float:
        return 0.8 * self.particle_diameter ** 3  # :607

    @property
    def mass(self) -> float:
        return self.particle_mass if self.particle_mass > 0 else self.volume0 * 1000.0

    @property
    def padding(self) -> float:
        return self.support_radius  # :615

    @property
    def grid_num(self) -> Tuple[int, int, int]:
        size = [e - s for s, e in zip(self.domain_start, self.domain_end)]
        return tuple(int(math.ceil(v / self.support_radius)) for v in size)  # :773-775

    def replace(self, **kw) -> "SceneParams":
        return dataclasses.replace(self, **kw)

    def as_dict(self) -> Dict[str, Any]:
        return dataclasses.asdict(self)


# ordered C layout shared by nn_params (include/nnsph.h) and the oracle's or_params
C_FIELDS: List[Tuple[str, str]] = [
    ("domain_start", "d3"),
    ("domain_end", "d3"),
    ("particle_radius", "d"),
    ("density0", "d"),
    ("gravity", "d3"),
    ("boundary_viscosity", "d"),
    ("dt", "d"),
    ("sticky_coeff", "d"),
    ("collision_coeff", "d"),
    ("viscosity", "d"),
    ("max_iterations_v", "i"),
    ("max_iterations", "i"),
    ("max_error_v", "d"),
    ("max_error", "d"),
    ("eps", "d"),
    ("enable_divergence_solver", "i"),
    ("use_surface_tension", "i"),
    ("use_viscosity", "i"),
    ("use_non_newtonian", "i"),
    ("viscosity_method", "i"),
    ("non_newton_method", "i"),
    ("surface_tension", "d"),
    ("power_index", "d"),
    ("consistency_index", "d"),
    ("viscosity0", "d"),
    ("viscosity_inf", "d"),
    ("critical_strain_rate", "d"),
    ("mu_c", "d"),
    ("yield_stress", "d"),
    ("max_viscosity", "d"),
    ("weiler_max_iter", "i"),
    ("compensated", "i"),
    ("weiler_tol", "d"),
    ("youngs_modulus", "d"),
    ("poisson_ratio", "d"),
    ("particle_mass", "d"),
    ("warm_start", "i"),
    ("coupling_interval", "i"),
]


def make_ctypes_struct(name: str):
    """A ctypes.Structure class with the C_FIELDS layout."""
    import ctypes

    fields = []
    for fname, kind in C_FIELDS:
        if kind == "d3":
            fields.append((fname, ctypes.c_double * 3))
        elif kind == "d":
            fields.append((fname, ctypes.c_double))
        else:
            fields.append((fname, ctypes.c_int32))
    return type(name, (ctypes.Structure,), {"_fields_": fields})


def fill_struct(struct, p: SceneParams):
    for fname, kind in C_FIELDS:
        val = getattr(p, fname)
        if kind == "d3":
            arr = getattr(struct, fname)
            for k in range(3):
                arr[k] = float(val[k])
        elif kind == "d":
            setattr(struct, fname, float(val))
        else:
            setattr(struct, fname, int(val))
    return struct


# ---------------------------------------------------------------------------
# JSON scene files (SimConfig.get_cfg, SPH/main.py:59-96)
# ---------------------------------------------------------------------------
_CONFIG_KEYS = {
    "domainStart": "domain_start",
    "domainEnd": "domain_end",
    "particleRadius": "particle_radius",
    "density0": "density0",
    "gravitation": "gravity",
    "boundaryViscosity": "boundary_viscosity",
    "timeStepSize": "dt",
    "stickyCoefficient": "sticky_coeff",
    "collisionCoefficient": "collision_coeff",
    "viscosity": "viscosity",
    "maxIterationsV": "max_iterations_v",
    "maxIterations": "max_iterations",
    "maxErrorV": "max_error_v",
    "maxError": "max_error",
    "useSurfaceTensionModel": "use_surface_tension",
    "useNonNewtonianModel": "use_non_newtonian",
    "useViscosityModel": "use_viscosity",
    "enableDivergenceSolver": "enable_divergence_solver",
    "couplingInterval": "coupling_interval",
}
# keys the reference reads but that do not change the step this build runs
_CONFIG_KEYS_HOST = {
    "simulationMethod",
    "solidTimeStepSize",
    "useDragForceModel",
    "useElasticityModel",
    "exportPly",
    "noGUI",
    "maxStep",
    "numberOfStepsPerRenderUpdate",
    "showWidget",
    "animeFile",
}
# keys present in the reference's scene files that its code ignores
_CONFIG_KEYS_IGNORED = {
    "boundaryHandlingMethod",
    "cflMethod",
    "cflFactor",
    "cflMaxTimeStepSize",
    "stiffness",
    "exponent",
    "velocityUpdateMethod",
    "domain_end",
}
_FLUID_KEYS = {
    "nonNewtonMethod": "non_newton_method",
    "power_index": "power_index",
    "consistency_index": "consistency_index",
    "viscosity0": "viscosity0",
    "viscosity_inf": "viscosity_inf",
    "criticalStrainRate": "critical_strain_rate",
    "muC": "mu_c",
    "yieldStress": "yield_stress",
}


def params_from_config(cfg: Dict[str, Any], strict_unknown: bool = False) -> Tuple[SceneParams, List[str]]:
    """SceneParams from a reference scene dict; returns (params, unknown keys)."""
    conf = cfg.get("Configuration", {})
    kw: Dict[str, Any] = {}
    unknown: List[str] = []
    for k, v in conf.items():
        if k in _CONFIG_KEYS:
            name = _CONFIG_KEYS[k]
            if name in ("domain_start", "domain_end", "gravity"):
                v = tuple(float(t) for t in v)
            kw[name] = v
        elif k in _CONFIG_KEYS_HOST or k in _CONFIG_KEYS_IGNORED:
            continue
        else:
            unknown.append(k)
    fluids = cfg.get("FluidParticles", [])
    if fluids:  # NonNewton reads the first fluid entry only, :335-337
        for k, name in _FLUID_KEYS.items():
            if k in fluids[0]:
                kw[name] = fluids[0][k]
        vm = fluids[0].get("viscosityMethod")
        if vm is not None and int(vm) == 13:  # SPlisHSPlasH's Weiler 2018 id
            kw["viscosity_method"] = VISCOSITY_WEILER
        if "viscoMaxError" in fluids[0]:
            kw["weiler_tol"] = float(fluids[0]["viscoMaxError"])
        if "viscoMaxIter" in fluids[0]:
            kw["weiler_max_iter"] = int(fluids[0]["viscoMaxIter"])
    for s in cfg.get("SolidParticles", []):
        if s.get("solidType") == "ELASTIC":
            if "youngsModulus" in s:
                kw["youngs_modulus"] = float(s["youngsModulus"])
            if "poissonsRatio" in s:
                kw["poisson_ratio"] = float(s["poissonsRatio"])
    if strict_unknown and unknown:
        raise KeyError(f"unknown configuration keys: {unknown}")
    return SceneParams(**kw), unknown


def load_scene_json(path: str) -> Dict[str, Any]:
    with open(path, "r") as f:
        return json.load(f)
