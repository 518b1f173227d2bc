"""MI355X-native SPH step (DFSPH, non-Newtonian implicit viscosity, Becker
elasticity) behind the object surface of chunleili/nonNewton SPH/main.py."""
from .params import SceneParams  # noqa: F401

__all__ = ["SceneParams"]
