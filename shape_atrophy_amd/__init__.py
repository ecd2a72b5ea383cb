"""MI355X-native forward shape-evolution engine (layered P1 meshes, frame /
isotropic elasticity, RKHS velocity solve, reaction-diffusion, varifold)."""
from . import mesh, workloads  # noqa: F401
