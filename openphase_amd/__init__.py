"""MI355X-native multiphase-field time step with OpenPhase's DoubleObstacle / PhaseField /
InterfaceProperties / BoundaryConditions interface.

The numerics run in libmpf_hip.so (HIP kernels for gfx950 behind the C ABI of include/mpf.h).
"""
from .engine import Engine, StepConfig  # noqa: F401
from .microstructure import Lists, voronoi_grains  # noqa: F401

__all__ = ["Engine", "StepConfig", "Lists", "voronoi_grains"]
