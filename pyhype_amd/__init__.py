"""pyhype_amd -- MI355X (gfx950) backend for pyHype's second-order MUSCL residual and explicit
Runge-Kutta time stepping, behind the reference's Euler2D / SolverConfig API.

The compute runs in hand-written HIP kernels (``lib/libpyhype_hip.so``, C ABI in
``include/pyhype_hip.h``) called through ctypes; PyTorch only carries device buffers.
"""
from .fluids import Air, Fluid
from .initial_conditions import (DoubleMachReflectionIC, ExplosionInitialCondition, InitialCondition,
                                 ObliqueShockIC, SquareExplosionIC)
from .mesh import BlockGeometry, QuadMeshGenerator, RectagularMeshGenerator
from .solver_config import SolverConfig
from .solvers import Euler2D, Euler2DBatch
from .states import ConservativeState, PrimitiveState, RealizabilityError
from .time_marching import butcher_table

__all__ = [
    "Air", "Fluid", "DoubleMachReflectionIC", "ExplosionInitialCondition", "InitialCondition", "ObliqueShockIC",
    "SquareExplosionIC",
    "BlockGeometry", "QuadMeshGenerator", "RectagularMeshGenerator", "SolverConfig", "Euler2D", "Euler2DBatch",
    "ConservativeState", "PrimitiveState", "RealizabilityError", "butcher_table",
]
