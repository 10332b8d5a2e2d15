"""Initial conditions (protocol of initial_conditions/base.py:28-31).

``ExplosionInitialCondition`` is the reference example (examples/explosion/initial_condition.py:
32-61).  ``SquareExplosionIC`` generalises its box, states and pressure ratio for batched cases;
``ObliqueShockIC`` is the double-Mach-reflection-style state split by a line at 60 degrees.
"""
from __future__ import annotations

from abc import ABC, abstractmethod

import numpy as np

from .states import ConservativeState, PrimitiveState


class InitialCondition(ABC):
    @abstractmethod
    def apply_to_block(self, block):
        raise NotImplementedError("Abstract Initial Condition")


def _cons(fluid, rho, u, v, p):
    return PrimitiveState(fluid=fluid, array=np.array([rho, u, v, p], dtype=np.float64).reshape((1, 1, 4))).to_type(
        ConservativeState)


class ExplosionInitialCondition(InitialCondition):
    """High pressure square [3,7]^2 (rho 4.6968, p 404400) in still air (rho 1.1742, p 101100)."""

    def apply_to_block(self, block):
        f = block.config.fluid
        left = _cons(f, 4.6968, 0.0, 0.0, 404400.0)
        right = _cons(f, 1.1742, 0.0, 0.0, 101100.0)
        xc = np.logical_and(block.mesh.x >= 3, block.mesh.x <= 7)
        yc = np.logical_and(block.mesh.y >= 3, block.mesh.y <= 7)
        block.state.data = np.where(np.logical_and(xc, yc), left.data, right.data)
        block.state.make_non_dimensional()


class SquareExplosionIC(InitialCondition):
    """Square [x0,x1]x[y0,y1] at pressure ratio*p_low and density density_ratio*rho_low, still air
    (rho_low, p_low) elsewhere."""

    def __init__(self, x0, x1, y0, y1, ratio, rho_low=1.1742, p_low=101100.0, density_ratio=None):
        self.box = (x0, x1, y0, y1)
        self.ratio = ratio
        self.density_ratio = ratio if density_ratio is None else density_ratio
        self.low = (rho_low, p_low)

    def apply_to_block(self, block):
        f = block.config.fluid
        rl, pl = self.low
        inside = _cons(f, self.density_ratio * rl, 0.0, 0.0, self.ratio * pl)
        outside = _cons(f, rl, 0.0, 0.0, pl)
        x0, x1, y0, y1 = self.box
        xc = np.logical_and(block.mesh.x >= x0, block.mesh.x <= x1)
        yc = np.logical_and(block.mesh.y >= y0, block.mesh.y <= y1)
        block.state.data = np.where(np.logical_and(xc, yc), inside.data, outside.data)
        block.state.make_non_dimensional()


class ObliqueShockIC(InitialCondition):
    """Post-shock state where x < x_shock + y/tan(60 deg), pre-shock state elsewhere."""

    def __init__(self, x_shock=1.0 / 6.0, post=(8.0, 8.25 * np.cos(np.pi / 6), -8.25 * np.sin(np.pi / 6), 116.5),
                 pre=(1.4, 0.0, 0.0, 1.0)):
        self.x_shock = x_shock
        self.post = post
        self.pre = pre

    def apply_to_block(self, block):
        f = block.config.fluid
        post = _cons(f, *self.post)
        pre = _cons(f, *self.pre)
        behind = block.mesh.x < self.x_shock + block.mesh.y / np.sqrt(3.0)
        block.state.data = np.where(behind, post.data, pre.data)
        block.state.make_non_dimensional()
