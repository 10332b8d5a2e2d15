"""Initial conditions (protocol of initial_conditions/base.py:28-31).

``ExplosionInitialCondition`` is the reference example (examples/explosion/initial_condition.py:
32-61).  ``SquareExplosionIC`` generalises its box, states and pressure ratio for batched cases;
``ObliqueShockIC`` is the double-Mach-reflection-style state split by a line at 60 degrees;
``DoubleMachReflectionIC`` is the reference's own double-Mach-reflection example.
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


class DoubleMachReflectionIC(InitialCondition):
    """The reference's double-Mach-reflection example (examples/dmr/initial_condition.py:32-61):
    primitive (8, 8.25, 0, 116.5) where x <= 0.95, (1.4, 0, 0, 1) elsewhere."""

    def apply_to_block(self, block):
        f = block.config.fluid
        left = _cons(f, 8.0, 8.25, 0.0, 116.5)
        right = _cons(f, 1.4, 0.0, 0.0, 1.0)
        block.state.data = np.where(block.mesh.x <= 0.95, left.data, right.data)
        block.state.make_non_dimensional()


class ImplosionIC(InitialCondition):
    """The reference's implosion example (examples/implosion/initial_condition.py): still air at the
    low state (rho 1.1742, p 101100) where x <= 5 and y <= 5, the high state (rho 4.6968, p 404400)
    elsewhere."""

    def apply_to_block(self, block):
        f = block.config.fluid
        high = _cons(f, 4.6968, 0.0, 0.0, 404400.0)
        low = _cons(f, 1.1742, 0.0, 0.0, 101100.0)
        block.state.data = np.where(np.logical_and(block.mesh.x <= 5, block.mesh.y <= 5), low.data, high.data)
        block.state.make_non_dimensional()


class ShockBoxIC(InitialCondition):
    """The reference's shock-box example (examples/shockbox/initial_condition.py): the low state in
    the quadrants x <= 5, y <= 5 and x > 5, y > 5, the high state in the other two."""

    def apply_to_block(self, block):
        f = block.config.fluid
        high = _cons(f, 4.6968, 0.0, 0.0, 404400.0)
        low = _cons(f, 1.1742, 0.0, 0.0, 101100.0)
        x, y = block.mesh.x, block.mesh.y
        low_quadrants = np.logical_or(np.logical_and(x <= 5, y <= 5), np.logical_and(x > 5, y > 5))
        block.state.data = np.where(low_quadrants, low.data, high.data)
        block.state.make_non_dimensional()


class SmoothExplosionIC(InitialCondition):
    """The explosion of ``ExplosionInitialCondition`` with differentiable data: the square's edges
    are tanh ramps of half-width ``width``; density and pressure carry a smooth background variation
    of relative amplitude ``amplitude`` and the air a smooth background velocity of ``amplitude``
    times 343 m/s (phases ``phases``), so that no two neighbouring cells hold exactly equal values
    in any variable.  The piecewise-constant original is a point where the discrete scheme is not
    differentiable (every cell of its flat regions, and the zero momentum everywhere, sits on the
    limiter's d = 0 branch or on a max/min tie), so the discrete adjoint is evaluated at this
    smoothed state."""

    def __init__(self, width=0.1, amplitude=0.01, phases=(0.3, 1.1, 2.0, 2.9, 0.7, 1.9)):
        self.width = float(width)
        self.amplitude = float(amplitude)
        phases = tuple(float(p) for p in phases)
        self.phases = phases + (0.7, 1.9)[len(phases) - 4:] if len(phases) < 6 else phases

    def _box(self, x, y):
        w = self.width
        sx = 0.25 * (1.0 + np.tanh((x - 3.0) / w)) * (1.0 + np.tanh((7.0 - x) / w))
        sy = 0.25 * (1.0 + np.tanh((y - 3.0) / w)) * (1.0 + np.tanh((7.0 - y) / w))
        return sx * sy

    def primitive(self, x, y):
        """(rho, u, v, p) in dimensional units at points x, y."""
        b = self._box(x, y)
        a, (p1, p2, p3, p4, p5, p6) = self.amplitude, self.phases
        rho = (1.1742 + (4.6968 - 1.1742) * b) * (1.0 + a * np.sin(0.7 * x + p1) * np.cos(0.9 * y + p2))
        p = (101100.0 + (404400.0 - 101100.0) * b) * (1.0 + a * np.cos(0.8 * x + p3) * np.sin(0.6 * y + p4))
        u = a * 343.0 * np.sin(0.5 * x + p5) * np.cos(0.4 * y + p6)
        v = a * 343.0 * np.cos(0.45 * x + p6) * np.sin(0.55 * y + p5)
        return rho, u, v, p

    def apply_to_block(self, block):
        x, y = block.mesh.x[..., 0], block.mesh.y[..., 0]
        rho, u, v, p = self.primitive(x, y)
        W = np.stack([rho, u, v, p], axis=-1)
        block.state.data = PrimitiveState(fluid=block.config.fluid, array=W).to_type(ConservativeState).data
        block.state.make_non_dimensional()


class DensityBumpIC(InitialCondition):
    """``base`` with the density multiplied by 1 + amplitude*exp(-|x - c|^2 / (2 sigma^2))
    (velocity and pressure unchanged)."""

    def __init__(self, base, amplitude, x0, y0, sigma):
        self.base, self.amplitude, self.center, self.sigma = base, float(amplitude), (float(x0), float(y0)), float(sigma)

    def primitive(self, x, y):
        rho, u, v, p = self.base.primitive(x, y)
        r2 = (x - self.center[0]) ** 2 + (y - self.center[1]) ** 2
        return rho * (1.0 + self.amplitude * np.exp(-r2 / (2.0 * self.sigma ** 2))), u, v, p

    def apply_to_block(self, block):
        x, y = block.mesh.x[..., 0], block.mesh.y[..., 0]
        W = np.stack(self.primitive(x, y), axis=-1)
        block.state.data = PrimitiveState(fluid=block.config.fluid, array=W).to_type(ConservativeState).data
        block.state.make_non_dimensional()
