"""Explicit Runge-Kutta tables (Butcher lower-triangular form).

Same names and coefficients as ``ExplicitRungeKutta`` and ``TimeIntegratorFactory``
(time_marching/explicit_runge_kutta.py:91-274, time_marching/factory.py:29-62).  Coefficients are
evaluated with the same Python expressions so they are the same doubles.  The stage kernel applies
a table in accumulator form; zero entries are skipped exactly as ``integrate`` skips them.
``"SSPRK2"`` is ``Generic2`` with alpha = 1, which the reference reaches only by setting
``config.alpha`` (a name outside its ``SolverConfig`` slots).
"""
from __future__ import annotations

from typing import List


def _generic2(a):
    return [[a], [1 - 1 / (2 * a), 1 / (2 * a)]]


def _generic3(a):
    if a in (0, 2 / 3, 1):
        raise ValueError("Value of alpha parameter is not allowd.")
    k = (1 - a) / a / (3 * a - 2)
    return [[a], [1 + k, -k], [0.5 - 1 / (6 * a), 1 / (6 * a * (1 - a)), (2 - 3 * a) / (6 * (1 - a))]]


TABLES = {
    "ExplicitEuler1": lambda cfg: [[1]],
    "RK2": lambda cfg: [[0.5], [0, 1]],
    "Generic2": lambda cfg: _generic2(_alpha(cfg)),
    "SSPRK2": lambda cfg: _generic2(1),
    "Ralston2": lambda cfg: [[2 / 3], [1 / 4, 3 / 4]],
    "Generic3": lambda cfg: _generic3(_alpha(cfg)),
    "RK3": lambda cfg: [[0.5], [-1, 2], [1 / 6, 2 / 3, 1 / 6]],
    "RK3SSP": lambda cfg: [[1], [1 / 4, 1 / 4], [1 / 6, 1 / 6, 2 / 3]],
    "Ralston3": lambda cfg: [[1 / 2], [0, 3 / 4], [2 / 9, 1 / 3, 4 / 9]],
    "RK4": lambda cfg: [[0.5], [0, 0.5], [0, 0, 1], [1 / 6, 1 / 3, 1 / 3, 1 / 6]],
    "Ralston4": lambda cfg: [
        [0.4],
        [0.29697761, 0.15875964],
        [0.21810040, -3.05096516, 3.83286476],
        [0.17476028, -0.55148066, 1.20553560, 0.17118478],
    ],
    "DormandPrince5": lambda cfg: [
        [1 / 5],
        [3 / 40, 9 / 40],
        [44 / 45, -56 / 15, 32 / 9],
        [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
        [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
        [35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84],
    ],
}


def _alpha(cfg):
    a = getattr(cfg, "alpha", None) if cfg is not None else None
    if a is None:
        raise ValueError("time integrator requires config.alpha")
    return a


def butcher_table(name, config=None) -> List[List[float]]:
    """TimeIntegratorFactory.create equivalent; a list-of-rows table is passed through."""
    if isinstance(name, (list, tuple)):
        return [list(r) for r in name]
    if name not in TABLES:
        raise ValueError(f"Time marching scheme {name} is not available.")
    return TABLES[name](config)


def stages(table) -> int:
    return len(table)
