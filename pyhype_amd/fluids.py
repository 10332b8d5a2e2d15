"""Fluids: the thermodynamic constants read by the hot path.

Mirrors ``Fluid``/``Air`` of the reference (fluids/base.py:27-64, fluids/air.py:24-28): a far-field
(a, rho) pair used for non-dimensionalisation and gamma-derived constants evaluated the same way.
"""
from __future__ import annotations

from collections import namedtuple

FarField = namedtuple("far_field", ["a", "rho"])


class Fluid:
    gas_constant = 8314.0
    _molecular_mass = float("-inf")

    def __init__(self, a_inf: float = 1.0, rho_inf: float = 1.0):
        self._far_field = FarField(a_inf, rho_inf)
        self._R = self.gas_constant / self.molecular_mass

    @property
    def far_field(self):
        return self._far_field

    @property
    def R(self):
        return self._R

    @property
    def molecular_mass(self):
        return self._molecular_mass

    def gamma(self, *args, temperature: float = None, **kwargs) -> float:
        raise NotImplementedError

    def g_over_gm1(self, *args, temperature: float = None, **kwargs) -> float:
        g = self.gamma(*args, temperature, **kwargs)
        return g / (g - 1.0)

    def one_over_gm1(self, *args, temperature: float = None, **kwargs) -> float:
        g = self.gamma(*args, temperature, **kwargs)
        return 1.0 / (g - 1.0)

    def gm1(self) -> float:
        """gamma - 1 as the conservative-state formulas evaluate it (states/conservative.py:110)."""
        return self.gamma() - 1


class Air(Fluid):
    _molecular_mass = 28.97

    def gamma(self, *args, temperature: float = None, **kwargs) -> float:
        return 1.4
