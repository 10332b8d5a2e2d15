"""Host-side state containers used by initial conditions.

A small mirror of the reference's ``State`` classes (states/base.py:37-313,
states/conservative.py:32-165, states/primitive.py:34-243) sufficient for the initial-condition
protocol: ``block.state.data`` assignment, ``make_non_dimensional`` and primitive -> conservative
conversion.  The hot path never touches these objects; it runs on device buffers.
"""
from __future__ import annotations

import numpy as np

from .fluids import Fluid


class RealizabilityError(FloatingPointError):
    """Raised when a state has rho <= 0 or E <= 0 (reference: RealizabilityException,
    states/base.py:33-34, 270-287)."""


class State:
    def __init__(self, fluid: Fluid, state: "State" = None, array: np.ndarray = None, shape=None, fill=None):
        if not isinstance(fluid, Fluid):
            raise TypeError("fluid must be of type Fluid.")
        self.fluid = fluid
        self._data = None
        if state is not None:
            self._data = np.zeros(state.shape)
            self.from_state(state)
        elif array is not None:
            self.from_array(array)
        elif shape is not None:
            self.data = np.full(shape=shape, fill_value=fill if fill is not None else 0.0)
        else:
            raise ValueError("State constructor must recieve either a state, array or a shape.")

    @property
    def data(self) -> np.ndarray:
        return self._data

    @data.setter
    def data(self, data: np.ndarray):
        self.from_array(data)

    @property
    def shape(self):
        return self._data.shape

    def from_array(self, array: np.ndarray):
        if not isinstance(array, np.ndarray):
            raise TypeError(f"Input array must be a Numpy array, but it is a {type(array)}.")
        if array.ndim != 3 or array.shape[-1] != 4:
            raise ValueError("Array must have 3 dims and a depth of 4.")
        if self._data is None or self._data.shape == array.shape:
            self._data = array
        else:
            self._data[:, :, :] = array

    def make_non_dimensional(self):
        """State.make_non_dimensional (states/base.py:93-97)."""
        ff = self.fluid.far_field
        self.data[:, :, 0] /= ff.rho
        self.data[:, :, 1] /= ff.rho * ff.a
        self.data[:, :, 2] /= ff.rho * ff.a
        self.data[:, :, 3] /= ff.rho * ff.a ** 2

    def to_type(self, to_type):
        return to_type(fluid=self.fluid, state=self)

    def from_state(self, state: "State"):
        if type(state) is type(self):
            self.data = state.data.copy()
        elif isinstance(self, ConservativeState) and isinstance(state, PrimitiveState):
            self.data = state.to_conservative_array()
        elif isinstance(self, PrimitiveState) and isinstance(state, ConservativeState):
            self.data = state.to_primitive_array()
        else:
            raise TypeError(f"cannot convert {type(state).__name__} to {type(self).__name__}")


class ConservativeState(State):
    """(rho, rho*u, rho*v, E) (states/conservative.py:32-40)."""

    def to_primitive_array(self) -> np.ndarray:
        """ConservativeConverter.to_primitive (states/converter/concrete_defs.py:85-100)."""
        d = self.data
        rho = d[:, :, 0]
        u = d[:, :, 1] / rho
        v = d[:, :, 2] / rho
        ek = rho * (0.5 * (u * u + v * v))
        return np.dstack((rho.copy(), u, v, self.fluid.gm1() * (d[:, :, 3] - ek)))


class PrimitiveState(State):
    """(rho, u, v, p) (states/primitive.py:34-42)."""

    def ek(self) -> np.ndarray:
        """PrimitiveState.ek_JIT (states/primitive.py:93-104)."""
        d = self.data
        return 0.5 * d[:, :, 0] * (d[:, :, 1] * d[:, :, 1] + d[:, :, 2] * d[:, :, 2])

    def to_conservative_array(self) -> np.ndarray:
        """PrimitiveConverter.to_conservative (states/converter/concrete_defs.py:126-141)."""
        d = self.data
        rho, u, v, p = d[:, :, 0], d[:, :, 1], d[:, :, 2], d[:, :, 3]
        return np.dstack((rho.copy(), rho * u, rho * v, p / self.fluid.gm1() + self.ek()))
