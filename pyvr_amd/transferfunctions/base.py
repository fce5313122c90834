"""Control-point transfer functions: shared base class and validation.

API of the reference's ``pyvr/transferfunctions/base.py:15-162``: control
points are ``(scalar, value)`` pairs kept sorted by scalar; ``to_lut(size)``
samples the piecewise-linear function at ``linspace(0, 1, size)``.
"""

from abc import ABC, abstractmethod
from typing import List, Optional, Tuple

import numpy as np


class BaseTransferFunction(ABC):
    def __init__(self, control_points: Optional[List[Tuple]] = None, lut_size: int = 256):
        if control_points is None:
            control_points = self._get_default_control_points()
        self.control_points = sorted(control_points, key=lambda p: p[0])
        self.lut_size = lut_size
        self._validate_control_points()

    @abstractmethod
    def _get_default_control_points(self) -> List[Tuple]:
        """Control points used when none are given."""

    @abstractmethod
    def _validate_control_points(self) -> None:
        """Raise on control points that do not fit this function type."""

    @abstractmethod
    def to_lut(self, size: Optional[int] = None) -> np.ndarray:
        """Lookup table sampled at ``linspace(0, 1, size)``."""

    def __call__(self, size: Optional[int] = None) -> np.ndarray:
        return self.to_lut(size)

    def _validate_scalar_range(self, min_val: float = 0.0, max_val: float = 1.0) -> None:
        for scalar, _ in self.control_points:
            if not (min_val <= scalar <= max_val):
                raise ValueError(
                    f"Control point scalar {scalar} outside valid range [{min_val}, {max_val}]"
                )

    def _get_scalar_values(self) -> np.ndarray:
        return np.array([p[0] for p in self.control_points])

    def _get_mapped_values(self) -> np.ndarray:
        return np.array([p[1] for p in self.control_points])


class TransferFunctionError(Exception):
    """Base class of transfer-function errors."""


class InvalidControlPointError(TransferFunctionError):
    """Raised for malformed or out-of-range control points."""


def validate_control_points_format(control_points: List[Tuple], expected_value_length: int) -> None:
    """Check ``(scalar, value)`` structure; value is a number (length 1) or a
    sequence of ``expected_value_length`` numbers."""
    if not control_points:
        raise InvalidControlPointError("Control points cannot be empty")
    for i, point in enumerate(control_points):
        if not isinstance(point, (tuple, list)) or len(point) != 2:
            raise InvalidControlPointError(
                f"Control point {i} must be a (scalar, value) tuple, got {point}"
            )
        scalar, value = point
        if not isinstance(scalar, (int, float)):
            raise InvalidControlPointError(
                f"Control point {i} scalar must be numeric, got {type(scalar)}"
            )
        if expected_value_length == 1:
            if not isinstance(value, (int, float)):
                raise InvalidControlPointError(
                    f"Control point {i} value must be numeric, got {type(value)}"
                )
            continue
        if not isinstance(value, (tuple, list, np.ndarray)) or len(value) != expected_value_length:
            raise InvalidControlPointError(
                f"Control point {i} value must be a {expected_value_length}-element sequence, got {value}"
            )
        for j, component in enumerate(value):
            if not isinstance(component, (int, float)):
                raise InvalidControlPointError(
                    f"Control point {i} value component {j} must be numeric, got {type(component)}"
                )
