"""Opacity transfer function: scalar -> alpha, piecewise linear.

API of the reference's ``pyvr/transferfunctions/opacity.py:20-215``; LUT rows
at ``linspace(0, 1, size)`` via ``np.interp`` (float64), stored float32.
"""

from typing import List, Optional, Tuple

import numpy as np

from .base import BaseTransferFunction, InvalidControlPointError, validate_control_points_format


class OpacityTransferFunction(BaseTransferFunction):
    def __init__(self, control_points: Optional[List[Tuple[float, float]]] = None,
                 lut_size: int = 256):
        super().__init__(control_points, lut_size)

    def _get_default_control_points(self):
        return [(0.0, 0.0), (1.0, 1.0)]

    def _validate_control_points(self) -> None:
        validate_control_points_format(self.control_points, expected_value_length=1)
        self._validate_scalar_range(0.0, 1.0)
        for i, (_, opacity) in enumerate(self.control_points):
            if not (0.0 <= opacity <= 1.0):
                raise InvalidControlPointError(
                    f"Control point {i} opacity {opacity} outside valid range [0.0, 1.0]"
                )

    @classmethod
    def linear(cls, low: float = 0.0, high: float = 1.0, lut_size: int = 256) -> "OpacityTransferFunction":
        return cls([(0.0, low), (1.0, high)], lut_size=lut_size)

    @classmethod
    def one_step(cls, step: float = 0.5, low: float = 0.0, high: float = 1.0,
                 lut_size: int = 256) -> "OpacityTransferFunction":
        """Step from ``low`` to ``high`` just after ``step`` (1e-12 wide ramp)."""
        eps = 1e-12
        return cls([(0.0, low), (step, low), (step + eps, high), (1.0, high)], lut_size=lut_size)

    @classmethod
    def peaks(cls, peaks: List[float], opacity: float = 1.0, eps: float = 0.02,
              lut_size: int = 256, base: float = 0.0) -> "OpacityTransferFunction":
        """Triangular peaks of half-width ``eps`` on a ``base`` level
        (reference opacity.py:113-171)."""
        if not peaks:
            raise ValueError("At least one peak position must be specified")
        for p in peaks:
            if not (0.0 <= p <= 1.0):
                raise ValueError(f"Peak position {p} must be between 0 and 1")
        pts = [(0.0, base)]
        for p in sorted(peaks):
            left = max(0.0, p - eps)
            right = min(1.0, p + eps)
            if left > pts[-1][0]:
                pts.append((left, base))
            pts.append((p, opacity))
            if right > p:
                pts.append((right, base))
        if pts[-1][0] < 1.0:
            pts.append((1.0, base))
        return cls(pts, lut_size=lut_size)

    def to_lut(self, size: Optional[int] = None) -> np.ndarray:
        """(size,) float32 LUT, rows at ``linspace(0, 1, size)``."""
        n = size or self.lut_size
        scalars, opacities = zip(*self.control_points)
        return np.interp(np.linspace(0, 1, n), scalars, opacities).astype(np.float32)

    def apply_to_array(self, scalar_array: np.ndarray) -> np.ndarray:
        scalars, opacities = zip(*self.control_points)
        return (np.interp(scalar_array.flatten(), scalars, opacities)
                .reshape(scalar_array.shape).astype(np.float32))

    def __repr__(self) -> str:
        return (f"OpacityTransferFunction(control_points={self.control_points}, "
                f"lut_size={self.lut_size})")
