"""Colour transfer function: scalar -> RGB, piecewise linear.

API of the reference's ``pyvr/transferfunctions/color.py:20-250``.  The LUT
row j holds the colour at ``x_j = j / (size - 1)``, interpolated with
``np.interp`` in float64 and stored as float32 (reference color.py:173-197).
"""

import itertools
from typing import List, Optional, Tuple

import numpy as np

from .base import BaseTransferFunction, InvalidControlPointError, validate_control_points_format


class ColorTransferFunction(BaseTransferFunction):
    def __init__(self, control_points: Optional[List[Tuple[float, Tuple[float, float, float]]]] = None,
                 lut_size: int = 256):
        super().__init__(control_points, lut_size)

    def _get_default_control_points(self):
        return [(0.0, (0.0, 0.0, 0.0)), (1.0, (1.0, 1.0, 1.0))]

    def _validate_control_points(self) -> None:
        validate_control_points_format(self.control_points, expected_value_length=3)
        self._validate_scalar_range(0.0, 1.0)
        for i, (_, rgb) in enumerate(self.control_points):
            for j, component in enumerate(rgb):
                if not (0.0 <= component <= 1.0):
                    raise InvalidControlPointError(
                        f"Control point {i} RGB component {j} value {component} outside valid range [0.0, 1.0]"
                    )

    @classmethod
    def grayscale(cls, lut_size: int = 256) -> "ColorTransferFunction":
        return cls([(0.0, (0.0, 0.0, 0.0)), (1.0, (1.0, 1.0, 1.0))], lut_size=lut_size)

    @classmethod
    def single_color(cls, color: Tuple[float, float, float], lut_size: int = 256) -> "ColorTransferFunction":
        return cls([(0.0, color), (1.0, color)], lut_size=lut_size)

    @classmethod
    def from_colormap(cls, colormap_name: str, value_range: Tuple[float, float] = (0.0, 1.0),
                      lut_size: int = 256) -> "ColorTransferFunction":
        """One control point per colour of a matplotlib colormap sampled at
        ``linspace(0, 1, lut_size)`` and placed at ``value_range`` (reference
        color.py:105-151).  With ``lut_size`` equal to a listed colormap's
        length this reproduces the listed colours exactly."""
        try:
            import matplotlib

            cmap = matplotlib.colormaps.get_cmap(colormap_name)
        except ImportError:
            raise ImportError("matplotlib is required for from_colormap method")
        except Exception as e:
            raise ValueError(f"Unknown colormap name '{colormap_name}': {e}")
        x = np.linspace(0, 1, lut_size)
        colors = cmap(x)
        if colors.shape[1] == 4:
            colors = colors[:, :3]
        lo, hi = value_range
        xs = lo + x * (hi - lo)
        points = [(float(xi), tuple(map(float, rgb))) for xi, rgb in zip(xs, colors)]
        return cls(points, lut_size=lut_size)

    @classmethod
    def two_color_ramp(cls, color1, color2, lut_size: int = 256) -> "ColorTransferFunction":
        return cls([(0.0, color1), (1.0, color2)], lut_size=lut_size)

    def _interp(self, x) -> np.ndarray:
        scalars, colors = zip(*self.control_points)
        # the colours as a float64 (n, 3) array (np.interp works in float64
        # whatever the input type); fromiter is twice as fast as np.array on
        # the nested tuples, which matters when every request has its own LUT
        colors = np.fromiter(itertools.chain.from_iterable(colors), dtype=np.float64,
                             count=3 * len(colors)).reshape(-1, 3)
        x = np.asarray(x)
        out = np.empty(x.shape + (3,), dtype=np.float32)
        for c in range(3):
            out[..., c] = np.interp(x, scalars, colors[:, c])
        return out

    def to_lut(self, size: Optional[int] = None) -> np.ndarray:
        """(size, 3) float32 LUT, rows at ``linspace(0, 1, size)``."""
        n = size or self.lut_size
        return self._interp(np.linspace(0, 1, n))

    def apply_to_array(self, scalar_array: np.ndarray) -> np.ndarray:
        return self._interp(scalar_array.flatten()).reshape(scalar_array.shape + (3,))

    def get_color_at(self, scalar: float) -> Tuple[float, float, float]:
        return tuple(self._interp(np.array([scalar]))[0].tolist())

    def __repr__(self) -> str:
        return (f"ColorTransferFunction({len(self.control_points)} control points, "
                f"lut_size={self.lut_size})")
