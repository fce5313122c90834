"""``Volume``: scalar field, optional normals and world-space bounds.

Same fields, defaults (bounds +-0.5), validation and helpers as the
reference's ``pyvr/volume/data.py:15-101``.
"""

from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np


@dataclass
class Volume:
    data: np.ndarray
    normals: Optional[np.ndarray] = None
    min_bounds: np.ndarray = field(
        default_factory=lambda: np.array([-0.5, -0.5, -0.5], dtype=np.float32)
    )
    max_bounds: np.ndarray = field(
        default_factory=lambda: np.array([0.5, 0.5, 0.5], dtype=np.float32)
    )
    name: Optional[str] = None

    def __post_init__(self):
        self.validate()

    def validate(self) -> None:
        if not isinstance(self.data, np.ndarray):
            raise ValueError("Volume data must be a numpy array")
        if len(self.data.shape) != 3:
            raise ValueError(f"Volume data must be 3D, got shape {self.data.shape}")
        if self.normals is not None:
            if not isinstance(self.normals, np.ndarray):
                raise ValueError("Normal volume must be a numpy array")
            expected = self.data.shape + (3,)
            if self.normals.shape != expected:
                raise ValueError(
                    f"Normal volume must have shape {expected}, got {self.normals.shape}"
                )
        if not isinstance(self.min_bounds, np.ndarray) or self.min_bounds.shape != (3,):
            raise ValueError("min_bounds must be a 3D numpy array")
        if not isinstance(self.max_bounds, np.ndarray) or self.max_bounds.shape != (3,):
            raise ValueError("max_bounds must be a 3D numpy array")
        if np.any(self.max_bounds <= self.min_bounds):
            raise ValueError("max_bounds must be greater than min_bounds")

    @property
    def shape(self) -> Tuple[int, int, int]:
        return self.data.shape

    @property
    def dimensions(self) -> np.ndarray:
        return self.max_bounds - self.min_bounds

    @property
    def center(self) -> np.ndarray:
        return (self.min_bounds + self.max_bounds) / 2.0

    @property
    def has_normals(self) -> bool:
        return self.normals is not None

    @property
    def voxel_spacing(self) -> np.ndarray:
        return self.dimensions / np.array(self.shape, dtype=np.float32)

    def compute_normals(self, method: str = "gradient") -> None:
        if method != "gradient":
            raise ValueError(f"Unsupported method: {method}")
        from ..datasets.synthetic import compute_normal_volume

        self.normals = compute_normal_volume(self.data)

    def normalize(self, method: str = "minmax") -> "Volume":
        if method == "minmax":
            lo, hi = self.data.min(), self.data.max()
            out = np.zeros_like(self.data) if hi - lo < 1e-9 else (self.data - lo) / (hi - lo)
        elif method == "zscore":
            mean, std = self.data.mean(), self.data.std()
            out = np.zeros_like(self.data) if std < 1e-9 else (self.data - mean) / std
        else:
            raise ValueError(f"Unsupported method: {method}")
        return Volume(
            data=out.astype(np.float32),
            normals=self.normals.copy() if self.normals is not None else None,
            min_bounds=self.min_bounds.copy(),
            max_bounds=self.max_bounds.copy(),
            name=f"{self.name}_normalized" if self.name else None,
        )

    def copy(self) -> "Volume":
        return Volume(
            data=self.data.copy(),
            normals=self.normals.copy() if self.normals is not None else None,
            min_bounds=self.min_bounds.copy(),
            max_bounds=self.max_bounds.copy(),
            name=self.name,
        )

    def __repr__(self) -> str:
        name = f"'{self.name}'" if self.name else "unnamed"
        normals = "with normals" if self.has_normals else "no normals"
        return (f"Volume({name}, shape={self.shape}, "
                f"bounds=[{self.min_bounds}, {self.max_bounds}], {normals})")


class VolumeError(Exception):
    """Raised for volume-data errors."""
