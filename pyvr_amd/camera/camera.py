"""``Camera``: spherical camera parameters, validation and GL-style matrices.

Same fields, defaults, presets, serialisation and errors as the reference's
``pyvr/camera/camera.py:16-433``.  The two matrices are built the way the
reference builds them, because the ray marcher consumes them exactly as
uploaded:

* ``get_view_matrix`` (reference camera.py:292-334) is a look-at matrix
  written *pre-transposed*: rows 0-2 hold (right, up, -forward) as columns and
  row 3 holds the translation.  Its row-major bytes read as a column-major
  GLSL ``mat4`` give the conventional view matrix.
* ``get_projection_matrix`` (camera.py:336-369) is the conventional
  perspective matrix in row-major order (so the shader sees its transpose;
  only its block-diagonal x/y scale factors matter for ray generation).

Both are float32, computed from float64 camera vectors.
"""

import json
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Dict, Optional, Tuple, Union

import numpy as np


def _cross3(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """``np.cross`` of two 3-vectors, with its arithmetic: both promoted to
    the common dtype, then (a1*b2 - a2*b1, a2*b0 - a0*b2, a0*b1 - a1*b0),
    each product and difference rounded in that dtype.  Without np.cross's
    general broadcasting it is ~50x faster, which matters on the per-frame
    set_camera path."""
    dt = np.promote_types(a.dtype, b.dtype)
    a0, a1, a2 = a.astype(dt, copy=False)
    b0, b1, b2 = b.astype(dt, copy=False)
    return np.array([a1 * b2 - a2 * b1, a2 * b0 - a0 * b2, a0 * b1 - a1 * b0], dtype=dt)


def _norm(x: np.ndarray):
    """``np.linalg.norm`` of a float vector, computed as it does:
    sqrt(x.dot(x)) in x's dtype."""
    return np.sqrt(x.dot(x))


@dataclass
class Camera:
    target: np.ndarray = field(
        default_factory=lambda: np.array([0.0, 0.0, 0.0], dtype=np.float32)
    )
    azimuth: float = 0.0
    elevation: float = 0.0
    roll: float = 0.0
    distance: float = 3.0
    init_pos: np.ndarray = field(
        default_factory=lambda: np.array([1.0, 0.0, 0.0], dtype=np.float32)
    )
    init_up: np.ndarray = field(
        default_factory=lambda: np.array([0.0, 0.0, 1.0], dtype=np.float32)
    )
    fov: float = np.pi / 4
    near_plane: float = 0.1
    far_plane: float = 100.0

    def __post_init__(self):
        self.validate()

    def validate(self) -> None:
        """Raise ``ValueError`` on invalid parameters (reference camera.py:63-113)."""
        if not isinstance(self.target, np.ndarray) or self.target.shape != (3,):
            raise ValueError("target must be a 3D numpy array")
        for name in ("azimuth", "elevation", "roll"):
            value = getattr(self, name)
            if not isinstance(value, (int, float)):
                raise ValueError(f"{name} must be numeric")
            if abs(value) > 4 * np.pi:
                print(
                    f"Warning: {name} = {value:.3f} rad ({np.degrees(value):.1f}°) is unusually large"
                )
        if not isinstance(self.distance, (int, float)) or self.distance <= 0:
            raise ValueError("distance must be positive")
        if not isinstance(self.init_pos, np.ndarray) or self.init_pos.shape != (3,):
            raise ValueError("init_pos must be a 3D numpy array")
        if not isinstance(self.init_up, np.ndarray) or self.init_up.shape != (3,):
            raise ValueError("init_up must be a 3D numpy array")
        if np.linalg.norm(self.init_pos - self.target) < 1e-9:
            raise ValueError("init_pos must not be at the same location as target")
        if np.linalg.norm(self.init_up) < 1e-9:
            raise ValueError("init_up must not be the zero vector")
        if self.fov <= 0 or self.fov >= np.pi:
            raise ValueError("fov must be between 0 and π radians")
        if self.near_plane <= 0:
            raise ValueError("near_plane must be positive")
        if self.far_plane <= self.near_plane:
            raise ValueError("far_plane must be greater than near_plane")

    @classmethod
    def from_spherical(cls, target: np.ndarray, azimuth: float, elevation: float,
                       roll: float, distance: float, **kwargs) -> "Camera":
        return cls(target=target, azimuth=azimuth, elevation=elevation, roll=roll,
                   distance=distance, **kwargs)

    @staticmethod
    def _origin() -> np.ndarray:
        return np.array([0.0, 0.0, 0.0], dtype=np.float32)

    @classmethod
    def front_view(cls, target: Optional[np.ndarray] = None, distance: float = 3.0) -> "Camera":
        return cls(target=cls._origin() if target is None else target,
                   azimuth=0.0, elevation=0.0, roll=0.0, distance=distance)

    @classmethod
    def side_view(cls, target: Optional[np.ndarray] = None, distance: float = 3.0) -> "Camera":
        return cls(target=cls._origin() if target is None else target,
                   azimuth=np.pi / 2, elevation=0.0, roll=0.0, distance=distance)

    @classmethod
    def top_view(cls, target: Optional[np.ndarray] = None, distance: float = 3.0) -> "Camera":
        return cls(target=cls._origin() if target is None else target,
                   azimuth=0.0, elevation=np.pi / 2, roll=0.0, distance=distance)

    @classmethod
    def isometric_view(cls, target: Optional[np.ndarray] = None, distance: float = 3.0) -> "Camera":
        return cls(target=cls._origin() if target is None else target,
                   azimuth=np.pi / 4, elevation=np.pi / 6, roll=0.0, distance=distance)

    def to_dict(self) -> Dict[str, Any]:
        return {
            "target": self.target.tolist(),
            "azimuth": float(self.azimuth),
            "elevation": float(self.elevation),
            "roll": float(self.roll),
            "distance": float(self.distance),
            "init_pos": self.init_pos.tolist(),
            "init_up": self.init_up.tolist(),
            "fov": float(self.fov),
            "near_plane": float(self.near_plane),
            "far_plane": float(self.far_plane),
        }

    @classmethod
    def from_dict(cls, data: Dict[str, Any]) -> "Camera":
        return cls(
            target=np.array(data["target"], dtype=np.float32),
            azimuth=data["azimuth"],
            elevation=data["elevation"],
            roll=data["roll"],
            distance=data["distance"],
            init_pos=np.array(data["init_pos"], dtype=np.float32),
            init_up=np.array(data["init_up"], dtype=np.float32),
            fov=data.get("fov", np.pi / 4),
            near_plane=data.get("near_plane", 0.1),
            far_plane=data.get("far_plane", 100.0),
        )

    def save_to_file(self, filepath: Union[str, Path]) -> None:
        with open(filepath, "w") as f:
            json.dump(self.to_dict(), f, indent=2)

    @classmethod
    def load_from_file(cls, filepath: Union[str, Path]) -> "Camera":
        with open(filepath, "r") as f:
            return cls.from_dict(json.load(f))

    def copy(self) -> "Camera":
        return Camera.from_dict(self.to_dict())

    def get_camera_vectors(self) -> Tuple[np.ndarray, np.ndarray]:
        """(position, up) as float64 arrays."""
        from .control import get_camera_pos_from_params

        return get_camera_pos_from_params(self)

    def get_view_matrix(self) -> np.ndarray:
        """Pre-transposed look-at matrix, float32 (layout in the module doc)."""
        return self._view_matrix(*self.get_camera_vectors())

    def _view_matrix(self, position: np.ndarray, up: np.ndarray) -> np.ndarray:
        """``get_view_matrix`` from already computed camera vectors."""
        fwd = self.target - position
        fwd = fwd / _norm(fwd)
        right = _cross3(fwd, up)
        right = right / _norm(right)
        true_up = _cross3(right, fwd)
        m = np.zeros((4, 4), dtype=np.float64)
        m[:3, 0] = right
        m[:3, 1] = true_up
        m[:3, 2] = -fwd
        m[3, 0] = -np.dot(right, position)
        m[3, 1] = -np.dot(true_up, position)
        m[3, 2] = np.dot(fwd, position)
        m[3, 3] = 1.0
        return m.astype(np.float32)

    def get_projection_matrix(self, aspect_ratio: float) -> np.ndarray:
        """Perspective matrix (row-major), float32."""
        near, far = self.near_plane, self.far_plane
        f = 1.0 / np.tan(self.fov / 2.0)
        m = np.zeros((4, 4), dtype=np.float64)
        m[0, 0] = f / aspect_ratio
        m[1, 1] = f
        m[2, 2] = (far + near) / (near - far)
        m[2, 3] = (2 * far * near) / (near - far)
        m[3, 2] = -1.0
        return m.astype(np.float32)

    def __repr__(self) -> str:
        return (
            f"Camera(target={self.target}, "
            f"azimuth={self.azimuth:.3f} rad ({np.degrees(self.azimuth):.1f}°), "
            f"elevation={self.elevation:.3f} rad ({np.degrees(self.elevation):.1f}°), "
            f"roll={self.roll:.3f} rad ({np.degrees(self.roll):.1f}°), "
            f"distance={self.distance:.2f})"
        )


class CameraError(Exception):
    """Raised for invalid camera parameters."""


def validate_camera_angles(azimuth: float, elevation: float, roll: float) -> None:
    """``CameraError`` unless every angle is a finite int/float."""
    for name, value in (("azimuth", azimuth), ("elevation", elevation), ("roll", roll)):
        if not isinstance(value, (int, float)):
            raise CameraError(f"{name} must be numeric, got {type(value)}")
        if not np.isfinite(value):
            raise CameraError(f"{name} must be finite, got {value}")


def degrees_to_radians(**kwargs) -> Dict[str, float]:
    return {k: np.radians(v) for k, v in kwargs.items()}


def radians_to_degrees(**kwargs) -> Dict[str, float]:
    return {k: np.degrees(v) for k, v in kwargs.items()}
