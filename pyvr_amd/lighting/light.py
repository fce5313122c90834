"""Light description consumed by the ray marcher.

Same dataclass fields, factories and validation as the reference's
``pyvr/lighting/light.py:14-431``.  The ray marcher reads ``position``,
``target``, ``ambient_intensity`` and ``diffuse_intensity``
(``pyvr/moderngl_renderer/renderer.py:311-316``) and shades each sample with
``ambient + diffuse * max(dot(n, normalize(target - position)), 0)``
(``pyvr/shaders/volume.frag.glsl:106-110``).
"""

from dataclasses import dataclass, field
from typing import Optional

import numpy as np


@dataclass
class Light:
    position: np.ndarray = field(
        default_factory=lambda: np.array([1.0, 1.0, 1.0], dtype=np.float32)
    )
    target: np.ndarray = field(
        default_factory=lambda: np.array([0.0, 0.0, 0.0], dtype=np.float32)
    )
    ambient_intensity: float = 0.2
    diffuse_intensity: float = 0.8

    _is_linked: bool = field(default=False, init=False, repr=False)
    _camera_offsets: Optional[dict] = field(default=None, init=False, repr=False)

    def __post_init__(self):
        self.validate()

    def validate(self) -> None:
        if not isinstance(self.position, np.ndarray) or self.position.shape != (3,):
            raise ValueError("position must be a 3D numpy array")
        if not isinstance(self.target, np.ndarray) or self.target.shape != (3,):
            raise ValueError("target must be a 3D numpy array")
        if not (0.0 <= self.ambient_intensity <= 1.0):
            raise ValueError("ambient_intensity must be between 0.0 and 1.0")
        if not (0.0 <= self.diffuse_intensity <= 1.0):
            raise ValueError("diffuse_intensity must be between 0.0 and 1.0")

    @classmethod
    def directional(cls, direction, ambient: float = 0.2, diffuse: float = 0.8,
                    distance: float = 10.0) -> "Light":
        """Light placed at ``-normalize(direction) * distance``, aimed at the origin
        (reference light.py:68-110)."""
        d = np.array(direction, dtype=np.float32)
        d = d / np.linalg.norm(d)
        return cls(
            position=-d * distance,
            target=np.array([0.0, 0.0, 0.0], dtype=np.float32),
            ambient_intensity=ambient,
            diffuse_intensity=diffuse,
        )

    @classmethod
    def point_light(cls, position, target=None, ambient: float = 0.2,
                    diffuse: float = 0.8) -> "Light":
        pos = np.array(position, dtype=np.float32)
        tgt = (np.array([0.0, 0.0, 0.0], dtype=np.float32) if target is None
               else np.array(target, dtype=np.float32))
        return cls(position=pos, target=tgt, ambient_intensity=ambient,
                   diffuse_intensity=diffuse)

    @classmethod
    def default(cls) -> "Light":
        return cls()

    @classmethod
    def ambient_only(cls, intensity: float = 0.5) -> "Light":
        return cls(
            position=np.array([0.0, 0.0, 0.0], dtype=np.float32),
            target=np.array([0.0, 0.0, 0.0], dtype=np.float32),
            ambient_intensity=intensity,
            diffuse_intensity=0.0,
        )

    @classmethod
    def camera_linked(cls, azimuth_offset: float = 0.0, elevation_offset: float = 0.0,
                      distance_offset: float = 0.0, ambient: float = 0.2,
                      diffuse: float = 0.8) -> "Light":
        light = cls(
            position=np.array([1.0, 1.0, 1.0], dtype=np.float32),
            target=np.array([0.0, 0.0, 0.0], dtype=np.float32),
            ambient_intensity=ambient,
            diffuse_intensity=diffuse,
        )
        return light.link_to_camera(azimuth_offset, elevation_offset, distance_offset)

    def get_direction(self) -> np.ndarray:
        """Unit vector from position to target; (0, 0, -1) when they coincide."""
        d = self.target - self.position
        n = np.linalg.norm(d)
        if n < 1e-9:
            return np.array([0.0, 0.0, -1.0], dtype=np.float32)
        return d / n

    def copy(self) -> "Light":
        out = Light(
            position=self.position.copy(),
            target=self.target.copy(),
            ambient_intensity=self.ambient_intensity,
            diffuse_intensity=self.diffuse_intensity,
        )
        out._is_linked = self._is_linked
        if self._camera_offsets is not None:
            out._camera_offsets = self._camera_offsets.copy()
        return out

    @property
    def is_linked(self) -> bool:
        return self._is_linked

    def link_to_camera(self, azimuth_offset: float = 0.0, elevation_offset: float = 0.0,
                       distance_offset: float = 0.0) -> "Light":
        """Headlight mode; the offsets are stored but unused (as in the reference)."""
        self._is_linked = True
        self._camera_offsets = {
            "azimuth": azimuth_offset,
            "elevation": elevation_offset,
            "distance": distance_offset,
        }
        return self

    def unlink_from_camera(self) -> "Light":
        self._is_linked = False
        self._camera_offsets = None
        return self

    def update_from_camera(self, camera) -> None:
        """Move the light to the camera position, aimed at the camera target."""
        if not self._is_linked:
            raise ValueError("Light is not linked to camera. Call link_to_camera() first.")
        if self._camera_offsets is None:
            raise ValueError("Camera offsets not set. Call link_to_camera() first.")
        from ..camera import Camera

        if not isinstance(camera, Camera):
            raise ValueError("camera must be a Camera instance")
        pos, _ = camera.get_camera_vectors()
        self.position = pos.copy()
        self.target = camera.target.copy()

    def get_offsets(self) -> Optional[dict]:
        return self._camera_offsets.copy() if self._camera_offsets else None

    def __repr__(self) -> str:
        return (
            f"Light(position={self.position}, target={self.target}, "
            f"ambient={self.ambient_intensity:.2f}, diffuse={self.diffuse_intensity:.2f})"
        )


class LightError(Exception):
    """Raised for lighting-configuration errors."""
