"""Camera parameters and placement (API of the reference's ``pyvr.camera``).

The interactive helpers of the reference (``CameraPath``, ``CameraController``,
trackball) drive the UI, not the render path, and are not part of this package.
"""

from .camera import (
    Camera,
    CameraError,
    degrees_to_radians,
    radians_to_degrees,
    validate_camera_angles,
)
from .control import get_camera_pos, get_camera_pos_from_params

__all__ = [
    "get_camera_pos",
    "get_camera_pos_from_params",
    "Camera",
    "CameraError",
    "validate_camera_angles",
    "degrees_to_radians",
    "radians_to_degrees",
]
