"""Camera parameters, placement, keyframed paths and interactive control
(API of the reference's ``pyvr.camera``)."""

from .camera import (
    Camera,
    CameraError,
    degrees_to_radians,
    radians_to_degrees,
    validate_camera_angles,
)
from .control import CameraController, CameraPath, get_camera_pos, get_camera_pos_from_params

__all__ = [
    "CameraController",
    "CameraPath",
    "get_camera_pos",
    "get_camera_pos_from_params",
    "Camera",
    "CameraError",
    "validate_camera_angles",
    "degrees_to_radians",
    "radians_to_degrees",
]
