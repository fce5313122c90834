"""HIP renderer: ``VolumeRenderer`` (drop-in for the reference's
``pyvr.moderngl_renderer.VolumeRenderer``), its device-resource manager, and
``BatchRenderer`` (many requests per launch)."""

from ..camera import Camera
from ..datasets import compute_normal_volume, create_sample_volume
from ..transferfunctions.color import ColorTransferFunction
from ..transferfunctions.opacity import OpacityTransferFunction
from .batch import BatchRenderer
from .manager import HipManager
from .renderer import HipVolumeRenderer, VolumeRenderer

__all__ = [
    "ColorTransferFunction",
    "OpacityTransferFunction",
    "HipVolumeRenderer",
    "VolumeRenderer",
    "HipManager",
    "BatchRenderer",
    "Camera",
    "create_sample_volume",
    "compute_normal_volume",
]
