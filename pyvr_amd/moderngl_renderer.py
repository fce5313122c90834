"""Import-compatible name for code written against the reference's
``pyvr.moderngl_renderer``: ``from pyvr_amd.moderngl_renderer import
VolumeRenderer`` gives the HIP renderer.  No OpenGL is involved."""

from .hip_renderer import *  # noqa: F401,F403
from .hip_renderer import HipManager as ModernGLManager  # noqa: F401
from .hip_renderer import HipVolumeRenderer as ModernGLVolumeRenderer  # noqa: F401
from .hip_renderer import __all__ as _all

__all__ = list(_all) + ["ModernGLManager", "ModernGLVolumeRenderer"]
