"""pyvr_amd -- direct volume rendering on AMD Instinct MI355X (gfx950).

The Python surface of the reference PyVR render path (``Volume``,
``Camera``, ``Light``, ``RenderConfig``, the transfer functions and
``VolumeRenderer``) over hand-written HIP kernels in ``libhipvr.so``.
"""

from . import camera, config, datasets, lighting, transferfunctions, volume
from .config import RenderConfig

__version__ = "0.1.0"


def build(verbose: bool = False) -> str:
    """Compile the HIP library in-tree (hipcc, gfx950); returns its path."""
    from ._build import build as _b

    return _b(verbose)


def __getattr__(name):
    # the renderer is imported lazily so that the host-side classes work
    # without the HIP runtime
    if name in ("hip_renderer", "moderngl_renderer", "VolumeRenderer"):
        import importlib

        mod = importlib.import_module(".hip_renderer", __name__)
        if name == "VolumeRenderer":
            return mod.VolumeRenderer
        if name == "moderngl_renderer":
            return importlib.import_module(".moderngl_renderer", __name__)
        return mod
    raise AttributeError(name)


__all__ = ["camera", "config", "datasets", "lighting", "transferfunctions", "volume",
           "RenderConfig", "build", "VolumeRenderer"]
