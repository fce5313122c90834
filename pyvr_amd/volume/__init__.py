"""Volume container (API of the reference's ``pyvr.volume``)."""

from .data import Volume, VolumeError

__all__ = ["Volume", "VolumeError"]
