"""Lighting for the HIP volume renderer (API of the reference's ``pyvr.lighting``)."""

from .light import Light, LightError

__all__ = ["Light", "LightError"]
