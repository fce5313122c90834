"""Volume loaders (reference ``pyvr/dataloaders``)."""

from .vtk_loader import load_vtk_volume

__all__ = ["load_vtk_volume"]
