"""Transfer functions (API of the reference's ``pyvr.transferfunctions``)."""

from .base import BaseTransferFunction, InvalidControlPointError, TransferFunctionError
from .color import ColorTransferFunction
from .opacity import OpacityTransferFunction

__all__ = [
    "BaseTransferFunction",
    "ColorTransferFunction",
    "OpacityTransferFunction",
    "TransferFunctionError",
    "InvalidControlPointError",
]
