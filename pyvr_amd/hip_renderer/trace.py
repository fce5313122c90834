"""roctx ranges around the host side of rendering and serving.

``rocprofv3 --marker-trace`` (with ``--kernel-trace``) shows them on the
timeline beside the kernels: ``hvr.render`` (one VolumeRenderer frame),
``hvr.batch.launch`` / ``hvr.batch.wait`` / ``hvr.batch.readback`` (a batch
slot), ``hvr.serve.collect`` and ``hvr.serve.admit`` (one pass of the
serving loop).  The ranges come from the ROCm profiler SDK's roctx library
(``librocprofiler-sdk-roctx``), which records nothing unless a profiler is
attached.  Without the library, or with ``HVR_ROCTX=0``, they are no-ops.
"""

import ctypes
import os
from contextlib import contextmanager

_CANDIDATES = ("librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so",
               os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "librocprofiler-sdk-roctx.so.1"))
_lib = None
_tried = False


def _roctx():
    global _lib, _tried
    if _tried:
        return _lib
    _tried = True
    if os.environ.get("HVR_ROCTX", "1") == "0":
        return None
    for name in _CANDIDATES:
        try:
            L = ctypes.CDLL(name)
        except OSError:
            continue
        L.roctxRangePushA.argtypes = [ctypes.c_char_p]
        L.roctxRangePushA.restype = ctypes.c_int
        L.roctxRangePop.argtypes = []
        L.roctxRangePop.restype = ctypes.c_int
        L.roctxMarkA.argtypes = [ctypes.c_char_p]
        L.roctxMarkA.restype = None
        _lib = L
        break
    return _lib


def enabled() -> bool:
    return _roctx() is not None


@contextmanager
def span(name: str):
    """A roctx range named ``name`` around the body (nested ranges nest)."""
    L = _roctx()
    if L is None:
        yield
        return
    L.roctxRangePushA(name.encode())
    try:
        yield
    finally:
        L.roctxRangePop()


def mark(message: str) -> None:
    """An instantaneous roctx marker."""
    L = _roctx()
    if L is not None:
        L.roctxMarkA(message.encode())
