"""One request per launch over resident volumes: the single-frame kernels
that ``VolumeRenderer.render()`` runs (ray setup, march, blend and
readback; reference ``pyvr/moderngl_renderer/renderer.py:209-219``), for
requests of any image size on volumes kept on the device.

The reference creates one ``ModernGLVolumeRenderer`` per image size and
uploads each volume into it (``manager.py:77-135``).  ``PlainRenderer`` keeps
one context that owns every resident volume and resizes its framebuffer per
request (``hvr_resize``), so a request stream over many volumes and sizes is
rendered one launch per request without re-uploading.  Each image is the one
a fresh ``VolumeRenderer`` of that size gives for the same volume, camera,
transfer functions and settings.  It is the "plain" side of the batched and
served measurements, and the reference renders the full-size batch and
serving tests compare against.
"""

import ctypes
from typing import Dict, Optional

import numpy as np

from ..config import REFERENCE_TERMINATION_ALPHA, RenderConfig
from ..lighting import Light
from . import _lib
from ._lib import check
from .batch import DeviceBlobVolume, ResidentVolume, request_params, rgba_lut  # noqa: F401 (re-export)
from .manager import _bytes_data, _new_bytes


class PlainRenderer:
    def __init__(self, device: int = 0, scalar_format: str = "f16", normal_format: str = "sn8x3",
                 brick_log2: int = 3):
        self.L = _lib.lib()
        self.scalar_format, self.normal_format = scalar_format, normal_format
        self.ctx = ctypes.c_void_p()
        check(self.L.hvr_create(device, 1, 1, ctypes.byref(self.ctx)), "hvr_create")
        self.size = (1, 1)
        self._resident: Dict[int, ResidentVolume] = {}
        self._bound = None      # handle of the bound volume
        self._lut = None        # bytes of the LUT last set
        self._last = None       # (volume, config, light, threshold) of the prepared request
        try:
            check(self.L.hvr_set_brick_log2(self.ctx, brick_log2), "hvr_set_brick_log2")
        except Exception:
            self.close()
            raise

    def upload(self, volume, device_normals: bool = False) -> ResidentVolume:
        """Make ``volume`` (``Volume`` or ``DeviceBlobVolume``) resident, once
        per object."""
        rv = self._resident.get(id(volume))
        if rv is None or rv.volume is not volume:
            rv = ResidentVolume(self.L, self.ctx, volume, self.scalar_format, self.normal_format, device_normals)
            self._resident[id(volume)] = rv
        return rv

    def release(self, volume):
        rv = self._resident.pop(id(volume), None)
        if rv is not None:
            if self._bound is not None and self._bound == rv.handle.value:
                check(self.L.hvr_bind(self.ctx, None, _lib.HVR_ROLE_SCALAR, None, _lib.HVR_ROLE_SCALAR), "hvr_bind")
                self._bound = None
            rv.release()

    def resident_bytes(self) -> int:
        return sum(rv.nbytes() for rv in self._resident.values())

    def prepare(self, volume, camera, color_transfer_function, opacity_transfer_function, width: int = 512,
                height: int = 512, config: Optional[RenderConfig] = None, light: Optional[Light] = None,
                opacity_threshold: float = REFERENCE_TERMINATION_ALPHA, lut_size: Optional[int] = None):
        """Bind the request's volume, LUT and parameters (sizing the
        framebuffer); the next ``launch`` renders it."""
        config = RenderConfig.balanced() if config is None else config
        light = Light.default() if light is None else light
        rv = self.upload(volume)
        if (width, height) != self.size:
            check(self.L.hvr_resize(self.ctx, int(width), int(height)), "hvr_resize")
            self.size = (int(width), int(height))
        if self._bound != rv.handle.value:
            nrole = _lib.HVR_ROLE_NORMAL if rv.has_normals else _lib.HVR_ROLE_SCALAR
            check(self.L.hvr_bind(self.ctx, rv.handle, _lib.HVR_ROLE_SCALAR, rv.handle, nrole), "hvr_bind")
            self._bound = rv.handle.value
        lut = rgba_lut(color_transfer_function, opacity_transfer_function, lut_size)
        if self._lut != lut.tobytes():
            check(self.L.hvr_set_lut(self.ctx, lut.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), lut.shape[0]),
                  "hvr_set_lut")
            self._lut = lut.tobytes()
        self.params = request_params(volume, camera, config, light, width, height, opacity_threshold)
        check(self.L.hvr_set_params(self.ctx, ctypes.byref(self.params)), "hvr_set_params")
        self._last = (volume, config, light, opacity_threshold)

    def set_camera(self, camera):
        """Only the camera of the prepared request changes (what
        ``VolumeRenderer.set_camera`` does between frames)."""
        if self._last is None:
            raise RuntimeError("no request prepared")
        volume, config, light, threshold = self._last
        w, h = self.size
        self.params = request_params(volume, camera, config, light, w, h, threshold)
        check(self.L.hvr_set_params(self.ctx, ctypes.byref(self.params)), "hvr_set_params")

    def read(self) -> bytes:
        """One frame of the prepared request, read back into a new bytes
        object (width*height*4 RGBA8, bottom row first), as
        ``VolumeRenderer.render()`` returns it."""
        w, h = self.size
        check(self.L.hvr_render_async(self.ctx), "hvr_render_async")
        out = _new_bytes(None, w * h * 4)
        check(self.L.hvr_read(self.ctx, ctypes.cast(_bytes_data(out), ctypes.POINTER(ctypes.c_uint8))), "hvr_read")
        return out

    def read_pinned(self) -> np.ndarray:
        """One frame of the prepared request, copied to the context's pinned
        host buffer: a read-only uint8 view (width*height*4) valid until the
        next frame -- the pixels in host memory without a further copy, as a
        served request's are."""
        w, h = self.size
        check(self.L.hvr_render_async(self.ctx), "hvr_render_async")
        ptr = ctypes.POINTER(ctypes.c_uint8)()
        check(self.L.hvr_read_pinned(self.ctx, ctypes.byref(ptr)), "hvr_read_pinned")
        a = np.ctypeslib.as_array(ptr, shape=(w * h * 4,))
        a.flags.writeable = False
        return a

    def render(self, *args, **kw) -> bytes:
        """``prepare`` then one frame, read back: width*height*4 RGBA8 bytes,
        bottom row first (what ``VolumeRenderer.render()`` returns)."""
        self.prepare(*args, **kw)
        return self.read()

    def launch_to(self, device_rgba8):
        """Enqueue the prepared request's frame into a device buffer of
        width*height*4 bytes (no wait, no readback)."""
        check(self.L.hvr_render_async_to(self.ctx, device_rgba8), "hvr_render_async_to")

    def sample_counts(self) -> np.ndarray:
        """Composited samples per pixel (H, W) for the prepared request."""
        w, h = self.size
        c = np.empty(w * h, np.uint32)
        check(self.L.hvr_sample_counts(self.ctx, c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
              "hvr_sample_counts")
        return c.reshape(h, w)

    def shaded_sample_counts(self) -> np.ndarray:
        """Of those, the samples that load texels (not stepped over in
        transparent bricks), per pixel (H, W)."""
        w, h = self.size
        c = np.empty(w * h, np.uint32)
        check(self.L.hvr_shaded_sample_counts(self.ctx, c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
              "hvr_shaded_sample_counts")
        return c.reshape(h, w)

    def touched_bytes(self) -> int:
        """Compulsory texel bytes of the prepared request's frame: the bricks
        its samples load texels from, times the bytes of a brick."""
        t, n, b = ctypes.c_ulonglong(), ctypes.c_ulonglong(), ctypes.c_ulonglong()
        check(self.L.hvr_touched_bricks(self.ctx, ctypes.byref(t), ctypes.byref(n), ctypes.byref(b)),
              "hvr_touched_bricks")
        return int(t.value) * int(b.value)

    def download(self, volume, normals: bool = True):
        """The resident volume's stored fields (scalar as the march reads it,
        normals as stored) as float32 arrays, for checking on the CPU."""
        rv = self.upload(volume)
        return rv.device.download(scalar=True, normals=normals and rv.has_normals)

    def close(self):
        for rv in getattr(self, "_resident", {}).values():
            rv.release()
        self._resident = {}
        if getattr(self, "ctx", None):
            self.L.hvr_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["PlainRenderer", "DeviceBlobVolume"]
