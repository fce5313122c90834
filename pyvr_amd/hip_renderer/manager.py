"""Device-resource manager: the HIP counterpart of the reference's
``ModernGLManager`` (``pyvr/moderngl_renderer/manager.py:8-256``).

It keeps the reference's programming model -- texture units, named shader
uniforms, clear/blend/draw/read -- so the renderer above it reads like the
reference, and maps it onto ``libhipvr.so``:

* a *texture unit* holds a 3-D texture (a field of a device volume) or the
  RGBA transfer-function LUT, numbered in creation order from 0
  (reference manager.py:232-236);
* the samplers ``volume_texture``, ``normal_volume`` and
  ``transfer_function_lut`` are int uniforms naming a unit; every uniform
  defaults to 0, exactly as in GL, so an unset ``normal_volume`` reads unit
  0 -- the first volume loaded -- as (scalar, 0, 0)
  (reference renderer.py:144-146);
* a sampler whose unit holds no texture of its kind reads like a GL
  incomplete texture: scalar 0, normal (0, 0, 0), LUT (0, 0, 0, 1);
* ``render_quad`` launches the ray-setup and march kernels with the current
  uniforms; the clear colour and blend state are applied in the kernel's
  store epilogue.

Resource policy (the reference never frees textures): whenever a texture is
created or a sampler uniform changes, textures that no sampler references
are released, except the most recently created one.  A sampler later
pointed at a released unit reads as unbound.
"""

import ctypes
from typing import Optional

import numpy as np

from . import _lib
from ._lib import HvrParams, check

# Uniforms of the march program (reference volume.frag.glsl:3-19), plus the
# early-termination alpha that the reference hard-codes at 0.99 (:87).
_SAMPLERS = ("volume_texture", "normal_volume", "transfer_function_lut")

# A bytes object of a given size whose contents the library fills before it
# is returned (the CPython C API allows writing a new bytes object's buffer
# until it is shared).
_new_bytes = ctypes.pythonapi.PyBytes_FromStringAndSize
_new_bytes.restype = ctypes.py_object
_new_bytes.argtypes = (ctypes.c_void_p, ctypes.c_ssize_t)
_bytes_data = ctypes.pythonapi.PyBytes_AsString
_bytes_data.restype = ctypes.c_void_p
_bytes_data.argtypes = (ctypes.py_object,)
_UNIFORMS = {
    "volume_texture": "int", "normal_volume": "int", "transfer_function_lut": "int",
    "view_matrix": "mat4", "projection_matrix": "mat4", "camera_pos": "vec3",
    "step_size": "float", "max_steps": "int", "reference_step_size": "float",
    "volume_min_bounds": "vec3", "volume_max_bounds": "vec3",
    "ambient_light": "float", "diffuse_light": "float",
    "light_position": "vec3", "light_target": "vec3",
    "opacity_threshold": "float",
}


class _MarchProgram:
    """Stands for the linked shader program: the march kernels compiled into
    libhipvr.so.  Uniform values live in ``values``."""

    def __init__(self):
        self.values = {}
        for name, kind in _UNIFORMS.items():
            if kind == "mat4":
                self.values[name] = np.zeros((4, 4), np.float32)
            elif kind == "vec3":
                self.values[name] = np.zeros(3, np.float32)
            elif kind == "int":
                self.values[name] = 0
            else:
                self.values[name] = 0.0
        self.values["opacity_threshold"] = 0.99
        self.matrices_written = set()

    def __getitem__(self, name):
        if name not in self.values:
            raise KeyError(name)
        return self.values[name]

    def release(self):
        pass


class _DeviceVolume:
    """One device volume (``hvr_volume``): scalar and/or normal field."""

    def __init__(self, L, ctx, shape, scalar_format, normal_format):
        self.L = L
        self.shape = tuple(int(s) for s in shape)
        self.scalar_format = scalar_format
        self.normal_format = normal_format
        h = ctypes.c_void_p()
        check(L.hvr_volume_create(ctx, self.shape[0], self.shape[1], self.shape[2],
                                  _lib.HVR_SCALAR[scalar_format], _lib.HVR_NORMAL[normal_format],
                                  ctypes.byref(h)), "hvr_volume_create")
        self.handle = h
        self.normals_loaded = False

    def upload_scalar(self, data: np.ndarray):
        a = np.ascontiguousarray(data, dtype=np.float32)
        check(self.L.hvr_volume_upload_scalar(self.handle, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
              "hvr_volume_upload_scalar")

    def upload_normals(self, data: np.ndarray):
        a = np.ascontiguousarray(data, dtype=np.float32)
        check(self.L.hvr_volume_upload_normals(self.handle, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
              "hvr_volume_upload_normals")
        self.normals_loaded = True

    def upload_scalar_normals(self, data: np.ndarray):
        """Scalar field, and its normals computed on the device exactly as
        compute_normal_volume computes them from the float32 array."""
        a = np.ascontiguousarray(data, dtype=np.float32)
        check(self.L.hvr_volume_upload_scalar_normals(self.handle,
                                                      a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
              "hvr_volume_upload_scalar_normals")
        self.normals_loaded = True

    def generate_blobs(self, centers, sigmas, amps, noise: float, seed: int):
        """Seeded Gaussian-blob field (and normals, if held) built on the device."""
        c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        sg = np.ascontiguousarray(sigmas, dtype=np.float64).reshape(-1)
        am = np.ascontiguousarray(amps, dtype=np.float64).reshape(-1)
        if not (len(c) == len(sg) == len(am)):
            raise ValueError("centers, sigmas and amps must describe the same blobs")
        dp = ctypes.POINTER(ctypes.c_double)
        check(self.L.hvr_volume_generate_blobs(self.handle, len(sg), c.ctypes.data_as(dp), sg.ctypes.data_as(dp),
                                               am.ctypes.data_as(dp), float(noise), int(seed)),
              "hvr_volume_generate_blobs")
        self.normals_loaded = self.normal_format != "none"

    def nbytes(self) -> int:
        v = ctypes.c_uint64()
        check(self.L.hvr_volume_bytes(self.handle, ctypes.byref(v)), "hvr_volume_bytes")
        return int(v.value)

    def download(self, scalar: bool = True, normals: bool = False):
        """The stored fields as float32 numpy arrays of the volume's shape
        (``hvr_volume_download``): the scalar as the march reads it, the
        normals as stored (SNORM8 as integers).  Returns (scalar, normals),
        None for a field not asked for."""
        fp = ctypes.POINTER(ctypes.c_float)
        s = np.empty(self.shape, np.float32) if scalar else None
        n = np.empty(self.shape + (3,), np.float32) if normals else None
        check(self.L.hvr_volume_download(self.handle, s.ctypes.data_as(fp) if s is not None else None,
                                         n.ctypes.data_as(fp) if n is not None else None), "hvr_volume_download")
        return s, n

    def destroy(self):
        if self.handle:
            self.L.hvr_volume_destroy(self.handle)
            self.handle = None


class _Texture3D:
    def __init__(self, volume: _DeviceVolume, role: int, components: int):
        self.volume = volume
        self.role = role
        self.components = components


class _TextureLUT:
    def __init__(self, rgba: np.ndarray):
        self.rgba = np.ascontiguousarray(rgba, dtype=np.float32)


_FORMAT_ALIASES = {
    None: None, "f32": "f32", "float32": "f32", "f16": "f16", "float16": "f16", "half": "f16",
    "u8": "u8", "unorm8": "u8", "uint8": "u8",
    "f32x3": "f32x3", "sn8x3": "sn8x3", "snorm8x3": "sn8x3", "f16x3": "f16x3", "float16x3": "f16x3",
}


class HipManager:
    """HIP device resources for one renderer: framebuffer, volumes, LUT,
    uniforms.  Method names and errors follow ``ModernGLManager``."""

    def __init__(self, width=512, height=512, device: int = 0):
        self.width = width
        self.height = height
        self.device = device
        self._released = True
        self.L = _lib.lib()  # the library this context lives in
        h = ctypes.c_void_p()
        check(self.L.hvr_create(int(device), int(width), int(height), ctypes.byref(h)), "hvr_create")
        self.ctx = h
        self.program = None
        self._next_texture_unit = 0
        self._units = {}
        self._texture_bindings = {}
        self._newest_unit = None
        self._volumes = []
        self._clear = (0, 0, 0, 0)
        self._blend = False
        self._lut_uploaded = None
        self._released = False

    # -- program -------------------------------------------------------------
    def load_shaders(self, vertex_shader_path=None, fragment_shader_path=None):
        """Make the march program current.  The kernels are compiled into
        libhipvr.so, so no shader source is read; the paths are accepted for
        API compatibility and ignored."""
        self.program = _MarchProgram()
        self._sync_bindings()

    def _need_program(self):
        if self.program is None:
            raise RuntimeError("Shader program not loaded")

    # -- textures ------------------------------------------------------------
    def _get_next_texture_unit(self):
        unit = self._next_texture_unit
        self._next_texture_unit += 1
        return unit

    def _add_texture(self, tex) -> int:
        unit = self._get_next_texture_unit()
        self._units[unit] = tex
        self._newest_unit = unit
        self._collect()
        return unit

    def create_volume_texture(self, volume_data, scalar_format: str = "f32",
                              normal_format: Optional[str] = None, device_normals: bool = False) -> int:
        """Upload a 3-D scalar field (reference manager.py:77-108) and return
        its texture unit.  Any dtype is converted to float32 first.
        ``scalar_format`` selects the device storage ("f32", "f16", "u8");
        ``normal_format`` ("f32x3", "sn8x3", "f16x3" with "f16") reserves an interleaved normal
        field that the next ``create_normal_texture`` of the same shape fills,
        or, with ``device_normals``, that the device computes from the field
        (``create_device_normal_texture`` then binds it)."""
        if len(volume_data.shape) != 3:
            raise ValueError("Volume data must be 3D")
        sf = _FORMAT_ALIASES.get(scalar_format)
        nf = _FORMAT_ALIASES.get(normal_format)
        if sf not in ("f32", "f16", "u8") or (nf not in (None, "f32x3", "sn8x3", "f16x3")):
            raise ValueError(f"unsupported formats scalar={scalar_format!r} normal={normal_format!r}")
        _lib.check_formats(sf, nf or "none")
        if device_normals and nf is None:
            raise ValueError("device_normals needs a normal_format")
        dv = _DeviceVolume(self.L, self.ctx, volume_data.shape, sf, nf or "none")
        self._volumes.append(dv)
        if device_normals:
            dv.upload_scalar_normals(np.asarray(volume_data, dtype=np.float32))
        else:
            dv.upload_scalar(np.asarray(volume_data, dtype=np.float32))
        unit = self._add_texture(_Texture3D(dv, _lib.HVR_ROLE_SCALAR, 1))
        self._texture_bindings["volume"] = (self._units[unit], unit)
        return unit

    def create_normal_texture(self, normal_data, normal_format: str = "f32x3") -> int:
        """Upload a normal field of shape (s0, s1, s2, 3) (reference
        manager.py:110-135) and return its texture unit."""
        if normal_data.shape[-1] != 3:
            raise ValueError("Normal volume must have 3 channels (last dimension).")
        if len(normal_data.shape) != 4:
            raise ValueError("Normal volume must have shape (s0, s1, s2, 3)")
        nf = _FORMAT_ALIASES.get(normal_format)
        if nf not in ("f32x3", "sn8x3", "f16x3"):
            raise ValueError(f"unsupported normal format {normal_format!r}")
        shape = tuple(normal_data.shape[:3])
        data = np.asarray(normal_data, dtype=np.float32)
        target = None
        newest = self._units.get(self._newest_unit)
        if (isinstance(newest, _Texture3D) and newest.role == _lib.HVR_ROLE_SCALAR
                and newest.volume.normal_format == nf and newest.volume.shape == shape
                and not newest.volume.normals_loaded):
            target = newest.volume           # interleave into the reserved texels
        if target is None:
            target = _DeviceVolume(self.L, self.ctx, shape, "none", nf)
            self._volumes.append(target)
        target.upload_normals(data)
        unit = self._add_texture(_Texture3D(target, _lib.HVR_ROLE_NORMAL, 3))
        self._texture_bindings["normal"] = (self._units[unit], unit)
        return unit

    def create_device_normal_texture(self) -> int:
        """Texture unit of the normals the device computed for the newest
        volume texture (``create_volume_texture(..., device_normals=True)``)."""
        newest = self._units.get(self._newest_unit)
        if not (isinstance(newest, _Texture3D) and newest.role == _lib.HVR_ROLE_SCALAR
                and newest.volume.normals_loaded):
            raise RuntimeError("the newest volume texture has no device-computed normals")
        unit = self._add_texture(_Texture3D(newest.volume, _lib.HVR_ROLE_NORMAL, 3))
        self._texture_bindings["normal"] = (self._units[unit], unit)
        return unit

    def create_rgba_transfer_function_texture(self, color_transfer_function,
                                              opacity_transfer_function,
                                              size: Optional[int] = None) -> int:
        """N x 4 RGBA LUT from the two transfer functions (reference
        manager.py:137-186): N = ``size`` or the larger ``lut_size``."""
        if size is None:
            size = max(color_transfer_function.lut_size, opacity_transfer_function.lut_size)
        color_lut = color_transfer_function.to_lut(size)
        opacity_lut = opacity_transfer_function.to_lut(size)
        rgba = np.empty((size, 4), dtype=np.float32)
        rgba[:, :3] = color_lut
        rgba[:, 3] = opacity_lut
        return self._add_texture(_TextureLUT(rgba))

    def _collect(self):
        """Release textures that no sampler references (keeping the newest)
        and device volumes that no texture references."""
        # a sampler keeps only a texture of its own kind alive
        keep = {self._newest_unit}
        if self.program is not None:
            v = self.program.values
            for name, kind in (("volume_texture", _Texture3D), ("normal_volume", _Texture3D),
                               ("transfer_function_lut", _TextureLUT)):
                if isinstance(self._units.get(int(v[name])), kind):
                    keep.add(int(v[name]))
        else:
            keep.add(0)
        for unit in list(self._units):
            if unit not in keep:
                del self._units[unit]
        live = {id(t.volume) for t in self._units.values() if isinstance(t, _Texture3D)}
        for dv in list(self._volumes):
            if id(dv) not in live:
                dv.destroy()
                self._volumes.remove(dv)
        for key, (tex, unit) in list(self._texture_bindings.items()):
            if unit not in self._units:
                del self._texture_bindings[key]

    # -- uniforms ------------------------------------------------------------
    def set_uniform_matrix(self, name, matrix):
        self._need_program()
        m = np.asarray(matrix, dtype=np.float32)
        if m.size != 16:
            raise ValueError(f"uniform {name} expects 16 values")
        self.program[name]  # KeyError for unknown uniforms, as in ModernGL
        self.program.values[name] = m.reshape(4, 4).copy()
        self.program.matrices_written.add(name)

    def set_uniform_vector(self, name, vector):
        self._need_program()
        self.program[name]
        v = np.asarray(tuple(vector), dtype=np.float32)
        if v.shape != (3,):
            raise ValueError(f"uniform {name} expects 3 values")
        self.program.values[name] = v

    def set_uniform_float(self, name, value):
        self._need_program()
        self.program[name]
        self.program.values[name] = float(value)

    def set_uniform_int(self, name, value):
        self._need_program()
        self.program[name]
        self.program.values[name] = int(value)
        if name in _SAMPLERS:
            self._collect()

    # -- drawing -------------------------------------------------------------
    def clear_framebuffer(self, r=0.0, g=0.0, b=0.0, a=0.0):
        """Set the clear colour of the next frame (8-bit, round to nearest)."""
        self._clear = tuple(int(np.floor(min(max(float(c), 0.0), 1.0) * 255.0 + 0.5)) for c in (r, g, b, a))

    def setup_blending(self):
        """Enable SRC_ALPHA, ONE_MINUS_SRC_ALPHA blending (applied in the
        kernel's store epilogue)."""
        self._blend = True

    def _texture_at(self, unit, kind):
        t = self._units.get(int(unit))
        return t if isinstance(t, kind) else None

    def _sync_bindings(self):
        L = self.L
        if self.program is None:
            check(L.hvr_bind(self.ctx, None, 0, None, 0), "hvr_bind")
            return
        v = self.program.values
        ts = self._texture_at(v["volume_texture"], _Texture3D)
        tn = self._texture_at(v["normal_volume"], _Texture3D)
        check(L.hvr_bind(self.ctx,
                         ts.volume.handle if ts else None, ts.role if ts else 0,
                         tn.volume.handle if tn else None, tn.role if tn else 0), "hvr_bind")
        lut = self._texture_at(v["transfer_function_lut"], _TextureLUT)
        key = id(lut) if lut is not None else None
        if key != self._lut_uploaded:
            if lut is None:
                check(L.hvr_set_lut(self.ctx, None, 0), "hvr_set_lut")
            else:
                check(L.hvr_set_lut(self.ctx, lut.rgba.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                    lut.rgba.shape[0]), "hvr_set_lut")
            self._lut_uploaded = key

    def _params(self) -> HvrParams:
        v = self.program.values
        p = HvrParams()
        p.view[:] = v["view_matrix"].ravel().tolist()
        p.proj[:] = v["projection_matrix"].ravel().tolist()
        p.cam_pos[:] = v["camera_pos"].tolist()
        p.bmin[:] = v["volume_min_bounds"].tolist()
        p.bmax[:] = v["volume_max_bounds"].tolist()
        p.step = v["step_size"]
        p.ref_step = v["reference_step_size"]
        p.max_steps = int(v["max_steps"])
        p.ambient = v["ambient_light"]
        p.diffuse = v["diffuse_light"]
        p.light_pos[:] = v["light_position"].tolist()
        p.light_target[:] = v["light_target"].tolist()
        p.et_threshold = v["opacity_threshold"]
        p.camera_set = int({"view_matrix", "projection_matrix"} <= self.program.matrices_written)
        p.clear_rgba[:] = list(self._clear)
        p.blend = int(self._blend)
        return p

    def render_quad(self):
        """Enqueue one frame: ray setup + march + blend/store."""
        if self.program is None:
            raise RuntimeError("Vertex array object not created")
        self._sync_bindings()
        p = self._params()
        if not (p.step > 0 and p.ref_step > 0):
            raise RuntimeError("step_size and reference_step_size must be set and positive")
        L = self.L
        check(L.hvr_set_params(self.ctx, ctypes.byref(p)), "hvr_set_params")
        check(L.hvr_render_async(self.ctx), "hvr_render_async")

    def read_pixels(self) -> bytes:
        """Framebuffer contents: width*height*4 bytes, RGBA8, bottom row first.
        The library copies the image straight into a new bytes object
        (hvr_read: chunked through pinned memory, each chunk's host copy
        overlapping the next chunk's transfer), so no further copy is made."""
        n = self.width * self.height * 4
        out = _new_bytes(None, n)
        check(self.L.hvr_read(self.ctx, ctypes.cast(_bytes_data(out), ctypes.POINTER(ctypes.c_uint8))), "hvr_read")
        return out

    def sample_counts(self) -> np.ndarray:
        """Samples composited per pixel for the current state (H, W) uint32,
        bottom row first (a separate counting pass)."""
        self._sync_bindings()
        p = self._params()
        L = self.L
        check(L.hvr_set_params(self.ctx, ctypes.byref(p)), "hvr_set_params")
        out = np.zeros((self.height, self.width), np.uint32)
        check(L.hvr_sample_counts(self.ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
              "hvr_sample_counts")
        return out

    def last_kernel_ms(self):
        """(march_ms, setup_ms) of the last frame, from HIP events."""
        a, b = ctypes.c_float(), ctypes.c_float()
        check(self.L.hvr_last_kernel_ms(self.ctx, ctypes.byref(a), ctypes.byref(b)), "hvr_last_kernel_ms")
        return a.value, b.value

    def cleanup(self):
        """Release every device resource (safe to call more than once)."""
        if self._released:
            return
        self._units.clear()
        self._texture_bindings.clear()
        for dv in self._volumes:
            dv.destroy()
        self._volumes = []
        if self.ctx:
            self.L.hvr_destroy(self.ctx)
        self.ctx = None
        self._released = True

    def __del__(self):
        try:
            self.cleanup()
        except Exception:
            pass
