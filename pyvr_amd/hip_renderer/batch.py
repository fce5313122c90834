"""Batched rendering: many render requests in one persistent, graph-captured
launch (``hvr_batch_*`` in ``include/hipvr.h``).

The reference renders one frame per ``VolumeRenderer.render()`` call
(``pyvr/moderngl_renderer/renderer.py:209-219``) and has no batching; this
module adds it on the same scene objects.  A request is what one
``VolumeRenderer`` would render after ``load_volume``, ``set_camera``,
``set_transfer_functions`` (and ``set_opacity_threshold``) on a fresh
renderer of that width and height: each request's bytes equal that
renderer's ``render()`` output, sample counts included.

    >>> batch = BatchRenderer(scalar_format="f16", normal_format="sn8x3")
    >>> i = batch.add(volume, camera, ctf, otf, width=512, height=512)
    >>> batch.run()
    >>> rgba = batch.result(i)        # width*height*4 bytes, bottom row first
"""

import ctypes
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from ..camera import Camera
from ..config import REFERENCE_TERMINATION_ALPHA, RenderConfig
from ..lighting import Light
from ..volume import Volume
from . import _lib, trace
from ._lib import HvrParams, check
from .manager import _FORMAT_ALIASES, _DeviceVolume

TILE_SHAPES = {"8x8": _lib.TILE_8X8, "16x4": _lib.TILE_16X4}


def rgba_lut(color_transfer_function, opacity_transfer_function, size: Optional[int] = None) -> np.ndarray:
    """The (N, 4) float32 RGBA LUT a renderer builds from the two transfer
    functions (reference ``manager.py:137-186``)."""
    if size is None:
        size = max(color_transfer_function.lut_size, opacity_transfer_function.lut_size)
    rgba = np.empty((size, 4), dtype=np.float32)
    rgba[:, :3] = color_transfer_function.to_lut(size)
    rgba[:, 3] = opacity_transfer_function.to_lut(size)
    return rgba


def _vec3(v) -> list:
    a = np.asarray(tuple(v), dtype=np.float32)
    if a.shape != (3,):
        raise ValueError("expected 3 values")
    return a.tolist()


def request_params(volume: Volume, camera: Camera, config: RenderConfig, light: Light,
                   width: int, height: int,
                   opacity_threshold: float = REFERENCE_TERMINATION_ALPHA) -> HvrParams:
    """``hvr_params`` of one request, converted exactly as the renderer's
    uniform setters convert them (float32 matrices and vectors)."""
    p = HvrParams()
    # the camera vectors once (get_view_matrix would compute them again)
    position, up = camera.get_camera_vectors()
    view = np.asarray(camera._view_matrix(position, up), dtype=np.float32).reshape(4, 4)
    proj = np.asarray(camera.get_projection_matrix(width / height), dtype=np.float32).reshape(4, 4)
    p.view[:] = view.ravel().tolist()
    p.proj[:] = proj.ravel().tolist()
    p.cam_pos[:] = _vec3(position)
    p.bmin[:] = _vec3(volume.min_bounds)
    p.bmax[:] = _vec3(volume.max_bounds)
    p.step = float(config.step_size)
    p.ref_step = float(config.reference_step_size)
    p.max_steps = int(config.max_steps)
    p.ambient = float(light.ambient_intensity)
    p.diffuse = float(light.diffuse_intensity)
    p.light_pos[:] = _vec3(light.position)
    p.light_target[:] = _vec3(light.target)
    p.et_threshold = float(opacity_threshold)
    p.camera_set = 1
    p.clear_rgba[:] = [0, 0, 0, 0]
    p.blend = 1
    return p


@dataclass
class DeviceBlobVolume:
    """A volume generated on the device (``hvr_volume_generate_blobs``):
    the seeded Gaussian-blob family of ``pyvr_amd.datasets.blob_field``
    (centres, widths and amplitudes drawn exactly as it draws them) with
    noise from a counter-based hash, normals computed on the device.  It
    has no host data; it stands in for a ``Volume`` in ``BatchRenderer``."""
    shape: Tuple[int, int, int]
    n_blobs: int = 8
    seed: int = 1
    noise: float = 1e-3
    with_normals: bool = True
    min_bounds: np.ndarray = field(default_factory=lambda: np.array([-1.0, -1.0, -1.0], np.float32))
    max_bounds: np.ndarray = field(default_factory=lambda: np.array([1.0, 1.0, 1.0], np.float32))

    @property
    def has_normals(self) -> bool:
        return self.with_normals

    def blob_params(self):
        from ..datasets.synthetic import blob_params
        _, centers, sigmas, amps = blob_params(self.n_blobs, self.seed)
        return centers, sigmas, amps


class ResidentVolume:
    """A ``Volume`` (or ``DeviceBlobVolume``) on the device, shared by
    requests.  ``device_normals``: compute the normals of a volume without
    normals on the device (the bytes compute_normal_volume would give)."""

    def __init__(self, L, ctx, volume, scalar_format: str, normal_format: str, device_normals: bool = False):
        generated = isinstance(volume, DeviceBlobVolume)
        self.has_normals = volume.has_normals or (device_normals and not generated)
        sf = _FORMAT_ALIASES.get(scalar_format)
        nf = _FORMAT_ALIASES.get(normal_format) if self.has_normals else "none"
        if sf not in _lib.HVR_SCALAR or sf == "none" or nf not in _lib.HVR_NORMAL:
            raise ValueError(f"unsupported formats scalar={scalar_format!r} normal={normal_format!r}")
        _lib.check_formats(sf, nf)
        self.volume = volume
        shape = tuple(volume.shape) if generated else volume.data.shape
        self.device = _DeviceVolume(L, ctx, shape, sf, nf)
        try:
            if generated:
                c, sg, am = volume.blob_params()
                self.device.generate_blobs(c, sg, am, volume.noise, volume.seed)
            elif volume.has_normals:
                self.device.upload_scalar(volume.data)
                self.device.upload_normals(volume.normals)
            elif device_normals:
                self.device.upload_scalar_normals(volume.data)
            else:
                self.device.upload_scalar(volume.data)
        except Exception:
            self.device.destroy()
            raise

    @property
    def handle(self):
        return self.device.handle

    def nbytes(self) -> int:
        return self.device.nbytes()

    def release(self):
        self.device.destroy()


# Kernel variants of the batched march (hvr_batch_set_kernel): waves per
# block, transparent-brick skipping, normals gated on non-zero alpha.
KERNEL_VARIANTS = {"default": (4, True, True), "waves2": (2, True, True), "waves8": (8, True, True),
                   "nogate": (4, True, False), "noskip": (4, False, True)}


# Where a batch's images go (hvr_batch_set_readback): a device pool copied
# to pinned host memory after the launch, or the pinned pool directly.
READBACK = {"copy": _lib.READBACK_COPY, "direct": _lib.READBACK_DIRECT}


class _Slot:
    """One ``hvr_batch`` (own stream and buffers) and its requests' sizes."""

    def __init__(self, L, ctx, tile_shape, kernel="default", readback="direct"):
        self.L = L
        self.handle = ctypes.c_void_p()
        check(L.hvr_batch_create(ctx, ctypes.byref(self.handle)), "hvr_batch_create")
        check(L.hvr_batch_set_tile_shape(self.handle, TILE_SHAPES[tile_shape]), "hvr_batch_set_tile_shape")
        w, sk, g = KERNEL_VARIANTS[kernel]
        check(L.hvr_batch_set_kernel(self.handle, w, int(sk), int(g)), "hvr_batch_set_kernel")
        check(L.hvr_batch_set_readback(self.handle, READBACK[readback]), "hvr_batch_set_readback")
        self.sizes: List[tuple] = []
        self.in_flight = False

    def destroy(self):
        if self.handle:
            self.L.hvr_batch_destroy(self.handle)
            self.handle = None


class BatchRenderer:
    """Renders many requests per launch on one GPU.

    ``scalar_format`` / ``normal_format`` select the device storage of the
    volumes it uploads ("f32"/"f16"/"u8", "f32x3"/"sn8x3", or "f16x3"
    with "f16"); ``brick_log2``
    the brick edge (3: 8^3, 4: 16^3), ``tile_shape`` the ray tile of a
    wavefront ("8x8" or "16x4") and ``kernel`` the kernel variant (a key of
    ``KERNEL_VARIANTS``: waves per block, skipping, gating; every variant
    gives the same images) -- the choices the autotuner makes.
    ``readback`` ("direct" or "copy", a key of ``READBACK``): whether the
    march writes the images straight to pinned host memory (no copy after
    the launch, and no copy engine reading HBM while the next batch
    marches), or to a device pool copied to host memory after the launch
    (same bytes).

    ``pipeline`` batches can be in flight at once (each has its own stream
    and buffers): ``submit()`` launches the batch being filled and moves on
    to the next, so that one batch's copy back and tail overlap the next
    one's work.  ``run()`` renders the batch being filled and waits.
    """

    def __init__(self, device: int = 0, scalar_format: str = "f16", normal_format: str = "sn8x3",
                 brick_log2: int = 3, tile_shape: str = "8x8", pipeline: int = 2, kernel: str = "default",
                 readback: str = "direct"):
        if tile_shape not in TILE_SHAPES:
            raise ValueError(f"tile_shape must be one of {sorted(TILE_SHAPES)}")
        if kernel not in KERNEL_VARIANTS:
            raise ValueError(f"kernel must be one of {sorted(KERNEL_VARIANTS)}")
        if pipeline < 1:
            raise ValueError("pipeline must be at least 1")
        if readback not in READBACK:
            raise ValueError(f"readback must be one of {sorted(READBACK)}")
        self.L = _lib.lib()
        self.scalar_format = scalar_format
        self.normal_format = normal_format
        self.brick_log2 = brick_log2
        self.tile_shape = tile_shape
        self.kernel = kernel
        self.readback = readback
        self.ctx = ctypes.c_void_p()
        check(self.L.hvr_create(device, 1, 1, ctypes.byref(self.ctx)), "hvr_create")
        self._resident: Dict[int, ResidentVolume] = {}
        self._slots: List[_Slot] = []
        try:
            check(self.L.hvr_set_brick_log2(self.ctx, brick_log2), "hvr_set_brick_log2")
            for _ in range(pipeline):
                self._slots.append(_Slot(self.L, self.ctx, tile_shape, kernel, readback))
        except Exception:
            self.close()
            raise
        self._cur = 0        # slot being filled
        self._last = None    # slot of the last run
        self._reuse = False  # the slot being filled still holds submitted requests

    def set_kernel(self, kernel: str):
        """Kernel variant for the following launches (a key of KERNEL_VARIANTS)."""
        if kernel not in KERNEL_VARIANTS:
            raise ValueError(f"kernel must be one of {sorted(KERNEL_VARIANTS)}")
        w, sk, g = KERNEL_VARIANTS[kernel]
        for s in self._slots:
            check(self.L.hvr_batch_set_kernel(s.handle, w, int(sk), int(g)), "hvr_batch_set_kernel")
        self.kernel = kernel

    def set_readback(self, readback: str):
        """Where the following launches write their images ("copy" or
        "direct"; same bytes)."""
        if readback not in READBACK:
            raise ValueError(f"readback must be one of {sorted(READBACK)}")
        for s in self._slots:
            check(self.L.hvr_batch_set_readback(s.handle, READBACK[readback]), "hvr_batch_set_readback")
        self.readback = readback

    def reserve(self, max_requests: int, max_pixels: int, max_bricks: int):
        """Grow every slot's pools now to what a batch of up to
        ``max_requests`` requests, ``max_pixels`` pixels and volumes of up to
        ``max_bricks`` bricks needs, so that such batches never allocate
        while others are in flight (hvr_batch_reserve)."""
        for s in self._slots:
            check(self.L.hvr_batch_reserve(s.handle, int(max_requests), int(max_pixels), int(max_bricks)),
                  "hvr_batch_reserve")

    def set_tile_shape(self, tile_shape: str):
        """Ray tile of a wavefront for the following launches ("8x8" or "16x4")."""
        if tile_shape not in TILE_SHAPES:
            raise ValueError(f"tile_shape must be one of {sorted(TILE_SHAPES)}")
        for s in self._slots:
            check(self.L.hvr_batch_set_tile_shape(s.handle, TILE_SHAPES[tile_shape]), "hvr_batch_set_tile_shape")
        self.tile_shape = tile_shape

    # -- volumes -------------------------------------------------------------
    def upload(self, volume, device_normals: bool = False) -> ResidentVolume:
        """Make ``volume`` (a ``Volume`` or ``DeviceBlobVolume``) resident,
        once per object.  ``device_normals``: a volume without normals gets
        normals computed on the device."""
        if not isinstance(volume, (Volume, DeviceBlobVolume)):
            raise TypeError(f"Expected Volume instance, got {type(volume)}")
        rv = self._resident.get(id(volume))
        if rv is None or rv.volume is not volume:
            rv = ResidentVolume(self.L, self.ctx, volume, self.scalar_format, self.normal_format, device_normals)
            self._resident[id(volume)] = rv
        return rv

    def release(self, volume: Volume):
        """Free a resident volume (after every batch using it completed and
        was cleared)."""
        rv = self._resident.pop(id(volume), None)
        if rv is not None:
            rv.release()

    def resident_bytes(self) -> int:
        return sum(rv.nbytes() for rv in self._resident.values())

    # -- requests ------------------------------------------------------------
    def add(self, volume: Volume, camera: Camera, color_transfer_function, opacity_transfer_function,
            width: int = 512, height: int = 512, config: Optional[RenderConfig] = None,
            light: Optional[Light] = None, opacity_threshold: float = REFERENCE_TERMINATION_ALPHA,
            lut_size: Optional[int] = None) -> int:
        """Queue one request in the batch being filled; returns its index
        there.  Defaults as the renderer's (balanced preset, default light,
        termination at 0.99)."""
        if not isinstance(camera, Camera):
            raise TypeError(f"Expected Camera instance, got {type(camera)}")
        config = RenderConfig.balanced() if config is None else config
        light = Light.default() if light is None else light
        if not isinstance(config, RenderConfig):
            raise TypeError(f"Expected RenderConfig instance, got {type(config)}")
        if not isinstance(light, Light):
            raise TypeError(f"Expected Light instance, got {type(light)}")
        rv = self.upload(volume)
        lut = rgba_lut(color_transfer_function, opacity_transfer_function, lut_size)
        p = request_params(volume, camera, config, light, width, height, opacity_threshold)
        if self._reuse:
            # first request after a submit: this slot's previous batch is done with
            self.wait(self._cur)
            self._slots[self._cur].in_flight = False
            check(self.L.hvr_batch_clear(self._slots[self._cur].handle), "hvr_batch_clear")
            self._slots[self._cur].sizes = []
            self._reuse = False
        # normals from the volume's interleaved field; without normals the
        # normal sampler reads the scalar texture (unit 0) as (rho, 0, 0)
        nrole = _lib.HVR_ROLE_NORMAL if rv.has_normals else _lib.HVR_ROLE_SCALAR
        slot = self._slots[self._cur]
        idx = ctypes.c_int()
        check(self.L.hvr_batch_add(slot.handle, rv.handle, _lib.HVR_ROLE_SCALAR, rv.handle, nrole,
                                   lut.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), lut.shape[0],
                                   ctypes.byref(p), int(width), int(height), ctypes.byref(idx)),
              "hvr_batch_add")
        slot.sizes.append((int(width), int(height)))
        return idx.value

    def __len__(self):
        return 0 if self._reuse else len(self._slots[self._cur].sizes)

    def clear(self):
        """Drop the requests of the batch being filled."""
        slot = self._slots[self._cur]
        if slot.in_flight:
            self.wait(self._cur)
        check(self.L.hvr_batch_clear(slot.handle), "hvr_batch_clear")
        slot.sizes = []
        self._reuse = False

    # -- running -------------------------------------------------------------
    def submit(self, count: bool = False) -> int:
        """Launch the batch being filled without waiting; returns its slot.
        The next ``add`` starts a new batch in the next slot, after waiting
        for the batch there (whose results and requests are then gone);
        until then every submitted slot can be ``rerun``."""
        k = self._cur
        self._launch(k, count)
        self._cur = (k + 1) % len(self._slots)
        self._reuse = True
        return k

    def rerun(self, slot: int, count: bool = False):
        """Launch slot ``slot`` again with its requests (no waiting)."""
        self._launch(slot, count)

    def _launch(self, k: int, count: bool):
        s = self._slots[k]
        with trace.span("hvr.batch.launch"):
            check(self.L.hvr_batch_run_async(s.handle, int(count)), "hvr_batch_run_async")
        s.in_flight = True
        self._last = k

    def wait(self, slot: Optional[int] = None):
        """Wait for slot ``slot`` (default: the last launched)."""
        k = self._last if slot is None else slot
        if k is None:
            return
        s = self._slots[k]
        with trace.span("hvr.batch.wait"):
            check(self.L.hvr_batch_wait(s.handle), "hvr_batch_wait")
        s.in_flight = False

    def done(self, slot: int) -> bool:
        """Non-blocking: whether slot ``slot``'s images are in host memory."""
        s = self._slots[slot]
        if not s.in_flight:
            return True
        d = ctypes.c_int()
        check(self.L.hvr_batch_query(s.handle, ctypes.byref(d)), "hvr_batch_query")
        if d.value:
            s.in_flight = False
        return bool(d.value)

    def free_slot(self) -> Optional[int]:
        """The slot being filled if it is free to fill (its previous batch,
        if any, finished), else None; never blocks."""
        k = self._cur
        if self._reuse and not self.done(k):
            return None
        return k

    def run(self, count: bool = False):
        """Render every request of the batch being filled (and, with
        ``count``, per-pixel sample counts); returns when the images are in
        host memory.  The requests stay, so ``run`` can be called again."""
        k = self._cur
        self._launch(k, count)
        self.wait(k)

    def _slot(self, slot):
        k = self._last if slot is None else slot
        if k is None:
            raise RuntimeError("no batch has been run")
        return k, self._slots[k]

    def view(self, index: int, slot: int) -> np.ndarray:
        """Request ``index``'s RGBA8 pixels in slot ``slot`` as a read-only
        array over pinned host memory (valid until the slot is reused);
        the slot must have completed."""
        s = self._slots[slot]
        w, h = s.sizes[index]
        ptr = ctypes.POINTER(ctypes.c_uint8)()
        check(self.L.hvr_batch_output(s.handle, index, ctypes.byref(ptr)), "hvr_batch_output")
        a = np.ctypeslib.as_array(ptr, shape=(h * w * 4,))
        a.flags.writeable = False
        return a

    def result(self, index: int, slot: Optional[int] = None) -> bytes:
        """Request ``index``'s RGBA8 bytes, bottom row first, from slot
        ``slot`` (default: the last launched; waits for it)."""
        k, s = self._slot(slot)
        w, h = s.sizes[index]
        if s.in_flight:
            self.wait(k)
        ptr = ctypes.POINTER(ctypes.c_uint8)()
        check(self.L.hvr_batch_output(s.handle, index, ctypes.byref(ptr)), "hvr_batch_output")
        with trace.span("hvr.batch.readback"):
            return ctypes.string_at(ptr, w * h * 4)

    def results(self, slot: Optional[int] = None) -> List[bytes]:
        k, s = self._slot(slot)
        return [self.result(i, k) for i in range(len(s.sizes))]

    def counts(self, index: int, slot: Optional[int] = None) -> np.ndarray:
        """Request ``index``'s composited-sample counts (H, W), from a run
        with ``count=True``."""
        k, s = self._slot(slot)
        w, h = s.sizes[index]
        if s.in_flight:
            self.wait(k)
        ptr = ctypes.POINTER(ctypes.c_uint32)()
        check(self.L.hvr_batch_counts(s.handle, index, ctypes.byref(ptr)), "hvr_batch_counts")
        return np.ctypeslib.as_array(ptr, shape=(h * w,)).reshape(h, w).copy()

    def last_ms(self, slot: Optional[int] = None) -> float:
        """Device time of a slot's last launch (ms)."""
        k, s = self._slot(slot)
        ms = ctypes.c_float()
        check(self.L.hvr_batch_last_ms(s.handle, ctypes.byref(ms)), "hvr_batch_last_ms")
        return float(ms.value)

    def close(self):
        for s in getattr(self, "_slots", []):
            s.destroy()
        self._slots = []
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
