"""``VolumeRenderer`` on MI355X: the drop-in counterpart of the reference's
``pyvr.moderngl_renderer.VolumeRenderer`` (``pyvr/moderngl_renderer/
renderer.py:19-320``).

Same constructor, methods, defaults, errors and state semantics; rendering
runs in hand-written HIP kernels (``libhipvr.so``) instead of an OpenGL
fragment shader.  Additions, all keyword-only or new methods with
reference-identical defaults:

* ``device`` -- HIP device index;
* ``scalar_format`` / ``normal_format`` -- device storage of the volume
  ("f32"/"f16"/"u8" and "f32x3"/"sn8x3", or "f16x3" with "f16"; default
  float32 like the reference);
* ``set_opacity_threshold`` -- early-termination alpha (default 0.99, the
  reference shader's constant);
* ``sample_counts`` / ``last_kernel_ms`` -- measurement helpers.
"""

import os
from typing import Optional

import numpy as np

from . import trace
from ..camera import Camera
from ..config import REFERENCE_TERMINATION_ALPHA, RenderConfig
from ..lighting import Light
from ..transferfunctions import ColorTransferFunction, OpacityTransferFunction
from ..volume import Volume
from .manager import HipManager


class HipVolumeRenderer:
    """Ray-marching direct volume renderer on an AMD Instinct GPU.

    Example:
        >>> renderer = VolumeRenderer(width=512, height=512, config=RenderConfig.balanced())
        >>> renderer.load_volume(volume)
        >>> renderer.set_camera(camera)
        >>> renderer.set_transfer_functions(ctf, otf)
        >>> rgba = renderer.render()   # width*height*4 bytes, bottom row first
    """

    def __init__(self, width=512, height=512, config=None, light=None, *, device: int = 0,
                 scalar_format: str = "f32", normal_format: str = "f32x3"):
        self.width = width
        self.height = height
        self.volume: Optional[Volume] = None
        self.camera: Optional[Camera] = None
        if config is None:
            self.config = RenderConfig.balanced()
        else:
            if not isinstance(config, RenderConfig):
                raise TypeError(f"Expected RenderConfig instance, got {type(config)}")
            self.config = config
        if light is None:
            self.light = Light.default()
        else:
            if not isinstance(light, Light):
                raise TypeError(f"Expected Light instance, got {type(light)}")
            self.light = light
        self.scalar_format = scalar_format
        self.normal_format = normal_format

        self.gl_manager = HipManager(width, height, device=device)
        self.gl_manager.load_shaders()
        self._update_render_config()
        self.gl_manager.set_uniform_vector("volume_min_bounds", (-0.5, -0.5, -0.5))
        self.gl_manager.set_uniform_vector("volume_max_bounds", (0.5, 0.5, 0.5))
        self._update_light()

    def load_volume(self, volume: Volume, *, scalar_format: Optional[str] = None,
                    normal_format: Optional[str] = None, device_normals: bool = False) -> None:
        """Upload ``volume`` (data, normals if present) and its bounds.

        Raises ``TypeError`` if ``volume`` is not a ``Volume``.  Without
        normals the normal sampler keeps its previous unit (by default unit
        0), as in the reference.  ``device_normals=True`` on a volume without
        normals computes them on the GPU instead -- the same bytes as
        ``volume.compute_normals()`` followed by ``load_volume``, without the
        host gradient."""
        if not isinstance(volume, Volume):
            raise TypeError(
                f"Expected Volume instance, got {type(volume)}. "
                "Create a Volume instance: from pyvr_amd.volume import Volume; "
                "volume = Volume(data=your_array)"
            )
        self.volume = volume
        sf = scalar_format or self.scalar_format
        nf = normal_format or self.normal_format
        on_device = device_normals and not volume.has_normals
        unit = self.gl_manager.create_volume_texture(
            volume.data, scalar_format=sf, normal_format=nf if (volume.has_normals or on_device) else None,
            device_normals=on_device)
        self.gl_manager.set_uniform_int("volume_texture", unit)
        self.gl_manager.set_uniform_vector("volume_min_bounds", tuple(volume.min_bounds))
        self.gl_manager.set_uniform_vector("volume_max_bounds", tuple(volume.max_bounds))
        if volume.has_normals:
            normal_unit = self.gl_manager.create_normal_texture(volume.normals, normal_format=nf)
            self.gl_manager.set_uniform_int("normal_volume", normal_unit)
        elif on_device:
            self.gl_manager.set_uniform_int("normal_volume", self.gl_manager.create_device_normal_texture())

    def set_camera(self, camera: Camera) -> None:
        """View/projection matrices and camera position from ``camera``
        (aspect = width/height at call time)."""
        if not isinstance(camera, Camera):
            raise TypeError(f"Expected Camera instance, got {type(camera)}")
        self.camera = camera
        aspect = self.width / self.height
        # the camera vectors once: get_view_matrix() would compute them again
        position, up = camera.get_camera_vectors()
        view_matrix = camera._view_matrix(position, up)
        projection_matrix = camera.get_projection_matrix(aspect)
        self.gl_manager.set_uniform_matrix("view_matrix", view_matrix)
        self.gl_manager.set_uniform_matrix("projection_matrix", projection_matrix)
        self.gl_manager.set_uniform_vector("camera_pos", tuple(position))

    def set_light(self, light: Light) -> None:
        if not isinstance(light, Light):
            raise TypeError(f"Expected Light instance, got {type(light)}")
        self.light = light
        self._update_light()

    def set_transfer_functions(self, color_transfer_function: ColorTransferFunction,
                               opacity_transfer_function: OpacityTransferFunction,
                               size: Optional[int] = None) -> None:
        """Combined RGBA LUT of ``size`` rows (default: the larger lut_size)."""
        unit = self.gl_manager.create_rgba_transfer_function_texture(
            color_transfer_function, opacity_transfer_function, size)
        self.gl_manager.set_uniform_int("transfer_function_lut", unit)

    def render(self) -> bytes:
        """Render one frame: width*height*4 bytes, RGBA8, bottom row first."""
        with trace.span("hvr.render"):
            self.gl_manager.clear_framebuffer(0.0, 0.0, 0.0, 0.0)
            self.gl_manager.setup_blending()
            self.gl_manager.render_quad()
            return self.gl_manager.read_pixels()

    def render_to_pil(self, data=None):
        """Render (or take ``data``) and return a top-row-first PIL image."""
        from PIL import Image

        if data is None:
            data = self.render()
        image = Image.frombytes("RGBA", (self.width, self.height), data)
        return image.transpose(Image.FLIP_TOP_BOTTOM)

    def set_config(self, config):
        if not isinstance(config, RenderConfig):
            raise TypeError(f"Expected RenderConfig instance, got {type(config)}")
        self.config = config
        self._update_render_config()

    def get_config(self):
        return self.config

    def get_light(self):
        return self.light

    def get_volume(self) -> Optional[Volume]:
        return self.volume

    def get_camera(self) -> Optional[Camera]:
        return self.camera

    # -- additions -------------------------------------------------------------
    def set_opacity_threshold(self, threshold: Optional[float] = None) -> None:
        """Accumulated alpha at which a ray stops.  ``None`` restores the
        reference shader's constant 0.99; a value above 1 disables early
        termination."""
        tau = REFERENCE_TERMINATION_ALPHA if threshold is None else float(threshold)
        if not tau > 0.0:
            raise ValueError("opacity threshold must be positive")
        self.gl_manager.set_uniform_float("opacity_threshold", tau)

    def sample_counts(self) -> np.ndarray:
        """Ray samples composited per pixel for the current state, (H, W)
        uint32, bottom row first."""
        return self.gl_manager.sample_counts()

    def last_kernel_ms(self):
        """(march_ms, setup_ms) of the last rendered frame."""
        return self.gl_manager.last_kernel_ms()

    def close(self) -> None:
        """Release the device resources now (also done on garbage collection)."""
        self.gl_manager.cleanup()

    # -- uniform plumbing (reference renderer.py:303-316) ---------------------
    def _update_render_config(self):
        self.gl_manager.set_uniform_float("step_size", self.config.step_size)
        self.gl_manager.set_uniform_int("max_steps", self.config.max_steps)
        self.gl_manager.set_uniform_float("reference_step_size", self.config.reference_step_size)

    def _update_light(self):
        self.gl_manager.set_uniform_float("ambient_light", self.light.ambient_intensity)
        self.gl_manager.set_uniform_float("diffuse_light", self.light.diffuse_intensity)
        self.gl_manager.set_uniform_vector("light_position", tuple(self.light.position))
        self.gl_manager.set_uniform_vector("light_target", tuple(self.light.target))


VolumeRenderer = HipVolumeRenderer
