"""Ray-marching quality settings (``RenderConfig``) and its presets.

Same fields, defaults, presets and validation errors as the reference's
``pyvr/config.py:11-374``.  Note on semantics kept from the reference: only
``step_size``, ``max_steps`` and ``reference_step_size`` reach the ray marcher
(``pyvr/moderngl_renderer/renderer.py:303-309``).  ``early_ray_termination``
and ``opacity_threshold`` are stored and validated but the reference shader
terminates at a fixed accumulated alpha of 0.99
(``pyvr/shaders/volume.frag.glsl:87``); the HIP renderer does the same unless
the caller opts in with ``VolumeRenderer.set_opacity_threshold``.
"""

import warnings
from dataclasses import dataclass

# Accumulated-alpha level at which the reference shader stops a ray
# (volume.frag.glsl:87).  It is a constant of the shader, not a config field.
REFERENCE_TERMINATION_ALPHA = 0.99


@dataclass
class RenderConfig:
    """Quality/performance trade-off for the ray marcher.

    Attributes:
        step_size: world-space distance between two samples on a ray.
        max_steps: upper bound on loop iterations per ray (counted from the
            ray's entry point, in or out of the box).
        early_ray_termination: stored for API compatibility (see module doc).
        opacity_threshold: stored for API compatibility (see module doc).
        reference_step_size: step length at which the opacity transfer
            function is defined; per-sample alpha is corrected with
            ``1 - exp(-alpha * step_size / reference_step_size)``.
    """

    step_size: float = 0.01
    max_steps: int = 500
    early_ray_termination: bool = True
    opacity_threshold: float = 0.95
    reference_step_size: float = 0.01

    def __post_init__(self):
        self.validate()

    def validate(self) -> None:
        """Raise ``ValueError`` on invalid settings (reference config.py:64-94)."""
        if self.step_size <= 0:
            raise ValueError("step_size must be positive")
        if self.step_size > 1.0:
            warnings.warn(
                f"step_size {self.step_size} is very large and may produce poor quality",
                UserWarning,
            )
        if self.max_steps < 1:
            raise ValueError("max_steps must be at least 1")
        if self.max_steps > 10000:
            warnings.warn(
                f"max_steps {self.max_steps} is very large and may be slow",
                UserWarning,
            )
        if not (0.0 <= self.opacity_threshold <= 1.0):
            raise ValueError("opacity_threshold must be between 0.0 and 1.0")

    # Presets: literal values of the reference (config.py:96-227).
    @classmethod
    def fast(cls) -> "RenderConfig":
        return cls(step_size=0.02, max_steps=100, early_ray_termination=True,
                   opacity_threshold=0.90)

    @classmethod
    def balanced(cls) -> "RenderConfig":
        return cls(step_size=0.01, max_steps=500, early_ray_termination=True,
                   opacity_threshold=0.95)

    @classmethod
    def high_quality(cls) -> "RenderConfig":
        return cls(step_size=0.005, max_steps=1000, early_ray_termination=True,
                   opacity_threshold=0.98)

    @classmethod
    def ultra_quality(cls) -> "RenderConfig":
        return cls(step_size=0.001, max_steps=2000, early_ray_termination=False,
                   opacity_threshold=1.0)

    @classmethod
    def preview(cls) -> "RenderConfig":
        return cls(step_size=0.05, max_steps=50, early_ray_termination=True,
                   opacity_threshold=0.80)

    @classmethod
    def custom(cls, step_size: float, max_steps: int,
               early_ray_termination: bool = True,
               opacity_threshold: float = 0.95) -> "RenderConfig":
        return cls(step_size=step_size, max_steps=max_steps,
                   early_ray_termination=early_ray_termination,
                   opacity_threshold=opacity_threshold)

    def copy(self) -> "RenderConfig":
        return RenderConfig(
            step_size=self.step_size,
            max_steps=self.max_steps,
            early_ray_termination=self.early_ray_termination,
            opacity_threshold=self.opacity_threshold,
            reference_step_size=self.reference_step_size,
        )

    def with_step_size(self, step_size: float) -> "RenderConfig":
        cfg = self.copy()
        cfg.step_size = step_size
        cfg.validate()
        return cfg

    def with_max_steps(self, max_steps: int) -> "RenderConfig":
        cfg = self.copy()
        cfg.max_steps = max_steps
        cfg.validate()
        return cfg

    def estimate_samples_per_ray(self) -> int:
        """Samples along a unit-cube diagonal (1.732), capped at ``max_steps``."""
        return min(int(1.732 / self.step_size), self.max_steps)

    def estimate_render_time_relative(self) -> float:
        """Sample-count ratio against the balanced preset."""
        return self.estimate_samples_per_ray() / RenderConfig.balanced().estimate_samples_per_ray()

    def __repr__(self) -> str:
        return (
            f"RenderConfig(step_size={self.step_size}, "
            f"max_steps={self.max_steps}, "
            f"early_termination={self.early_ray_termination}, "
            f"~{self.estimate_samples_per_ray()} samples/ray)"
        )


class RenderConfigError(Exception):
    """Raised for rendering-configuration errors."""
