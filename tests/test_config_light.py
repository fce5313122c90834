"""RenderConfig and Light: values and errors of the reference's own tests
(tests/test_config.py:23-217, tests/test_config_opacity_correction.py,
tests/test_lighting/test_light.py:21-30, 167-175 of the reference)."""

import warnings

import numpy as np
import pytest

from pyvr_amd.config import REFERENCE_TERMINATION_ALPHA, RenderConfig, RenderConfigError
from pyvr_amd.lighting import Light, LightError


class TestRenderConfig:
    def test_presets_literal_values(self):
        # config.py:96-227 of the reference
        assert (RenderConfig.preview().step_size, RenderConfig.preview().max_steps) == (0.05, 50)
        assert (RenderConfig.fast().step_size, RenderConfig.fast().max_steps) == (0.02, 100)
        assert (RenderConfig.balanced().step_size, RenderConfig.balanced().max_steps) == (0.01, 500)
        hq = RenderConfig.high_quality()
        assert (hq.step_size, hq.max_steps, hq.opacity_threshold) == (0.005, 1000, 0.98)
        uq = RenderConfig.ultra_quality()
        assert (uq.step_size, uq.max_steps, uq.early_ray_termination, uq.opacity_threshold) == (0.001, 2000, False, 1.0)
        for cfg in (RenderConfig.preview(), RenderConfig.fast(), RenderConfig.balanced(), hq, uq):
            assert cfg.reference_step_size == 0.01

    def test_defaults(self):
        c = RenderConfig()
        assert (c.step_size, c.max_steps, c.early_ray_termination, c.opacity_threshold,
                c.reference_step_size) == (0.01, 500, True, 0.95, 0.01)

    def test_preset_ordering(self):
        steps = [RenderConfig.preview(), RenderConfig.fast(), RenderConfig.balanced(),
                 RenderConfig.high_quality(), RenderConfig.ultra_quality()]
        assert all(a.step_size > b.step_size for a, b in zip(steps, steps[1:]))

    @pytest.mark.parametrize("kw,match", [
        ({"step_size": 0.0}, "step_size"), ({"step_size": -0.01}, "step_size"),
        ({"max_steps": 0}, "max_steps"), ({"max_steps": -10}, "max_steps"),
        ({"opacity_threshold": 1.5}, "opacity_threshold"), ({"opacity_threshold": -0.1}, "opacity_threshold"),
    ])
    def test_validation(self, kw, match):
        with pytest.raises(ValueError, match=match):
            RenderConfig(**kw)

    def test_warnings(self):
        with pytest.warns(UserWarning):
            RenderConfig(step_size=2.0)
        with pytest.warns(UserWarning):
            RenderConfig(max_steps=20000)

    def test_copy_and_with(self):
        c = RenderConfig(step_size=0.015, max_steps=300, early_ray_termination=False,
                         opacity_threshold=0.85, reference_step_size=0.008)
        d = c.copy()
        assert d == c and d is not c
        assert c.with_step_size(0.005).step_size == 0.005
        assert c.with_step_size(0.005).reference_step_size == 0.008
        assert c.with_max_steps(600).max_steps == 600
        with pytest.raises(ValueError):
            c.with_max_steps(0)

    def test_estimates(self):
        # reference config.py:319-358: diagonal 1.732 / step, capped
        assert RenderConfig.balanced().estimate_samples_per_ray() == 173
        assert RenderConfig.fast().estimate_samples_per_ray() == 86
        assert RenderConfig.preview().estimate_samples_per_ray() == 34
        assert RenderConfig.high_quality().estimate_render_time_relative() == pytest.approx(346 / 173)
        assert "samples/ray" in repr(RenderConfig())

    def test_custom(self):
        c = RenderConfig.custom(step_size=0.015, max_steps=300, early_ray_termination=False)
        assert (c.step_size, c.max_steps, c.early_ray_termination) == (0.015, 300, False)

    def test_reference_termination_constant(self):
        assert REFERENCE_TERMINATION_ALPHA == 0.99
        assert issubclass(RenderConfigError, Exception)


class TestOpacityCorrection:
    """The correction the march applies per sample: 1 - exp(-alpha*step/ref)."""

    @staticmethod
    def corr(a, s, r):
        return 1.0 - np.exp(-a * s / r)

    def test_two_half_steps_equal_one_full_step(self):
        # exact Beer-Lambert identity: two steps of s composite to one of 2s
        for a in (0.1, 0.4, 0.9):
            one = self.corr(a, 0.02, 0.01)
            h = self.corr(a, 0.01, 0.01)
            assert h + (1 - h) * h == pytest.approx(one, rel=1e-12)

    def test_bounds_and_monotone(self):
        steps = [0.001, 0.005, 0.01, 0.02, 0.05]
        vals = [self.corr(0.4, s, 0.01) for s in steps]
        assert all(b > a for a, b in zip(vals, vals[1:]))
        assert self.corr(0.0, 0.005, 0.01) == 0.0
        assert 0.99 < self.corr(0.5, 1.0, 0.01) <= 1.0


class TestLight:
    def test_default(self):
        L = Light.default()
        assert (L.ambient_intensity, L.diffuse_intensity) == (0.2, 0.8)
        assert np.array_equal(L.position, [1, 1, 1]) and np.array_equal(L.target, [0, 0, 0])
        assert L.position.dtype == np.float32

    def test_directional_closed_form(self):
        L = Light.directional(direction=[1, -1, 0], ambient=0.2, diffuse=0.8)
        d = np.array([1, -1, 0], np.float32) / np.float32(np.sqrt(2))
        assert np.allclose(L.position, -d * 10.0, atol=1e-6)
        assert np.allclose(L.get_direction(), d, atol=1e-6)
        L2 = Light.directional(direction=np.array([1, 0, 0]), distance=50.0)
        assert np.allclose(L2.position, [-50, 0, 0])

    def test_directional_normalises(self):
        L = Light.directional(direction=np.array([3, 4, 0]))
        assert np.linalg.norm(L.get_direction()) == pytest.approx(1.0)

    def test_point_and_ambient(self):
        L = Light.point_light(position=np.array([5, 5, 5]))
        assert np.allclose(L.position, [5, 5, 5]) and np.allclose(L.target, 0)
        A = Light.ambient_only(intensity=0.5)
        assert (A.ambient_intensity, A.diffuse_intensity) == (0.5, 0.0)
        assert np.allclose(A.get_direction(), [0, 0, -1])

    @pytest.mark.parametrize("kw", [{"ambient_intensity": -0.1}, {"ambient_intensity": 1.1},
                                    {"diffuse_intensity": -0.1}, {"diffuse_intensity": 1.5},
                                    {"position": np.zeros(2)}, {"target": [0, 0, 0]}])
    def test_validation(self, kw):
        with pytest.raises(ValueError):
            Light(**kw)

    def test_copy_and_linking(self):
        from pyvr_amd.camera import Camera

        L = Light.camera_linked(ambient=0.3)
        assert L.is_linked and L.get_offsets() == {"azimuth": 0.0, "elevation": 0.0, "distance": 0.0}
        cam = Camera.isometric_view(distance=4.0)
        L.update_from_camera(cam)
        pos, _ = cam.get_camera_vectors()
        assert np.allclose(L.position, pos)
        c = L.copy()
        assert c.is_linked and c.ambient_intensity == 0.3
        L.unlink_from_camera()
        with pytest.raises(ValueError):
            L.update_from_camera(cam)
        assert issubclass(LightError, Exception)
