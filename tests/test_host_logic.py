"""Host logic of VolumeRenderer / HipManager on the CPU, with the native layer
replaced by tests/fake_hvr.py.  Mirrors the reference's mocked renderer tests
(tests/test_moderngl_renderer/test_volume_renderer.py and
test_moderngl_manager.py): API surface, defaults, errors, texture-unit
allocation and the GL sampler semantics the manager models."""

import numpy as np
import pytest

from fake_hvr import FakeHvr
from pyvr_amd.camera import Camera
from pyvr_amd.config import RenderConfig
from pyvr_amd.hip_renderer import HipManager, VolumeRenderer, _lib
from pyvr_amd.lighting import Light
from pyvr_amd.transferfunctions import ColorTransferFunction, OpacityTransferFunction
from pyvr_amd.volume import Volume


@pytest.fixture
def fake(monkeypatch):
    f = FakeHvr()
    monkeypatch.setattr(_lib, "_lib", f)
    return f


def vol(shape=(8, 8, 8), normals=True, seed=0):
    rng = np.random.default_rng(seed)
    d = rng.random(shape).astype(np.float32)
    n = rng.random(shape + (3,)).astype(np.float32) if normals else None
    return Volume(data=d, normals=n)


def tfs():
    return ColorTransferFunction.from_colormap("viridis"), OpacityTransferFunction.linear(0.0, 0.5)


class TestRendererSurface:
    def test_defaults(self, fake):
        r = VolumeRenderer()
        assert (r.width, r.height) == (512, 512)
        assert r.config.step_size == 0.01 and r.config.max_steps == 500
        assert r.light.ambient_intensity == 0.2 and r.gl_manager is not None
        assert r.volume is None and r.camera is None
        v = r.gl_manager.program.values
        assert np.allclose(v["volume_min_bounds"], -0.5) and np.allclose(v["volume_max_bounds"], 0.5)
        assert v["opacity_threshold"] == pytest.approx(0.99)

    def test_custom(self, fake):
        r = VolumeRenderer(width=1024, height=768, config=RenderConfig(step_size=0.005, max_steps=500))
        assert (r.width, r.height, r.config.step_size) == (1024, 768, 0.005)
        assert fake.contexts[r.gl_manager.ctx.value] == (1024, 768)

    @pytest.mark.parametrize("call", [
        lambda r: r.load_volume(np.zeros((4, 4, 4))),
        lambda r: r.set_camera("camera"),
        lambda r: r.set_light({}),
        lambda r: r.set_config(object()),
    ])
    def test_type_errors(self, fake, call):
        r = VolumeRenderer()
        with pytest.raises(TypeError):
            call(r)

    def test_ctor_type_errors(self, fake):
        with pytest.raises(TypeError, match="RenderConfig"):
            VolumeRenderer(config="fast")
        with pytest.raises(TypeError, match="Light"):
            VolumeRenderer(light="sun")

    def test_getters_and_config_uniforms(self, fake):
        r = VolumeRenderer()
        cfg = RenderConfig(step_size=0.01, max_steps=500, reference_step_size=0.008)
        r.set_config(cfg)
        assert r.get_config() is cfg
        v = r.gl_manager.program.values
        assert v["reference_step_size"] == pytest.approx(0.008) and v["max_steps"] == 500
        L = Light.directional([1, -1, 0])
        r.set_light(L)
        assert r.get_light() is L and v["diffuse_light"] == pytest.approx(0.8)
        vv = vol()
        r.load_volume(vv)
        assert r.get_volume() is vv
        cam = Camera.front_view(distance=3.0)
        r.set_camera(cam)
        assert r.get_camera() is cam

    def test_render_bytes_and_pil_flip(self, fake):
        r = VolumeRenderer(width=6, height=4)
        r.load_volume(vol())
        r.set_camera(Camera.isometric_view())
        r.set_transfer_functions(*tfs())
        data = r.render()
        assert isinstance(data, bytes) and len(data) == 6 * 4 * 4
        img = np.asarray(r.render_to_pil(data))
        assert img.shape == (4, 6, 4)
        assert img[0, 0, 0] == 3 and img[-1, 0, 0] == 0  # top row first after the flip
        p = fake.params[-1]
        assert p["camera_set"] == 1 and p["blend"] == 1 and p["clear_rgba"] == [0, 0, 0, 0]
        assert p["et_threshold"] == pytest.approx(0.99)

    def test_render_without_camera_marks_matrices_unset(self, fake):
        r = VolumeRenderer(width=8, height=8)
        r.render()
        assert fake.params[-1]["camera_set"] == 0

    def test_opacity_threshold(self, fake):
        r = VolumeRenderer(width=8, height=8)
        r.set_opacity_threshold(0.95)
        r.render()
        assert fake.params[-1]["et_threshold"] == pytest.approx(0.95)
        r.set_opacity_threshold()
        r.render()
        assert fake.params[-1]["et_threshold"] == pytest.approx(0.99)
        with pytest.raises(ValueError):
            r.set_opacity_threshold(0.0)

    def test_transfer_function_size(self, fake):
        r = VolumeRenderer()
        ctf, otf = tfs()
        otf.lut_size = 512
        r.set_transfer_functions(ctf, otf)
        r.render()
        assert fake.luts[-1] == 512
        r.set_transfer_functions(ctf, otf, size=100)
        r.render()
        assert fake.luts[-1] == 100

    def test_multi_frame_and_replacement(self, fake):
        r = VolumeRenderer(width=16, height=16)
        r.set_transfer_functions(*tfs())
        r.load_volume(vol((16, 16, 16), normals=False))
        r.set_config(RenderConfig.balanced())
        a = r.render()
        r.set_config(RenderConfig.high_quality())
        b = r.render()
        assert len(a) == len(b) == 16 * 16 * 4
        r.load_volume(vol((32, 32, 32), normals=False, seed=1))
        r.render()
        assert fake.frames == 3

    def test_tiny_volume(self, fake):
        r = VolumeRenderer(width=4, height=4)
        r.load_volume(Volume(data=np.random.random((1, 1, 1)).astype(np.float32)))
        r.render()


class TestReferenceRendererBehaviour:
    """The reference's renderer tests that the classes above do not already
    cover (tests/test_moderngl_renderer/test_volume_renderer.py:423-790):
    the TypeError texts, uniforms written by the light, volume and config
    setters, and the presets."""

    @pytest.mark.parametrize("call,kind", [
        (lambda r: r.load_volume(np.random.random((10, 10, 10))), "Volume"),
        (lambda r: r.set_light("not a light"), "Light"),
        (lambda r: r.set_light({"ambient": 0.5}), "Light"),
        (lambda r: r.set_config("not a config"), "RenderConfig"),
        (lambda r: r.set_config({"step_size": 0.01}), "RenderConfig"),
    ])
    def test_type_error_messages(self, fake, call, kind):
        r = VolumeRenderer(width=16, height=16)
        with pytest.raises(TypeError, match=f"Expected {kind} instance"):
            call(r)
        with pytest.raises(TypeError, match="Expected RenderConfig instance"):
            VolumeRenderer(config="invalid")

    def test_light_uniforms(self, fake):
        r = VolumeRenderer(width=16, height=16)
        v = r.gl_manager.program.values
        assert v["ambient_light"] == pytest.approx(0.2) and v["diffuse_light"] == pytest.approx(0.8)
        L = Light.directional(direction=[1, -1, 0], ambient=0.3, diffuse=0.9)
        r.set_light(L)
        assert r.light is L
        assert v["ambient_light"] == pytest.approx(0.3) and v["diffuse_light"] == pytest.approx(0.9)
        assert np.allclose(v["light_position"], L.position) and np.allclose(v["light_target"], L.target)
        P = Light.point_light(position=[5, 5, 5])
        r.set_light(P)
        assert np.allclose(r.light.position, [5, 5, 5]) and np.allclose(v["light_position"], [5, 5, 5])
        r.set_light(Light.ambient_only(intensity=0.4))
        assert v["ambient_light"] == pytest.approx(0.4) and v["diffuse_light"] == 0.0
        custom = Light.directional(direction=[1, 0, 0], ambient=0.5)
        r2 = VolumeRenderer(width=16, height=16, light=custom)
        assert r2.light is custom and r2.gl_manager.program.values["ambient_light"] == pytest.approx(0.5)

    def test_volume_uniforms_and_replacement(self, fake):
        r = VolumeRenderer(width=16, height=16)
        assert r.get_volume() is None
        v = r.gl_manager.program.values
        a = Volume(data=np.random.rand(32, 32, 32).astype(np.float32), name="volume1",
                   min_bounds=np.array([-1.0, -2.0, -3.0], np.float32),
                   max_bounds=np.array([1.0, 2.0, 3.0], np.float32))
        r.load_volume(a)
        assert r.get_volume() is a and v["volume_texture"] == 0
        assert np.allclose(v["volume_min_bounds"], (-1, -2, -3)) and np.allclose(v["volume_max_bounds"], (1, 2, 3))
        b = Volume(data=np.random.rand(64, 64, 64).astype(np.float32), name="volume2")
        r.load_volume(b)
        assert r.get_volume() is b and v["volume_texture"] == 1
        assert np.allclose(v["volume_min_bounds"], -0.5) and np.allclose(v["volume_max_bounds"], 0.5)
        n = Volume(data=np.random.rand(32, 32, 32).astype(np.float32),
                   normals=np.random.rand(32, 32, 32, 3).astype(np.float32))
        r.load_volume(n)
        assert v["volume_texture"] == 2 and v["normal_volume"] == 3

    def test_config_presets_and_changes(self, fake):
        for preset in (RenderConfig.preview(), RenderConfig.fast(), RenderConfig.balanced(),
                       RenderConfig.high_quality(), RenderConfig.ultra_quality()):
            r = VolumeRenderer(width=16, height=16, config=preset)
            assert r.config is preset and r.get_config() is preset
            v = r.gl_manager.program.values
            assert v["step_size"] == pytest.approx(preset.step_size) and v["max_steps"] == preset.max_steps
        r = VolumeRenderer(width=16, height=16)
        v = r.gl_manager.program.values
        for cfg in (RenderConfig.fast(), RenderConfig.high_quality(), RenderConfig.ultra_quality()):
            r.set_config(cfg)
            assert r.config is cfg
            assert v["step_size"] == pytest.approx(cfg.step_size) and v["max_steps"] == cfg.max_steps
            assert v["reference_step_size"] == pytest.approx(cfg.reference_step_size)
        assert r.config.step_size == 0.001 and r.config.max_steps == 2000
        m = RenderConfig.balanced().with_step_size(0.008)
        r = VolumeRenderer(width=16, height=16, config=m)
        assert r.config.step_size == 0.008 and r.config.max_steps == 500
        r = VolumeRenderer(width=16, height=16, config=RenderConfig(step_size=0.015, max_steps=300))
        assert r.gl_manager.program.values["max_steps"] == 300
        r.render()
        assert fake.params[-1]["step"] == pytest.approx(0.015) and fake.params[-1]["max_steps"] == 300


class TestManager:
    def test_program_required(self, fake):
        m = HipManager()
        with pytest.raises(RuntimeError, match="Shader program not loaded"):
            m.set_uniform_float("step_size", 1.0)
        with pytest.raises(RuntimeError, match="Vertex array object not created"):
            m.render_quad()
        m.load_shaders("ignored.vert", "ignored.frag")
        with pytest.raises(KeyError):
            m.set_uniform_float("nonexistent_uniform", 1.0)

    def test_texture_units_increment(self, fake):
        m = HipManager()
        m.load_shaders()
        ctf, otf = tfs()
        u1 = m.create_volume_texture(np.random.rand(16, 16, 16).astype(np.float32))
        u2 = m.create_rgba_transfer_function_texture(ctf, otf)
        u3 = m.create_normal_texture(np.random.rand(16, 16, 16, 3).astype(np.float32))
        assert [u1, u2, u3] == [0, 1, 2]

    def test_input_validation(self, fake):
        m = HipManager()
        m.load_shaders()
        with pytest.raises(ValueError, match="Volume data must be 3D"):
            m.create_volume_texture(np.array([1, 2, 3]))
        with pytest.raises(ValueError, match="Normal volume must have 3 channels"):
            m.create_normal_texture(np.random.random((10, 10, 10, 2)).astype(np.float32))
        with pytest.raises(ValueError):
            m.create_normal_texture(np.random.random((2, 10, 10, 10, 3)).astype(np.float32))
        with pytest.raises(ValueError):
            m.create_volume_texture(np.zeros((2, 2, 2), np.float32), scalar_format="f64")
        # any dtype converts to float32
        m.create_volume_texture(np.random.randint(0, 255, (10, 10, 10), dtype=np.uint8))

    def test_cleanup_twice(self, fake):
        m = HipManager()
        m.create_volume_texture(np.random.rand(4, 4, 4).astype(np.float32))
        m.cleanup()
        m.cleanup()
        assert fake.volumes == {}


class TestSamplerSemantics:
    def test_interleaved_normals(self, fake):
        r = VolumeRenderer(width=8, height=8)
        r.load_volume(vol())
        r.render()
        s, sr, n, nr = fake.binds[-1]
        assert s == n and (sr, nr) == (_lib.HVR_ROLE_SCALAR, _lib.HVR_ROLE_NORMAL)
        v = fake.volumes[s]
        assert v["sf"] == _lib.HVR_SCALAR["f32"] and v["nf"] == _lib.HVR_NORMAL["f32x3"]
        assert v["scalar"] and v["normals"]

    def test_no_normals_alias_unit0(self, fake):
        """Without normals the normal sampler keeps its default unit 0 -- the
        first volume -- and reads (rho, 0, 0) of it."""
        r = VolumeRenderer(width=8, height=8)
        r.load_volume(vol(normals=False))
        r.render()
        s, sr, n, nr = fake.binds[-1]
        assert s == n and sr == nr == _lib.HVR_ROLE_SCALAR

    def test_lut_first_leaves_normal_sampler_unbound(self, fake):
        r = VolumeRenderer(width=8, height=8)
        r.set_transfer_functions(*tfs())           # unit 0 is the LUT
        r.load_volume(vol(normals=False))          # unit 1
        r.render()
        s, sr, n, nr = fake.binds[-1]
        assert s is not None and n is None

    def test_stale_normals_survive_reload_without_normals(self, fake):
        r = VolumeRenderer(width=8, height=8)
        r.load_volume(vol())                      # units 0 (scalar) and 1 (normals): one device volume
        first = fake.binds and None
        r.render()
        a, _, _, _ = fake.binds[-1]
        r.load_volume(vol(normals=False, seed=3))  # unit 2; normal sampler still unit 1
        r.render()
        s, sr, n, nr = fake.binds[-1]
        assert s != a and n == a and nr == _lib.HVR_ROLE_NORMAL
        assert a in fake.volumes                   # kept alive: still sampled
        assert first is None

    def test_unreferenced_volumes_are_released(self, fake):
        r = VolumeRenderer(width=8, height=8)
        r.load_volume(vol())
        r.render()
        a = fake.binds[-1][0]
        r.load_volume(vol(seed=5))
        r.render()
        assert a in fake.destroyed and a not in fake.volumes
        assert len(fake.volumes) == 1

    def test_formats_passed_through(self, fake):
        r = VolumeRenderer(width=8, height=8, scalar_format="f16", normal_format="sn8x3")
        r.load_volume(vol())
        r.load_volume(vol(), scalar_format="u8")
        hs = list(fake.volumes.values())
        assert hs[-1]["sf"] == _lib.HVR_SCALAR["u8"] and hs[-1]["nf"] == _lib.HVR_NORMAL["sn8x3"]

    def test_f16x3_normals_pair_with_f16_scalars_only(self, fake):
        r = VolumeRenderer(width=8, height=8, scalar_format="f16", normal_format="f16x3")
        r.load_volume(vol())
        h = list(fake.volumes.values())[-1]
        assert h["sf"] == _lib.HVR_SCALAR["f16"] and h["nf"] == _lib.HVR_NORMAL["f16x3"]
        n = len(fake.volumes)
        for sf in ("f32", "u8"):
            with pytest.raises(ValueError, match="f16x3"):
                r.load_volume(vol(), scalar_format=sf)
        assert len(fake.volumes) == n              # rejected before any device call
        m = HipManager()
        m.load_shaders()
        m.create_normal_texture(np.zeros((4, 4, 4, 3), np.float32), normal_format="f16x3")
        assert list(fake.volumes.values())[-1]["nf"] == _lib.HVR_NORMAL["f16x3"]

    def test_normals_with_other_shape_get_their_own_volume(self, fake):
        m = HipManager()
        m.load_shaders()
        u0 = m.create_volume_texture(np.zeros((4, 4, 4), np.float32), normal_format="f32x3")
        u1 = m.create_normal_texture(np.zeros((5, 4, 4, 3), np.float32))
        m.set_uniform_int("volume_texture", u0)
        m.set_uniform_int("normal_volume", u1)
        m._sync_bindings()
        s, sr, n, nr = fake.binds[-1]
        assert s != n and fake.volumes[n]["sf"] == _lib.HVR_SCALAR["none"]

    def test_clear_colour_quantised(self, fake):
        m = HipManager(width=4, height=4)
        m.load_shaders()
        m.clear_framebuffer(0.5, 0.3, 0.1, 1.0)
        m.setup_blending()
        m.set_uniform_float("step_size", 0.01)
        m.set_uniform_float("reference_step_size", 0.01)
        m.render_quad()
        assert fake.params[-1]["clear_rgba"] == [128, 77, 26, 255]
