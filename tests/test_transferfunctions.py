"""Transfer functions: known answers of the reference's tests
(tests/test_transferfunctions/test_color.py:17-176, test_opacity.py:16-100)
and the colormap fixtures in tests/golden/colormaps.npz (the 12 maps of the
reference's colour selector)."""

import os

import numpy as np
import pytest

from pyvr_amd.transferfunctions import (ColorTransferFunction, InvalidControlPointError,
                                        OpacityTransferFunction)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "colormaps.npz")


def test_color_defaults_and_sorting():
    ctf = ColorTransferFunction()
    assert ctf.control_points == [(0.0, (0.0, 0.0, 0.0)), (1.0, (1.0, 1.0, 1.0))]
    assert ctf.lut_size == 256
    pts = [(1.0, (0, 0, 1.0)), (0.0, (1.0, 0, 0)), (0.5, (0, 1.0, 0))]
    assert [p[0] for p in ColorTransferFunction(pts).control_points] == [0.0, 0.5, 1.0]


def test_grayscale_lut():
    lut = ColorTransferFunction.grayscale(lut_size=128).to_lut()
    assert lut.shape == (128, 3) and lut.dtype == np.float32
    assert np.allclose(lut[:, 0], lut[:, 1]) and np.allclose(lut[:, 1], lut[:, 2])
    assert np.allclose(lut[0], 0) and np.allclose(lut[-1], 1)
    # rows sit at linspace(0, 1, N)
    assert np.allclose(lut[:, 0], np.linspace(0, 1, 128), atol=1e-7)


def test_two_color_ramp_and_single():
    blue, yellow = (0.0, 0.0, 1.0), (1.0, 1.0, 0.0)
    lut = ColorTransferFunction.two_color_ramp(blue, yellow).to_lut(256)
    assert np.allclose(lut[0], blue) and np.allclose(lut[-1], yellow)
    assert np.allclose(lut[128], (np.array(blue) + np.array(yellow)) / 2, atol=1e-2)
    s = ColorTransferFunction.single_color((1.0, 0.0, 0.0), lut_size=100).to_lut()
    assert s.shape == (100, 3) and np.allclose(s, [1, 0, 0])


def test_color_queries():
    ctf = ColorTransferFunction.two_color_ramp((1.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    assert np.allclose(ctf.get_color_at(0.5), (0.5, 0.5, 0.0))
    g = ColorTransferFunction.grayscale()
    a = np.array([[0.0, 0.5], [0.25, 1.0]])
    out = g.apply_to_array(a)
    assert out.shape == (2, 2, 3) and np.allclose(out[..., 0], a)
    assert np.allclose(g(64), g.to_lut(64))


UI_COLORMAPS = ["viridis", "plasma", "inferno", "magma", "cividis", "gray", "bone", "copper", "hot", "cool",
                "turbo", "jet"]


@pytest.mark.parametrize("name", UI_COLORMAPS)
def test_from_colormap_reproduces_listed_colours(name):
    ref = np.load(GOLDEN)[name]
    ctf = ColorTransferFunction.from_colormap(name)
    assert len(ctf.control_points) == 256
    lut = ctf.to_lut(256)
    assert np.array_equal(lut, ref.astype(np.float32))


def test_segmented_colormaps_closed_forms():
    """gray is (x, x, x) and cool is (x, 1 - x, 1) at x = i / 255: the
    fixtures and from_colormap agree with the formulas, not only with the
    installed matplotlib."""
    x = np.linspace(0.0, 1.0, 256)
    g = np.load(GOLDEN)
    assert np.allclose(g["gray"], np.stack([x, x, x], 1), atol=1e-12)
    assert np.allclose(g["cool"], np.stack([x, 1 - x, np.ones_like(x)], 1), atol=1e-12)
    lut = ColorTransferFunction.from_colormap("cool").to_lut(256)
    assert np.array_equal(lut, np.stack([x, 1 - x, np.ones_like(x)], 1).astype(np.float32))


def test_from_colormap_range_and_errors():
    ctf = ColorTransferFunction.from_colormap("viridis", value_range=(0.2, 0.8), lut_size=16)
    assert ctf.control_points[0][0] == pytest.approx(0.2) and ctf.control_points[-1][0] == pytest.approx(0.8)
    with pytest.raises(ValueError):
        ColorTransferFunction.from_colormap("no_such_map")


@pytest.mark.parametrize("pts", [[], [(0.0, (1.0, 0.0))], [(1.5, (0, 0, 0))], [(0.0, (2.0, 0, 0))],
                                 [(0.0, "red")], [("a", (0, 0, 0))]])
def test_color_validation(pts):
    with pytest.raises((InvalidControlPointError, ValueError)):
        ColorTransferFunction(pts)


def test_opacity_linear_kat():
    otf = OpacityTransferFunction.linear(0.2, 0.8, lut_size=128)
    assert otf.control_points == [(0.0, 0.2), (1.0, 0.8)]
    lut = otf.to_lut()
    assert lut.shape == (128,) and lut.dtype == np.float32
    assert lut[0] == pytest.approx(0.2, abs=1e-6) and lut[-1] == pytest.approx(0.8, abs=1e-6)


def test_opacity_linear_lut_closed_form():
    # linear(0, 0.1) at N = 256: row j holds 0.1 * j / 255
    lut = OpacityTransferFunction.linear(0.0, 0.1).to_lut(256)
    assert np.allclose(lut, (0.1 * np.arange(256) / 255).astype(np.float32), atol=1e-8)


def test_one_step_and_peaks():
    otf = OpacityTransferFunction.one_step(step=0.5, low=0.1, high=0.9)
    assert len(otf.control_points) == 4
    lut = otf.to_lut(1000)
    assert np.all(lut[:500] <= 0.1 + 1e-6) and np.all(lut[500:] >= 0.9 - 1e-6)
    pk = OpacityTransferFunction.peaks([0.5], opacity=0.9, eps=0.1, base=0.1).to_lut(1000)
    assert pk[500] >= 0.8 and pk[0] == pytest.approx(0.1) and pk[-1] == pytest.approx(0.1)
    multi = OpacityTransferFunction.peaks([0.2, 0.8], opacity=0.9, eps=0.05).to_lut(1000)
    assert multi[200] >= 0.8 and multi[800] >= 0.8
    with pytest.raises(ValueError, match="At least one peak"):
        OpacityTransferFunction.peaks([])
    with pytest.raises(ValueError):
        OpacityTransferFunction.peaks([1.5])


def test_opacity_validation_and_apply():
    with pytest.raises(InvalidControlPointError):
        OpacityTransferFunction([(0.0, 1.5)])
    with pytest.raises(ValueError):
        OpacityTransferFunction([(-0.5, 0.5)])
    otf = OpacityTransferFunction.linear(0.0, 1.0)
    a = np.array([0.0, 0.25, 1.0])
    assert np.allclose(otf.apply_to_array(a), a)
    assert "control_points" in repr(otf)
