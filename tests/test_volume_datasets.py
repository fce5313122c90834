"""Volume container and synthetic datasets.

Pins: the double-sphere closed form evaluated independently (with the 'xy'
meshgrid convention: array axis 0 runs along y), np.gradient-based normals
against central differences written out by hand, and the blob-field
definition evaluated directly (non-separably)."""

import numpy as np
import pytest

from pyvr_amd.datasets import blob_field, blob_params, compute_normal_volume, create_sample_volume
from pyvr_amd.volume import Volume, VolumeError


def test_double_sphere_closed_form():
    n = 17
    v = create_sample_volume(n, "double_sphere")
    lin = np.linspace(-1, 1, n)
    i, j, k = 3, 8, 11   # axis 0 -> y, axis 1 -> x, axis 2 -> z
    x, y, z = lin[j], lin[i], lin[k]
    d1 = np.sqrt((x - 0.3) ** 2 + y ** 2 + z ** 2)
    d2 = np.sqrt((x + 0.3) ** 2 + y ** 2 + z ** 2)
    assert v[i, j, k] == np.float32(max(np.exp(-(4 * d1) ** 2), np.exp(-(4 * d2) ** 2)))
    assert v.dtype == np.float32 and v.shape == (n, n, n)


@pytest.mark.parametrize("shape", ["sphere", "torus", "double_sphere", "cube", "helix", "random_blob"])
def test_all_shapes(shape):
    v = create_sample_volume(16, shape)
    assert v.shape == (16, 16, 16) and v.dtype == np.float32 and np.isfinite(v).all()


def test_unknown_shape():
    with pytest.raises(ValueError, match="Unknown shape"):
        create_sample_volume(8, "teapot")


def test_normals_central_differences():
    rng = np.random.default_rng(0)
    v = rng.random((6, 7, 8)).astype(np.float32)
    n = compute_normal_volume(v)
    assert n.shape == (6, 7, 8, 3) and n.dtype == np.float32
    g = np.array([(v[3, 4, 5] - v[1, 4, 5]) / 2, (v[2, 5, 5] - v[2, 3, 5]) / 2, (v[2, 4, 6] - v[2, 4, 4]) / 2])
    assert np.allclose(n[2, 4, 5], g / (np.linalg.norm(g) + 1e-8), atol=1e-6)
    # one-sided at the faces
    g0 = np.array([v[1, 4, 5] - v[0, 4, 5], (v[0, 5, 5] - v[0, 3, 5]) / 2, (v[0, 4, 6] - v[0, 4, 4]) / 2])
    assert np.allclose(n[0, 4, 5], g0 / (np.linalg.norm(g0) + 1e-8), atol=1e-6)
    # flat field -> exact zero vectors
    assert np.all(compute_normal_volume(np.ones((4, 4, 4), np.float32)) == 0)


def test_blob_field_definition():
    shape = (9, 11, 13)
    v = blob_field(shape, n_blobs=3, seed=7, noise=1e-3, slab=4)
    rng, c, s, a = blob_params(3, 7)
    axes = [np.linspace(-1, 1, m) for m in shape]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    ref = sum(a[b] * np.exp(-((X - c[b, 0]) ** 2 + (Y - c[b, 1]) ** 2 + (Z - c[b, 2]) ** 2) / (2 * s[b] ** 2))
              for b in range(3))
    ref = np.clip(ref + rng.uniform(-1e-3, 1e-3, shape), 0, 1).astype(np.float32)
    assert np.abs(v.astype(np.float64) - ref).max() <= 2e-7
    assert v.dtype == np.float32 and v.min() >= 0 and v.max() <= 1


def test_volume_container():
    d = np.zeros((4, 5, 6), np.float32)
    vol = Volume(data=d, normals=np.zeros((4, 5, 6, 3), np.float32), name="x")
    assert vol.shape == (4, 5, 6) and vol.has_normals
    assert np.allclose(vol.min_bounds, -0.5) and np.allclose(vol.dimensions, 1)
    assert np.allclose(vol.voxel_spacing, 1 / np.array([4, 5, 6]))
    assert np.allclose(vol.center, 0)
    c = vol.copy()
    assert c.data is not vol.data and c.name == "x"
    assert "with normals" in repr(vol)
    nv = Volume(data=np.arange(8, dtype=np.float32).reshape(2, 2, 2)).normalize()
    assert nv.data.min() == 0 and nv.data.max() == 1
    z = Volume(data=np.arange(8, dtype=np.float32).reshape(2, 2, 2)).normalize("zscore")
    assert abs(z.data.mean()) < 1e-6
    with pytest.raises(ValueError):
        vol.normalize("bogus")
    v2 = Volume(data=create_sample_volume(8, "sphere"))
    v2.compute_normals()
    assert v2.normals.shape == (8, 8, 8, 3)
    with pytest.raises(ValueError):
        v2.compute_normals("sobel")
    assert issubclass(VolumeError, Exception)


@pytest.mark.parametrize("kw,match", [
    ({"data": [1, 2]}, "numpy array"), ({"data": np.zeros((2, 2))}, "3D"),
    ({"data": np.zeros((2, 2, 2)), "normals": np.zeros((2, 2, 2, 2))}, "shape"),
    ({"data": np.zeros((2, 2, 2)), "min_bounds": np.zeros(2)}, "min_bounds"),
    ({"data": np.zeros((2, 2, 2)), "max_bounds": -np.ones(3, np.float32)}, "greater"),
])
def test_volume_validation(kw, match):
    with pytest.raises(ValueError, match=match):
        Volume(**kw)
