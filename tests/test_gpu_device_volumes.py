"""Volumes built on the device (GPU).

* Normals computed on the device (hvr_volume_upload_scalar_normals) are the
  bytes compute_normal_volume gives on the host: renders with either are
  identical, image and sample counts, for FP32x3, SNORM8x3 and FP16x3 storage and
  non-cubic shapes.
* Blob volumes generated on the device (hvr_volume_generate_blobs) match a
  NumPy restatement of the generator (same axes, product order and hash
  noise) uploaded from the host: only the last bits of exp may differ, so
  the renders agree to the 1-LSB bar.
"""

import numpy as np
import pytest

import scenes
from parity import compare

pytestmark = pytest.mark.gpu


def render(volume, sc, sf, nf, **load_kw):
    from pyvr_amd.hip_renderer import VolumeRenderer

    r = VolumeRenderer(sc.width, sc.height, config=sc.config, light=sc.light, scalar_format=sf, normal_format=nf)
    try:
        r.load_volume(volume, **load_kw)
        r.set_camera(sc.camera)
        r.set_transfer_functions(sc.ctf, sc.otf)
        r.set_opacity_threshold(sc.tau)
        return r.render(), r.sample_counts()
    finally:
        r.close()


@pytest.mark.parametrize("shape", [(48, 48, 48), (20, 33, 17)])
@pytest.mark.parametrize("sf,nf", [("f32", "f32x3"), ("f16", "sn8x3"), ("f16", "f16x3")])
def test_device_normals_equal_host_normals(hip, shape, sf, nf):
    from pyvr_amd.datasets import blob_field, compute_normal_volume
    from pyvr_amd.volume import Volume

    sc = scenes.c1(size=8, image=96)
    data = blob_field(shape, n_blobs=8, seed=1, noise=1e-3)
    host = Volume(data=data, normals=compute_normal_volume(data), min_bounds=-scenes.BOX, max_bounds=scenes.BOX)
    bare = Volume(data=data, min_bounds=-scenes.BOX, max_bounds=scenes.BOX)
    a, ca = render(host, sc, sf, nf)
    b, cb = render(bare, sc, sf, nf, device_normals=True)
    assert a == b and np.array_equal(ca, cb)
    assert ca.sum() > 0


def splitmix64(z):
    z = (z + np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)).astype(np.uint64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)).astype(np.uint64)
    return z ^ (z >> np.uint64(31))


def generator_reference(g):
    """NumPy restatement of k_blob_field (float64, then float32)."""
    centers, sigmas, amps = g.blob_params()
    axes = [np.linspace(-1, 1, s) for s in g.shape]
    acc = np.zeros(g.shape, np.float64)
    for b in range(len(sigmas)):
        k = 2.0 * sigmas[b] * sigmas[b]
        f = [np.exp(-((axes[a] - centers[b, a]) ** 2) / k) for a in range(3)]
        acc += amps[b] * (f[0][:, None, None] * (f[1][:, None] * f[2][None, :])[None])
    n = np.prod(g.shape)
    with np.errstate(over="ignore"):
        h = splitmix64(np.uint64(g.seed) ^ splitmix64(np.arange(n, dtype=np.uint64)))
    u = (h >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    acc += ((2.0 * u - 1.0) * g.noise).reshape(g.shape)
    return np.clip(acc, 0.0, 1.0).astype(np.float32)


def test_generated_blobs_match_reference(hip):
    from pyvr_amd.datasets import compute_normal_volume
    from pyvr_amd.hip_renderer import BatchRenderer
    from pyvr_amd.hip_renderer.batch import DeviceBlobVolume
    from pyvr_amd.volume import Volume

    g = DeviceBlobVolume(shape=(40, 48, 36), n_blobs=8, seed=3, noise=1e-3)
    ref = generator_reference(g)
    host = Volume(data=ref, normals=compute_normal_volume(ref), min_bounds=g.min_bounds, max_bounds=g.max_bounds)
    reqs = scenes.c3_requests(4, seed=31, image_scale=8)
    b = BatchRenderer(scalar_format="f32", normal_format="f32x3")
    try:
        for q in reqs:
            b.add(g, q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
            b.add(host, q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
        b.run()
        for k, q in enumerate(reqs):
            a = np.frombuffer(b.result(2 * k), np.uint8).reshape(q.height, q.width, 4)
            e = np.frombuffer(b.result(2 * k + 1), np.uint8).reshape(q.height, q.width, 4)
            st = compare(a, e)
            print(f"[generated] request {k}: {st}", flush=True)
            assert st["frac_le1"] >= 0.999 and e[..., 3].max() > 0
    finally:
        b.close()
