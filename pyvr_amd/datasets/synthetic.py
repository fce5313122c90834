"""Synthetic volumes and gradient normals.

``create_sample_volume`` and ``compute_normal_volume`` follow the reference's
``pyvr/datasets/synthetic.py:10-122`` (same shapes and formulas; note the
grid comes from ``np.meshgrid`` with its default 'xy' indexing, so array axis
0 runs along y and axis 1 along x).

``blob_field`` is the seeded Gaussian-blob family used by the larger render
workloads (256^3 .. 1024^3).  It is defined here, not in the reference:

* grid: voxel centres at ``linspace(-1, 1, n)`` per axis, 'ij' indexing
  (array axis k <-> coordinate k);
* ``rng = numpy.random.default_rng(seed)``; draws in this order:
  ``centers = rng.uniform(-0.7, 0.7, (n_blobs, 3))``,
  ``sigmas = rng.uniform(0.05, 0.2, n_blobs)``,
  ``amps = rng.uniform(0.3, 1.0, n_blobs)``, then, if ``noise > 0``,
  ``rng.uniform(-noise, noise, shape)``;
* value = ``sum_b amps[b] * exp(-|x - c_b|^2 / (2 sigma_b^2))`` + noise,
  clipped to [0, 1], float32.  The Gaussian is evaluated as the product of
  three 1-D factors in float64 (separable), slab by slab to bound memory.
"""

import numpy as np


def create_sample_volume(size=64, shape="sphere"):
    """(size, size, size) float32 test volume: 'sphere', 'torus',
    'double_sphere', 'cube', 'helix' or 'random_blob'."""
    lin = np.linspace(-1, 1, size)
    x, y, z = np.meshgrid(lin, lin, lin)

    if shape == "sphere":
        r = np.sqrt(x * x + y * y + z * z)
        vol = np.exp(-((r * 3) ** 2)).astype(np.float32)
    elif shape == "torus":
        ring = np.sqrt(x * x + y * y)
        tube = np.sqrt((ring - 0.6) ** 2 + z * z)
        vol = np.exp(-((tube / 0.3 * 4) ** 2)).astype(np.float32)
    elif shape == "double_sphere":
        r1 = np.sqrt((x - 0.3) ** 2 + y * y + z * z)
        r2 = np.sqrt((x + 0.3) ** 2 + y * y + z * z)
        vol = np.maximum(np.exp(-((r1 * 4) ** 2)), np.exp(-((r2 * 4) ** 2))).astype(np.float32)
    elif shape == "cube":
        box = np.maximum(np.maximum(np.abs(x), np.abs(y)), np.abs(z))
        vol = np.exp(-(((box - 0.4) * 8) ** 2)).astype(np.float32)
        vol[box > 0.6] = 0
    elif shape == "helix":
        phase = z * 3 * 2 * np.pi
        hx = 0.5 * np.cos(phase)
        hy = 0.5 * np.sin(phase)
        dist = np.sqrt((x - hx) ** 2 + (y - hy) ** 2)
        vol = np.exp(-((dist / 0.15 * 3) ** 2)).astype(np.float32)
    elif shape == "random_blob":
        from scipy.ndimage import gaussian_filter

        np.random.seed(42)
        off = np.random.uniform(-1, 1, size=3)
        xo, yo, zo = x + off[0] * 0.5, y + off[1] * 0.5, z + off[2] * 0.5
        noise = np.random.random((size, size, size)).astype(np.float32)
        noise = gaussian_filter(noise, sigma=size / 18)
        ramp = (xo + 1.5) * (yo + 1.2) * (zo + 0.8)
        ramp = ramp / np.max(np.abs(ramp))
        vol = noise * (0.7 + 0.3 * ramp)
        vol = np.maximum(0, vol - 0.25) * 2.5
        for _ in range(3):
            cx, cy, cz = np.random.uniform(-0.7, 0.7, 3)
            dist = np.sqrt((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2)
            vol += np.exp(-((dist * 6) ** 2)) * np.random.uniform(0.5, 1.2)
        vol = np.clip(vol, 0, 1).astype(np.float32)
    else:
        raise ValueError(
            f"Unknown shape: {shape}. Available shapes: sphere, torus, double_sphere, cube, helix, random_blob"
        )
    return vol


def compute_normal_volume(volume):
    """Unit gradient per voxel: ``np.gradient`` (central differences inside,
    one-sided at the faces, unit spacing) divided by ``|g| + 1e-8``; float32,
    shape ``volume.shape + (3,)``.  Flat regions give exact zero vectors."""
    gx, gy, gz = np.gradient(volume)
    g = np.stack((gx, gy, gz), axis=-1)
    g = g / (np.linalg.norm(g, axis=-1, keepdims=True) + 1e-8)
    return g.astype(np.float32)


def blob_params(n_blobs: int, seed: int):
    """(rng, centers, sigmas, amps) of the blob family, in the draw order above."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-0.7, 0.7, (n_blobs, 3))
    sigmas = rng.uniform(0.05, 0.2, n_blobs)
    amps = rng.uniform(0.3, 1.0, n_blobs)
    return rng, centers, sigmas, amps


def blob_field(shape, n_blobs: int = 8, seed: int = 1, noise: float = 1e-3,
               slab: int = 32) -> np.ndarray:
    """Seeded Gaussian-blob volume (definition in the module doc)."""
    if isinstance(shape, int):
        shape = (shape, shape, shape)
    shape = tuple(int(s) for s in shape)
    rng, centers, sigmas, amps = blob_params(n_blobs, seed)
    axes = [np.linspace(-1, 1, s) for s in shape]
    out = np.empty(shape, dtype=np.float32)
    # 1-D factors per blob and axis
    fac = [
        [np.exp(-((axes[k] - centers[b, k]) ** 2) / (2.0 * sigmas[b] ** 2)) for k in range(3)]
        for b in range(n_blobs)
    ]
    noise_gen = None
    if noise > 0:
        noise_gen = rng
    for i0 in range(0, shape[0], slab):
        i1 = min(shape[0], i0 + slab)
        acc = np.zeros((i1 - i0, shape[1], shape[2]), dtype=np.float64)
        for b in range(n_blobs):
            fx, fy, fz = fac[b]
            yz = np.multiply.outer(fy, fz)
            acc += (amps[b] * fx[i0:i1])[:, None, None] * yz[None, :, :]
        if noise_gen is not None:
            acc += noise_gen.uniform(-noise, noise, acc.shape)
        out[i0:i1] = np.clip(acc, 0.0, 1.0)
    return out
