"""Render scenes shared by the tests, the smoke check and the benchmark.

Each scene fixes everything the reference path consumes: volume (data,
normals, bounds), camera, light, transfer functions, ``RenderConfig``,
early-termination alpha and image size.  ``uniforms()`` produces the uniform
block exactly as ``VolumeRenderer.set_camera`` / ``_update_*`` would set it,
in the form the oracle takes.

Workloads (sizes as in the benchmark configurations):
  c0  128^3 double_sphere (float32, normals), 512x512, balanced, ET 0.99
  c1  256^3 8-blob field (seed 1), 1024x1024, high_quality, ET 0.95,
      6-point seeded OTF
  c2  512^3 32-blob field (seed 2), 1024x1024, high_quality, 24 cameras
  c3  64 batched requests (seed 3): FP16 blob volumes of 64^3, 128^3,
      256^3 and 96x160x224, images of 256^2, 512^2, 768x432 and 1024^2,
      cameras around the box, own OTFs and one of four colormaps
All take a scale-down (``size``, ``image``, ``scale``) for fast tests.
"""

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from pyvr_amd.camera import Camera
from pyvr_amd.config import RenderConfig
from pyvr_amd.datasets import blob_field, compute_normal_volume, create_sample_volume
from pyvr_amd.lighting import Light
from pyvr_amd.transferfunctions import ColorTransferFunction, OpacityTransferFunction
from pyvr_amd.volume import Volume

BOX = np.array([1.0, 1.0, 1.0], dtype=np.float32)


@dataclass
class Scene:
    name: str
    data: np.ndarray
    normals: Optional[np.ndarray]
    camera: Camera
    light: Light
    ctf: ColorTransferFunction
    otf: OpacityTransferFunction
    config: RenderConfig
    width: int
    height: int
    tau: float = 0.99
    bmin: np.ndarray = field(default_factory=lambda: -BOX.copy())
    bmax: np.ndarray = field(default_factory=lambda: BOX.copy())
    lut_size: Optional[int] = None

    def volume(self) -> Volume:
        return Volume(data=self.data, normals=self.normals, min_bounds=self.bmin, max_bounds=self.bmax)

    def lut(self) -> np.ndarray:
        n = self.lut_size or max(self.ctf.lut_size, self.otf.lut_size)
        rgba = np.empty((n, 4), np.float32)
        rgba[:, :3] = self.ctf.to_lut(n)
        rgba[:, 3] = self.otf.to_lut(n)
        return rgba

    def uniforms(self, camera: Optional[Camera] = None) -> dict:
        cam = camera or self.camera
        pos, _ = cam.get_camera_vectors()
        return dict(
            width=self.width, height=self.height,
            view=cam.get_view_matrix(),
            proj=cam.get_projection_matrix(self.width / self.height),
            cam_pos=np.asarray(pos, dtype=np.float32),
            bmin=self.bmin, bmax=self.bmax,
            step=self.config.step_size, ref_step=self.config.reference_step_size,
            max_steps=self.config.max_steps,
            ambient=self.light.ambient_intensity, diffuse=self.light.diffuse_intensity,
            light_pos=np.asarray(self.light.position, np.float32),
            light_target=np.asarray(self.light.target, np.float32),
            et_threshold=self.tau,
        )


def standard_camera(azimuth=np.pi / 4, elevation=np.pi / 6, distance=3.0) -> Camera:
    return Camera.from_spherical(target=np.array([0.0, 0.0, 0.0], dtype=np.float32),
                                 azimuth=azimuth, elevation=elevation, roll=0.0, distance=distance)


def standard_light() -> Light:
    return Light.directional(direction=[1, -1, 0], ambient=0.2, diffuse=0.8)


def seeded_otf(n_points: int, seed: int, y_max: float = 0.5) -> OpacityTransferFunction:
    """Piecewise-linear OTF with control points x = [0, sorted U(0,1)^(n-2), 1]
    and y = [0, U(0, y_max)^(n-1)], drawn from ``default_rng(100 + seed)``.

    Density 0 (empty space) is transparent.  With a random opacity at 0 the
    whole box is a fog and every ray reaches the termination alpha within
    ~20 samples of the box face, whatever the volume holds."""
    rng = np.random.default_rng(100 + seed)
    xs = np.concatenate([[0.0], np.sort(rng.uniform(0.0, 1.0, n_points - 2)), [1.0]])
    ys = np.concatenate([[0.0], rng.uniform(0.0, y_max, n_points - 1)])
    return OpacityTransferFunction([(float(x), float(y)) for x, y in zip(xs, ys)])


def c0(size: int = 128, image: int = 512) -> Scene:
    data = create_sample_volume(size, "double_sphere")
    return Scene(
        name=f"c0_double_sphere_{size}_{image}",
        data=data, normals=compute_normal_volume(data),
        camera=standard_camera(), light=standard_light(),
        ctf=ColorTransferFunction.from_colormap("viridis"),
        otf=OpacityTransferFunction.linear(0.0, 0.1),
        config=RenderConfig.balanced(), width=image, height=image, tau=0.99,
    )


def c1(size: int = 256, image: int = 1024) -> Scene:
    data = blob_field(size, n_blobs=8, seed=1, noise=1e-3)
    return Scene(
        name=f"c1_blobs8_{size}_{image}",
        data=data, normals=compute_normal_volume(data),
        camera=standard_camera(), light=standard_light(),
        ctf=ColorTransferFunction.from_colormap("viridis"),
        otf=seeded_otf(6, seed=1),
        config=RenderConfig.high_quality(), width=image, height=image, tau=0.95,
    )


def c2_cameras(n: int = 24, seed: int = 2):
    """Cameras on the sphere of radius 3: azimuth U[0, 2pi), elevation
    asin(U[-1, 1]), drawn from default_rng(200 + seed)."""
    rng = np.random.default_rng(200 + seed)
    az = rng.uniform(0.0, 2 * np.pi, n)
    el = np.arcsin(rng.uniform(-1.0, 1.0, n))
    return [standard_camera(float(a), float(e), 3.0) for a, e in zip(az, el)]


def c2(size: int = 512, image: int = 1024, with_normals: bool = True) -> Scene:
    data = blob_field(size, n_blobs=32, seed=2, noise=1e-3)
    return Scene(
        name=f"c2_blobs32_{size}_{image}",
        data=data, normals=compute_normal_volume(data) if with_normals else None,
        camera=c2_cameras(1)[0], light=standard_light(),
        ctf=ColorTransferFunction.from_colormap("viridis"),
        otf=seeded_otf(6, seed=2),
        config=RenderConfig.high_quality(), width=image, height=image, tau=0.95,
    )


# ---------------------------------------------------------------------------
# c3: batched requests

C3_SHAPES = [(64, 64, 64), (128, 128, 128), (256, 256, 256), (96, 160, 224)]
C3_IMAGES = [(256, 256), (512, 512), (768, 432), (1024, 1024)]
C3_COLORMAPS = ["viridis", "plasma", "inferno", "magma"]


@dataclass
class Request:
    """One render request: what a fresh renderer of (width, height) renders
    after load_volume(volumes[volume]), set_camera, set_light,
    set_transfer_functions and set_opacity_threshold(tau)."""
    volume: int
    camera: Camera
    ctf: ColorTransferFunction
    otf: OpacityTransferFunction
    width: int
    height: int
    config: RenderConfig
    light: Light
    tau: float

    def lut(self) -> np.ndarray:
        n = max(self.ctf.lut_size, self.otf.lut_size)
        rgba = np.empty((n, 4), np.float32)
        rgba[:, :3] = self.ctf.to_lut(n)
        rgba[:, 3] = self.otf.to_lut(n)
        return rgba


def request_uniforms(q: Request, volume) -> dict:
    """The oracle's uniform block for request ``q`` on ``volume`` (a
    ``Volume`` or a device volume with bounds): what a fresh renderer of the
    request's size sets after set_camera, set_light and
    set_opacity_threshold."""
    pos, _ = q.camera.get_camera_vectors()
    return dict(
        width=q.width, height=q.height,
        view=q.camera.get_view_matrix(),
        proj=q.camera.get_projection_matrix(q.width / q.height),
        cam_pos=np.asarray(pos, dtype=np.float32),
        bmin=np.asarray(volume.min_bounds, np.float32), bmax=np.asarray(volume.max_bounds, np.float32),
        step=q.config.step_size, ref_step=q.config.reference_step_size, max_steps=q.config.max_steps,
        ambient=q.light.ambient_intensity, diffuse=q.light.diffuse_intensity,
        light_pos=np.asarray(q.light.position, np.float32),
        light_target=np.asarray(q.light.target, np.float32),
        et_threshold=q.tau,
    )


def c3_volumes(scale: int = 1, seed: int = 3, with_normals: bool = True, shapes=None):
    """The batch's volumes: blob fields (8 blobs, seed 10*seed + k) over
    [-1, 1]^3 with central-difference normals; sizes divided by ``scale``."""
    vols = []
    for k, shape in enumerate(shapes or C3_SHAPES):
        shp = tuple(max(8, s // scale) for s in shape)
        data = blob_field(shp, n_blobs=8, seed=10 * seed + k, noise=1e-3)
        vols.append(Volume(data=data, normals=compute_normal_volume(data) if with_normals else None,
                           min_bounds=-BOX, max_bounds=BOX))
    return vols


def c3_requests(n: int = 64, seed: int = 3, image_scale: int = 1, nvolumes: int = 4):
    """``n`` requests drawn from ``default_rng(300 + seed)``: volume index,
    image size, azimuth U[0, 2pi), elevation U[-pi/3, pi/3], distance
    U[2.5, 4], OTF with 4-8 points (its own seed), colormap; balanced
    preset, default termination alpha."""
    rng = np.random.default_rng(300 + seed)
    out = []
    for _ in range(n):
        v = int(rng.integers(nvolumes))
        w, h = C3_IMAGES[int(rng.integers(len(C3_IMAGES)))]
        az = float(rng.uniform(0.0, 2 * np.pi))
        el = float(rng.uniform(-np.pi / 3, np.pi / 3))
        d = float(rng.uniform(2.5, 4.0))
        npts = int(rng.integers(4, 9))
        otf_seed = int(rng.integers(1 << 20))
        cmap = C3_COLORMAPS[int(rng.integers(len(C3_COLORMAPS)))]
        out.append(Request(volume=v, camera=standard_camera(az, el, d),
                           ctf=ColorTransferFunction.from_colormap(cmap), otf=seeded_otf(npts, otf_seed),
                           width=max(8, w // image_scale), height=max(8, h // image_scale),
                           config=RenderConfig.balanced(), light=standard_light(), tau=0.99))
    return out


# ---------------------------------------------------------------------------
# c4: serving

C4_PRESETS = ("fast", "balanced", "high_quality")


def c4_sizes(n: int = 16):
    """N_k = 256 * 4^(k/15) rounded to a multiple of 16 (256 ... 1024)."""
    return [int(round(256 * 4 ** (k / (n - 1)) / 16)) * 16 for k in range(n)]


def c4_volumes(n: int = 16, seed: int = 4, scale: int = 1):
    """Device-generated blob volumes (8 blobs, seed 100*seed + k, noise
    1e-3) of sizes c4_sizes(n) / scale, normals computed on the device."""
    from pyvr_amd.hip_renderer.batch import DeviceBlobVolume

    return [DeviceBlobVolume(shape=(max(8, s // scale),) * 3, n_blobs=8, seed=100 * seed + k, noise=1e-3)
            for k, s in enumerate(c4_sizes(n))]


def c4_requests(count: int, nvol: int = 16, seed: int = 4, image_scale: int = 1):
    """Requests drawn from default_rng(400 + seed): volume by a truncated
    Zipf(1.1) over ranks 1..nvol (rank r -> a seeded permutation of the
    volumes), square image side U[512, 2048] rounded to 16, preset uniform
    over fast / balanced / high_quality, camera, OTF and colormap as c3;
    default termination alpha."""
    rng = np.random.default_rng(400 + seed)
    perm = rng.permutation(nvol)
    p = 1.0 / np.arange(1, nvol + 1) ** 1.1
    p /= p.sum()
    out = []
    for _ in range(count):
        v = int(perm[rng.choice(nvol, p=p)])
        side = int(round(rng.uniform(512, 2048) / 16)) * 16
        preset = C4_PRESETS[int(rng.integers(3))]
        az = float(rng.uniform(0.0, 2 * np.pi))
        el = float(rng.uniform(-np.pi / 3, np.pi / 3))
        d = float(rng.uniform(2.5, 4.0))
        npts = int(rng.integers(4, 9))
        otf_seed = int(rng.integers(1 << 20))
        cmap = C3_COLORMAPS[int(rng.integers(len(C3_COLORMAPS)))]
        side = max(16, side // image_scale)
        out.append(Request(volume=v, camera=standard_camera(az, el, d),
                           ctf=ColorTransferFunction.from_colormap(cmap), otf=seeded_otf(npts, otf_seed),
                           width=side, height=side, config=getattr(RenderConfig, preset)(),
                           light=standard_light(), tau=0.99))
    return out


# ---------------------------------------------------------------------------
# storage-format round trips as the device performs them (float32 results);
# the oracle consumes these to check a quantised render against the same data
def dequant_scalar(data: np.ndarray, fmt: str) -> np.ndarray:
    d = np.asarray(data, np.float32)
    if fmt == "f32":
        return d
    if fmt == "f16":
        return d.astype(np.float16).astype(np.float32)
    if fmt == "u8":
        q = np.rint(np.clip(d, 0.0, 1.0) * np.float32(255.0)).astype(np.float32)
        return q * np.float32(1.0 / 255.0)
    raise ValueError(fmt)


def dequant_normals(n: np.ndarray, fmt: str) -> np.ndarray:
    a = np.asarray(n, np.float32)
    if fmt == "f32x3":
        return a
    if fmt == "sn8x3":
        q = np.rint(np.clip(a, -1.0, 1.0) * np.float32(127.0)).astype(np.float32)
        return q  # direction only matters: the renderer normalises
    if fmt == "f16x3":
        return a.astype(np.float16).astype(np.float32)
    raise ValueError(fmt)
