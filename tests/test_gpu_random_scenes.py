"""Seeded random scenes through the single-frame and batched paths (GPU).

Each scene draws what the fixed-scene tests hold constant: a non-cubic
volume (3 to 40 voxels per axis) in an offset box of unequal extents, with
or without normals, stored in any format pair the library supports, in 8^3
or 16^3 bricks; a LUT of 1 to 2048 entries from a random colormap and a
seeded or linear opacity function; step, reference step and max_steps (down to 1); a
termination alpha from 0.5 to above 1 (no early termination); a directional
or point light; a camera outside or inside the box, with roll; an image of
1 to 96 pixels per side.

Bar, per scene: against the FP32-mode oracle on the stored (dequantised)
values, at most max(1, 0.1 %) pixels more than 1 LSB off and sample counts
equal on all but max(1, 0.1 %) pixels; against the FP64 oracle, the
pixels more than 1 LSB off are the FP32-mode oracle's own, pixel for
pixel, plus at most max(1, 0.01 %) others.  The batched path renders every scene of a storage
configuration in one launch and must give the single-frame path's bytes and
counts exactly.
"""

import numpy as np
import pytest

import scenes
from parity import compare, oracle_render, render_hip
from pyvr_amd.camera import Camera
from pyvr_amd.config import RenderConfig
from pyvr_amd.datasets import blob_field, compute_normal_volume
from pyvr_amd.lighting import Light
from pyvr_amd.transferfunctions import ColorTransferFunction, OpacityTransferFunction

pytestmark = pytest.mark.gpu

FORMATS = [("f32", "f32x3"), ("f16", "f16x3"), ("f16", "sn8x3"), ("u8", "sn8x3"), ("f32", "sn8x3"),
           ("f16", "f32x3"), ("u8", "f32x3")]
COLORMAPS = ["viridis", "plasma", "inferno", "magma", "cividis", "twilight", "coolwarm", "gray"]
NSCENES = 48


def random_scene(seed):
    rng = np.random.default_rng(1000 + seed)
    shape = tuple(int(x) for x in rng.integers(3, 41, 3))
    data = blob_field(shape, n_blobs=int(rng.integers(1, 9)), seed=seed, noise=1e-3)
    normals = compute_normal_volume(data) if rng.random() < 0.75 else None
    lo = rng.uniform(-1.5, 0.5, 3).astype(np.float32)
    hi = (lo + rng.uniform(0.4, 2.5, 3)).astype(np.float32)
    lut_n = int(rng.choice([1, 2, 7, 64, 256, 1000, 2048]))
    ctf = ColorTransferFunction.from_colormap(str(rng.choice(COLORMAPS)), lut_size=lut_n)
    if rng.random() < 0.5:
        otf = scenes.seeded_otf(int(rng.integers(2, 8)), seed=seed, y_max=float(rng.uniform(0.05, 1.0)))
    else:   # opaque at density 0 too: a 1-entry LUT is then a uniform fog
        otf = OpacityTransferFunction.linear(float(rng.uniform(0.0, 0.3)), float(rng.uniform(0.3, 1.0)))
    otf.lut_size = lut_n
    config = RenderConfig(step_size=float(rng.uniform(0.004, 0.05)),
                          max_steps=int(rng.choice([1, 3, 17, 100, 500, 1000])),
                          reference_step_size=float(rng.uniform(0.004, 0.02)))
    ambient, diffuse = float(rng.uniform(0.0, 0.5)), float(rng.uniform(0.2, 1.0))
    if rng.random() < 0.5:
        light = Light.directional(direction=rng.normal(size=3), ambient=ambient, diffuse=diffuse)
    else:
        light = Light.point_light(position=rng.normal(size=3) * 3.0, ambient=ambient, diffuse=diffuse)
    centre = ((lo + hi) / 2).astype(np.float32)
    camera = Camera.from_spherical(target=centre, azimuth=float(rng.uniform(0, 2 * np.pi)),
                                   elevation=float(rng.uniform(-1.4, 1.4)), roll=float(rng.uniform(-0.5, 0.5)),
                                   distance=float(rng.choice([0.05, 0.3, 2.0, 4.0])))
    sf, nf = FORMATS[int(rng.integers(len(FORMATS)))]
    sc = scenes.Scene(name=f"random{seed}", data=data, normals=normals, camera=camera, light=light, ctf=ctf,
                      otf=otf, config=config, width=int(rng.integers(1, 97)), height=int(rng.integers(1, 97)),
                      tau=float(rng.choice([0.5, 0.9, 0.95, 0.99, 1.0, 1.5])), bmin=lo, bmax=hi)
    return sc, sf, nf, int(rng.choice([3, 4]))


def oracles(sc, sf, nf):
    s = scenes.dequant_scalar(sc.data, sf)
    n = scenes.dequant_normals(sc.normals, nf) if sc.normals is not None else s
    return (oracle_render(sc, mode="f32", scalar=s, normals=n), oracle_render(sc, mode="f64", scalar=s, normals=n))


def test_random_scenes_single_and_batched(hip):
    from pyvr_amd.hip_renderer import BatchRenderer

    cases = [random_scene(k) for k in range(NSCENES)]
    singles = []
    for sc, sf, nf, lb in cases:
        img, cnt = render_hip(sc, scalar_format=sf, normal_format=nf, brick_log2=lb, counts=True)
        r32, r64 = oracles(sc, sf, nf)
        npix = sc.width * sc.height
        s32, floor = compare(img, r32["rgba8"]), compare(r32["rgba8"], r64["rgba8"])
        diff = int((cnt.astype(np.int64) != r32["counts"].astype(np.int64)).sum())
        err = lambda a: np.abs(a.astype(np.int16) - r64["rgba8"].astype(np.int16)).max(axis=-1) > 1
        own = int((err(img) & ~err(r32["rgba8"])).sum())   # beyond 1 LSB where FP32 arithmetic is not
        print(f"[random] {sc.name} {sc.data.shape} {sf}+{nf if sc.normals is not None else 'none'} lb{lb} "
              f"lut{sc.lut().shape[0]} {sc.width}x{sc.height} steps{sc.config.max_steps} tau{sc.tau}: "
              f"vs32 max {s32['max']} n_gt1 {s32['n_gt1']}; vs64 kernel-only {own} (floor {floor['n_gt1']}); "
              f"counts differ {diff}, samples {int(cnt.sum())}", flush=True)
        assert s32["n_gt1"] <= max(1, int(1e-3 * npix)), (sc.name, s32)
        assert diff <= max(1, int(1e-3 * npix)), (sc.name, diff)
        assert own <= max(1, int(1e-4 * npix)), (sc.name, own, floor)
        singles.append((img, cnt))
    # the batched path: one launch per storage configuration
    groups = {}
    for k, (sc, sf, nf, lb) in enumerate(cases):
        groups.setdefault((sf, nf, lb), []).append(k)
    for (sf, nf, lb), ks in groups.items():
        b = BatchRenderer(scalar_format=sf, normal_format=nf, brick_log2=lb)
        try:
            idx = []
            for k in ks:
                sc = cases[k][0]
                idx.append(b.add(sc.volume(), sc.camera, sc.ctf, sc.otf, sc.width, sc.height, sc.config,
                                 sc.light, sc.tau))
            b.run(count=True)
            for k, i in zip(ks, idx):
                sc = cases[k][0]
                got = np.frombuffer(b.result(i), np.uint8).reshape(sc.height, sc.width, 4)
                assert np.array_equal(got, singles[k][0]), f"{sc.name}: batched image differs"
                assert np.array_equal(b.counts(i), singles[k][1]), f"{sc.name}: batched counts differ"
            print(f"[random] batch {sf}+{nf} lb{lb}: {len(ks)} scenes equal to single renders", flush=True)
        finally:
            b.close()
