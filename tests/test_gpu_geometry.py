"""Ray-geometry edge cases of the march against the FP64 oracle (GPU).

The in-box run is walked per axis in the direction of its step with one
bound tested per axis, under an iteration bound; these cases stress exactly
that: a camera inside the box (the run starts at the eye), views along the
axes (step components of zero or near zero, odd image sizes so that the
centre pixel's ray is on the axis), an off-centre box with unequal extents
on a non-cubic volume, and presets whose max_steps cuts rays short.

Bar: against the FP32-mode oracle (the shader's own FP32 arithmetic), per
RGBA8 channel |kernel - oracle| <= 1 on >= 99.9 % of pixels and composited
sample counts equal on >= 99.9 %; against the FP64 oracle, no more pixels
beyond 1 LSB than the FP32-mode oracle itself has (the FP32 floor: where a
sample sits on a box face, FP32 and FP64 disagree on whether it is in the
box, and such a pixel may differ by several LSB -- 11 of 6300 pixels in the
offset-box case, for the shader's FP32 arithmetic as for the kernel).
"""

import dataclasses

import numpy as np
import pytest

import scenes
from parity import assert_parity, compare, oracle_render, render_hip
from pyvr_amd.camera import Camera
from pyvr_amd.config import RenderConfig
from pyvr_amd.datasets import compute_normal_volume

pytestmark = pytest.mark.gpu


def check(sc, camera=None):
    print(f"[case] {sc.name} {sc.width}x{sc.height}", flush=True)
    img, cnt = render_hip(sc, camera=camera, counts=True)
    r64 = oracle_render(sc, camera=camera)
    r32 = oracle_render(sc, camera=camera, mode="f32")
    assert_parity(img, r32["rgba8"], label=sc.name + " vs FP32 oracle")
    floor = compare(r32["rgba8"], r64["rgba8"])
    st = compare(img, r64["rgba8"])
    print(f"[parity] {sc.name} vs FP64 oracle {st}; FP32 floor {floor}", flush=True)
    assert st["n_gt1"] <= floor["n_gt1"] + max(1, int(1e-4 * img.shape[0] * img.shape[1])), (st, floor)
    ref32 = r32["counts"]
    same = (cnt.astype(np.int64) == ref32.astype(np.int64)).mean()
    print(f"[counts] {sc.name}: kernel {int(cnt.sum())} oracle32 {int(ref32.sum())} equal {same:.5f}", flush=True)
    assert same >= 0.999 and ref32.sum() > 0


def blob_scene(name, size=(40, 40, 40), image=(97, 97), **kw):
    base = scenes.c1(size=40, image=64)
    data = base.data if tuple(size) == (40, 40, 40) else scenes.blob_field(tuple(size), n_blobs=8, seed=7, noise=1e-3)
    return dataclasses.replace(base, name=name, data=data, normals=compute_normal_volume(data),
                               width=image[0], height=image[1], **kw)


@pytest.mark.parametrize("distance", [0.5, 0.05])
def test_camera_inside_the_box(hip, distance):
    sc = blob_scene(f"inside_d{distance}")
    cam = Camera.from_spherical(target=np.zeros(3, np.float32), azimuth=0.7, elevation=0.3, roll=0.0,
                                distance=distance)
    check(sc, cam)


@pytest.mark.parametrize("az,el", [(0.0, 0.0), (np.pi / 2, 0.0), (np.pi, 0.0), (0.0, -np.pi / 4)])
def test_views_along_the_axes(hip, az, el):
    sc = blob_scene(f"axis_az{az:.2f}_el{el:.2f}")
    cam = Camera.from_spherical(target=np.zeros(3, np.float32), azimuth=az, elevation=el, roll=0.0, distance=3.0)
    check(sc, cam)


def test_offset_box_non_cubic_volume(hip):
    sc = blob_scene("offset_box", size=(20, 33, 17), image=(90, 70),
                    bmin=np.array([-0.3, -1.2, -0.7], np.float32), bmax=np.array([0.9, 0.4, 0.5], np.float32))
    check(sc)


@pytest.mark.parametrize("preset", ["preview", "fast"])
def test_max_steps_cuts_rays(hip, preset):
    sc = blob_scene(f"preset_{preset}", config=getattr(RenderConfig, preset)())
    # rays crossing the box need more than max_steps samples at this step
    assert 2.0 * np.sqrt(3.0) / sc.config.step_size > sc.config.max_steps
    check(sc)


def test_edge_cases_batched_equal_single(hip):
    """The same edge cases as requests of one batch (the batched kernel's
    copy of the march): images and counts equal the single renders byte for
    byte."""
    import test_gpu_batch as tb

    cams = [Camera.from_spherical(target=np.zeros(3, np.float32), azimuth=0.7, elevation=0.3, roll=0.0,
                                  distance=0.05),
            Camera.from_spherical(target=np.zeros(3, np.float32), azimuth=np.pi / 2, elevation=0.0, roll=0.0,
                                  distance=3.0)]
    scs = [blob_scene("inside"), blob_scene("axis"),
           blob_scene("offset_box", size=(20, 33, 17), image=(90, 70),
                      bmin=np.array([-0.3, -1.2, -0.7], np.float32), bmax=np.array([0.9, 0.4, 0.5], np.float32)),
           blob_scene("preset_fast", config=RenderConfig.fast())]
    reqs, vols = [], []
    for k, sc in enumerate(scs):
        vols.append(sc.volume())
        cam = cams[k] if k < len(cams) else sc.camera
        reqs.append(scenes.Request(volume=k, camera=cam, ctf=sc.ctf, otf=sc.otf, width=sc.width,
                                   height=sc.height, config=sc.config, light=sc.light, tau=sc.tau))
    b, idx = tb.make_batch(vols, reqs, scalar_format="f16", normal_format="sn8x3")
    try:
        b.run(count=True)
        tb.check_against_single(b, idx, vols, reqs, "f16", "sn8x3")
    finally:
        b.close()
