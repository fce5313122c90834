"""The serving loop on the GPU: requests submitted in waves while earlier
batches are in flight, under a memory budget that forces evictions; every
result equals what a fresh single-request renderer gives for it (each
result is the same alone and in the mix)."""

import time

import numpy as np
import pytest

import scenes
from test_gpu_batch import single

pytestmark = pytest.mark.gpu


def test_served_results_equal_single_renders(hip):
    from pyvr_amd.hip_renderer.serve import (RenderServer, ServeRequest, batch_pool_bytes, volume_bricks,
                                             volume_device_bytes)

    vols = scenes.c3_volumes(scale=4)
    sizes = sorted(volume_device_bytes(v.data.shape, "f16", "sn8x3", True) for v in vols)
    pipeline, max_px = 2, 1 << 16
    pools = pipeline * batch_pool_bytes(5, max_px, max(volume_bricks(v.data.shape) for v in vols))
    budget = pools + sizes[-1] + sizes[-2] + 1      # two largest volumes at most
    srv = RenderServer(scalar_format="f16", normal_format="sn8x3", memory_budget=budget,
                       max_batch_requests=5, max_batch_pixels=max_px, pipeline=pipeline)
    try:
        for k, v in enumerate(vols):
            srv.register(k, v)
        reqs = scenes.c3_requests(24, seed=41, image_scale=4)
        sub = []
        for w in range(3):
            for q in reqs[8 * w: 8 * w + 8]:
                sub.append((q, srv.submit(ServeRequest(volume=q.volume, camera=q.camera, ctf=q.ctf, otf=q.otf,
                                                       width=q.width, height=q.height, config=q.config,
                                                       light=q.light, tau=q.tau))))
            srv.poll()
        t0 = time.perf_counter()
        srv.drain()
        print(f"[serve] {srv.batches} batches, {srv.evictions} evictions, drained in "
              f"{(time.perf_counter() - t0) * 1e3:.1f} ms", flush=True)
        assert srv.evictions > 0 and srv.batches >= 5
        for q, r in sub:
            img, _ = single(q, vols[q.volume], "f16", "sn8x3")
            assert r.rgba is not None and r.rgba.tobytes() == img
            assert r.latency >= 0
    finally:
        srv.close()


def test_served_camera_path_equals_single_renders(hip):
    """A camera animation (CameraController moves and a CameraPath through
    them) served as a frame stream: every frame is the image a single
    renderer gives for that camera, and the frames come back in order."""
    from pyvr_amd.camera import Camera, CameraController, CameraPath
    from pyvr_amd.hip_renderer.serve import RenderServer, path_requests

    vol = scenes.c3_volumes(scale=4)[1]
    c = CameraController(scenes.standard_camera())
    keys = [c.params.copy()]
    c.orbit(0.6, 0.2)
    c.zoom(1.2)
    keys.append(c.params.copy())
    c.trackball(dx=80, dy=-40, viewport_width=96, viewport_height=64)
    keys.append(c.params.copy())
    frames = CameraPath(keys).generate_frames(9)
    base = scenes.c3_requests(1, seed=43, image_scale=8)[0]
    srv = RenderServer(scalar_format="f16", normal_format="sn8x3", memory_budget=1 << 30, max_batch_requests=4,
                       max_batch_pixels=1 << 16, pipeline=2)
    try:
        srv.register("v", vol)
        reqs = path_requests("v", frames, base.ctf, base.otf, 96, 64, config=base.config, light=base.light,
                             tau=base.tau, tag="path")
        for r in reqs:
            srv.submit(r)
        done = srv.drain()
        assert [r.tag[1] for r in done] == list(range(9)) and srv.batches >= 3
        for r, cam in zip(done, frames):
            q = scenes.Request(0, cam, base.ctf, base.otf, 96, 64, base.config, base.light, base.tau)
            img, cnt = single(q, vol, "f16", "sn8x3")
            assert r.rgba.tobytes() == img and cnt.sum() > 0
    finally:
        srv.close()


def test_rejected_request_does_not_stop_the_server(hip):
    """A request the batched path cannot take (a LUT of 4096 entries; the
    batch kernels hold at most 2048 in LDS) comes back with an error and
    no pixels; the requests around it render as they do alone."""
    import copy

    from pyvr_amd.hip_renderer.serve import RenderServer, ServeRequest

    vols = scenes.c3_volumes(scale=4)
    reqs = scenes.c3_requests(5, seed=43, image_scale=4)
    srv = RenderServer(scalar_format="f16", normal_format="sn8x3", max_batch_requests=8, pipeline=2)
    try:
        for k, v in enumerate(vols):
            srv.register(k, v)
        sub = []
        for i, q in enumerate(reqs):
            otf = q.otf
            if i == 2:
                otf = copy.deepcopy(q.otf)
                otf.lut_size = 4096
            sub.append(srv.submit(ServeRequest(volume=q.volume, camera=q.camera, ctf=q.ctf, otf=otf,
                                               width=q.width, height=q.height, config=q.config, light=q.light,
                                               tau=q.tau)))
        done = srv.drain()
        assert sorted(map(id, done)) == sorted(map(id, sub)) and srv.failed == 1
        assert sub[2].error is not None and "2048" in sub[2].error and sub[2].rgba is None
        for i, (q, r) in enumerate(zip(reqs, sub)):
            if i == 2:
                continue
            img, _ = single(q, vols[q.volume], "f16", "sn8x3")
            assert r.error is None and r.rgba.tobytes() == img
    finally:
        srv.close()
