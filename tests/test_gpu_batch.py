"""Batched rendering on the GPU (hvr_batch_*): every request's image and
sample counts equal, byte for byte, what a fresh single-request renderer
gives for the same volume, camera, transfer functions and settings -- for
both ray-tile shapes, both brick sizes, alone and in the mix, on first
launch and on graph replays with new contents."""

import numpy as np
import pytest

import scenes

pytestmark = pytest.mark.gpu


def single(req, vol, sf, nf, brick_log2=3):
    from pyvr_amd.hip_renderer import VolumeRenderer, _lib

    r = VolumeRenderer(req.width, req.height, config=req.config, light=req.light,
                       scalar_format=sf, normal_format=nf)
    try:
        _lib.check(_lib.lib().hvr_set_brick_log2(r.gl_manager.ctx, brick_log2), "set_brick")
        r.load_volume(vol)
        r.set_camera(req.camera)
        r.set_transfer_functions(req.ctf, req.otf)
        r.set_opacity_threshold(req.tau)
        img = r.render()
        cnt = r.sample_counts()
    finally:
        r.close()
    return img, cnt


def make_batch(vols, reqs, **kw):
    from pyvr_amd.hip_renderer import BatchRenderer

    b = BatchRenderer(**kw)
    idx = [b.add(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
           for q in reqs]
    return b, idx


def check_against_single(b, idx, vols, reqs, sf, nf, brick_log2=3):
    for i, q in zip(idx, reqs):
        img, cnt = single(q, vols[q.volume], sf, nf, brick_log2)
        got = b.result(i)
        assert len(got) == q.width * q.height * 4
        if got != img:
            a = np.frombuffer(got, np.uint8).astype(int)
            e = np.frombuffer(img, np.uint8).astype(int)
            raise AssertionError(f"request {i} ({q.width}x{q.height}, volume {q.volume}): "
                                 f"{int((a != e).sum())} bytes differ, max {int(np.abs(a - e).max())}")
        assert np.array_equal(b.counts(i), cnt), f"request {i}: sample counts differ"
        assert cnt.sum() > 0


@pytest.mark.parametrize("tile", ["8x8", "16x4"])
def test_batch_equals_single_renders(hip, tile):
    vols = scenes.c3_volumes(scale=4)
    reqs = scenes.c3_requests(12, image_scale=4)
    b, idx = make_batch(vols, reqs, scalar_format="f16", normal_format="sn8x3", tile_shape=tile)
    try:
        b.run(count=True)
        print(f"[batch] {len(reqs)} requests, {tile}: {b.last_ms():.3f} ms", flush=True)
        check_against_single(b, idx, vols, reqs, "f16", "sn8x3")
    finally:
        b.close()


def test_batch_alone_and_in_mix(hip):
    vols = scenes.c3_volumes(scale=4)
    reqs = scenes.c3_requests(10, seed=5, image_scale=4)
    b, idx = make_batch(vols, reqs)
    try:
        b.run()
        mixed = b.results()
    finally:
        b.close()
    for q, want in zip(reqs, mixed):
        b1, (i1,) = make_batch(vols, [q])
        try:
            b1.run()
            assert b1.result(i1) == want
        finally:
            b1.close()


def test_batch_replays_and_reuse(hip):
    vols = scenes.c3_volumes(scale=4)
    reqs = scenes.c3_requests(8, seed=7, image_scale=4)
    b, idx = make_batch(vols, reqs)
    try:
        b.run()
        first = b.results()
        for _ in range(3):           # graph replays of the same contents
            b.run()
            assert b.results() == first
        # same shape, new cameras and transfer functions: the captured graph
        # is replayed with the new staged contents
        reqs2 = scenes.c3_requests(8, seed=8, image_scale=4)
        b.clear()
        idx2 = [b.add(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
                for q in reqs2]
        b.run(count=True)
        check_against_single(b, idx2, vols, reqs2, "f16", "sn8x3")
    finally:
        b.close()


@pytest.mark.parametrize("sf,nf,lb,normals", [("f32", "f32x3", 4, True), ("u8", "sn8x3", 3, True),
                                               ("f16", "sn8x3", 3, False), ("f32", "f32x3", 3, False),
                                               ("f16", "f16x3", 3, True), ("f16", "f16x3", 4, True)])
def test_batch_formats_bricks_and_no_normals(hip, sf, nf, lb, normals):
    vols = scenes.c3_volumes(scale=4, with_normals=normals)
    reqs = scenes.c3_requests(6, seed=11, image_scale=4)
    b, idx = make_batch(vols, reqs, scalar_format=sf, normal_format=nf, brick_log2=lb)
    try:
        b.run(count=True)
        check_against_single(b, idx, vols, reqs, sf, nf, lb)
    finally:
        b.close()


def test_batch_mixed_classes(hip):
    """Volumes with and without normals in one batch (two kernel
    instances, so two classes launched in the same graph), interleaved in
    request order: every request equals its single render."""
    import dataclasses

    vols = scenes.c3_volumes(scale=4, with_normals=True) + scenes.c3_volumes(scale=4, with_normals=False)
    reqs = scenes.c3_requests(8, seed=31, image_scale=4)
    reqs = [dataclasses.replace(q, volume=q.volume + (4 if k % 2 else 0)) for k, q in enumerate(reqs)]
    b, idx = make_batch(vols, reqs, scalar_format="f16", normal_format="sn8x3")
    try:
        b.run(count=True)
        check_against_single(b, idx, vols, reqs, "f16", "sn8x3")
    finally:
        b.close()


def test_batch_global_load_instances(hip, monkeypatch):
    """The 64-bit-load batch kernels (volumes of 4 GiB and more; forced here
    with HVR_GATHER=global) give the same bytes and counts as the
    buffer-load single-request path."""
    vols = scenes.c3_volumes(scale=4)
    reqs = scenes.c3_requests(6, seed=13, image_scale=4)
    monkeypatch.setenv("HVR_GATHER", "global")
    b, idx = make_batch(vols, reqs)
    monkeypatch.delenv("HVR_GATHER")
    try:
        b.run(count=True)
        check_against_single(b, idx, vols, reqs, "f16", "sn8x3")
    finally:
        b.close()


def test_batch_pipeline_slots(hip):
    """Two batches in flight on their own streams: each slot's results equal
    a synchronous run of the same requests."""
    from pyvr_amd.hip_renderer import BatchRenderer

    vols = scenes.c3_volumes(scale=4)
    ra = scenes.c3_requests(6, seed=21, image_scale=4)
    rb = scenes.c3_requests(5, seed=22, image_scale=4)
    want = []
    for reqs in (ra, rb):
        b, _ = make_batch(vols, reqs, pipeline=1)
        try:
            b.run()
            want.append(b.results())
        finally:
            b.close()
    b = BatchRenderer(pipeline=2)
    try:
        for q in ra:
            b.add(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
        sa = b.submit()
        for q in rb:
            b.add(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
        sb = b.submit()
        for _ in range(3):
            b.rerun(sa)
            b.rerun(sb)
        assert b.results(sa) == want[0]
        assert b.results(sb) == want[1]
    finally:
        b.close()


def test_batch_rejects_unbatchable_requests(hip):
    from pyvr_amd.hip_renderer import BatchRenderer
    from pyvr_amd.transferfunctions import ColorTransferFunction

    vols = scenes.c3_volumes(scale=8)
    q = scenes.c3_requests(1, image_scale=8)[0]
    b = BatchRenderer()
    try:
        with pytest.raises(RuntimeError):
            b.add(vols[0], q.camera, ColorTransferFunction.from_colormap("viridis"), q.otf, 32, 32,
                  lut_size=4096)
        with pytest.raises(RuntimeError):      # nothing run yet
            b.result(0)
        b.add(vols[0], q.camera, q.ctf, q.otf, 32, 32)
        with pytest.raises(RuntimeError):      # not run yet
            b.result(0)
        b.run()
        assert len(b.result(0)) == 32 * 32 * 4
    finally:
        b.close()


@pytest.mark.parametrize("kernel", ["waves2", "waves8", "nogate", "noskip"])
def test_batch_kernel_variants_equal_default(hip, kernel):
    """Every kernel variant the autotuner may pick (waves per block,
    skipping, gating) renders the same bytes as the default variant, which
    the tests above hold equal to single renders."""
    vols = scenes.c3_volumes(scale=4)
    reqs = scenes.c3_requests(10, seed=17, image_scale=4)
    out = {}
    for k in ("default", kernel):
        b, idx = make_batch(vols, reqs, kernel=k)
        try:
            b.run()
            out[k] = [b.result(i) for i in idx]
        finally:
            b.close()
    assert out[kernel] == out["default"]


def test_batch_reserved_pools(hip):
    """hvr_batch_reserve grows a batch's pools to given limits up front:
    batches within them render the same bytes as without the reservation,
    a smaller reservation afterwards changes nothing, and a reservation the
    device cannot hold fails with an error instead of faulting -- leaving no
    error behind for later calls on the thread (a failed allocation also
    sets HIP's last error, which the library's launch checks read)."""
    vols = scenes.c3_volumes(scale=4)
    reqs = scenes.c3_requests(8, seed=11, image_scale=4)
    b, idx = make_batch(vols, reqs)
    try:
        b.run()
        plain = b.results()
        bricks = max(int(np.prod([(s >> 3) + 1 for s in v.data.shape])) for v in vols)
        pixels = sum(q.width * q.height for q in reqs)
        b.reserve(16, 2 * pixels, bricks)
        b.run()
        assert b.results() == plain
        b.reserve(1, 1, 1)
        b.run()
        assert b.results() == plain
        with pytest.raises(RuntimeError):
            b.reserve(16, 1 << 40, bricks)      # 4 TiB of output pool
        b.run(count=True)
        # a new renderer (volume upload, flags and march launches) and the
        # batch itself after the failure
        check_against_single(b, idx, vols, reqs[:2], "f16", "sn8x3")
    finally:
        b.close()


def test_batch_direct_output_equals_copy(hip):
    """hvr_batch_set_readback: the march writing straight into the pinned
    pool (the default) gives the bytes of the copied-out device pool -- for
    a batch, after its pools grew (new pinned pool, new device address),
    after switching modes, and on two slots in flight."""
    from pyvr_amd.hip_renderer import BatchRenderer

    vols = scenes.c3_volumes(scale=4)
    small = scenes.c3_requests(5, seed=41, image_scale=4)
    large = scenes.c3_requests(12, seed=42, image_scale=2)
    want = {}
    for name, reqs in (("small", small), ("large", large)):
        b, _ = make_batch(vols, reqs, pipeline=1, readback="copy")
        try:
            b.run(count=True)
            want[name] = (b.results(), [b.counts(i).copy() for i in range(len(reqs))])
        finally:
            b.close()
    b = BatchRenderer(pipeline=2, readback="direct")
    try:
        for name, reqs in (("small", small), ("large", large), ("small", small)):
            b.clear()
            for q in reqs:
                b.add(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
            b.run(count=True)
            assert b.results() == want[name][0], name
            assert all(np.array_equal(b.counts(i), c) for i, c in enumerate(want[name][1])), name
        b.set_readback("copy")
        b.run()
        assert b.results() == want["small"][0]
        b.set_readback("direct")
        b.clear()
        for q in small:
            b.add(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
        sa = b.submit()
        for q in large:
            b.add(vols[q.volume], q.camera, q.ctf, q.otf, q.width, q.height, q.config, q.light, q.tau)
        sb = b.submit()
        assert b.results(sa) == want["small"][0]
        assert b.results(sb) == want["large"][0]
    finally:
        b.close()
