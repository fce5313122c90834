"""HIP renderer vs the FP64 oracle on the benchmark scenes (GPU).

Bar: per RGBA8 channel |kernel - oracle| <= 1 on >= 99.9 % of pixels, with
the maximum error printed.  Sample counts are compared with the oracle's
(FP32-arithmetic mode, the same stepping as the kernel).
"""

import numpy as np
import pytest

import scenes
from parity import assert_parity, compare, oracle_render, render_hip

pytestmark = pytest.mark.gpu


def test_c0_small(hip):
    sc = scenes.c0(size=48, image=96)
    img, cnt = render_hip(sc, counts=True)
    ref = oracle_render(sc)
    assert_parity(img, ref["rgba8"], label=sc.name)
    ref32 = oracle_render(sc, mode="f32")
    # same FP32 stepping as the kernel: sample counts agree except where a
    # face test rounds differently
    diff = np.abs(cnt.astype(np.int64) - ref32["counts"].astype(np.int64))
    assert (diff == 0).mean() > 0.999, diff.sum()


def test_c0_full(hip):
    sc = scenes.c0()
    img, cnt = render_hip(sc, counts=True)
    ref = oracle_render(sc)
    st = assert_parity(img, ref["rgba8"], label=sc.name)
    print(f"[counts] kernel {int(cnt.sum())} oracle64 {int(ref['counts'].sum())}", flush=True)
    assert abs(int(cnt.sum()) - int(ref["counts"].sum())) < 1e-4 * ref["counts"].sum()


def test_c1_small(hip):
    sc = scenes.c1(size=96, image=128)
    img = render_hip(sc)
    ref = oracle_render(sc)
    assert_parity(img, ref["rgba8"], label=sc.name)


@pytest.mark.slow
def test_c1_full(hip):
    sc = scenes.c1()
    img, cnt = render_hip(sc, counts=True)
    ref = oracle_render(sc)
    assert_parity(img, ref["rgba8"], label=sc.name)
    print(f"[counts] kernel {int(cnt.sum())} oracle64 {int(ref['counts'].sum())}", flush=True)


def test_deterministic(hip):
    sc = scenes.c0(size=32, image=64)
    a = render_hip(sc)
    b = render_hip(sc)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", ["on", "off"])
def test_global_and_buffer_gather_identical(hip, monkeypatch, mode):
    """The 64-bit global-load instance (used for volumes of 4 GiB and more)
    and the buffer-load instance give the same bytes and counts."""
    from pyvr_amd.hip_renderer import _lib

    sc = scenes.c2(size=80, image=96)
    m = {"on": _lib.SKIP_ON, "off": _lib.SKIP_OFF}[mode]
    out = []
    for gather in ("", "global"):
        monkeypatch.setenv("HVR_GATHER", gather)
        r = _prepared(sc)
        try:
            out.append(_render_raw(r, sc.lut(), m)[:2])
        finally:
            r.close()
    (a, ca), (b, cb) = out
    assert np.array_equal(a, b) and np.array_equal(ca, cb)


@pytest.mark.parametrize("lb", [2, 4])
def test_brick_layouts_identical(hip, lb):
    """Brick edge 4 (generic path), 8 (default) and 16 give the same bytes."""
    sc = scenes.c1(size=40, image=64)
    base = render_hip(sc, brick_log2=3)
    img = render_hip(sc, brick_log2=lb)
    assert np.array_equal(base, img)


def _render_raw(r, lut, mode):
    """One frame and its sample counts through the C ABI with LUT `lut` and
    skip mode `mode` (camera, volume and parameters as last set on `r`)."""
    import ctypes
    from pyvr_amd.hip_renderer import _lib

    L, ctx = _lib.lib(), r.gl_manager.ctx
    lut = np.ascontiguousarray(lut, np.float32)
    _lib.check(L.hvr_set_empty_skip(ctx, mode), "skip")
    _lib.check(L.hvr_set_lut(ctx, lut.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(lut)), "lut")
    h, w = r.height, r.width
    img = np.zeros((h, w, 4), np.uint8)
    cnt = np.zeros((h, w), np.uint32)
    _lib.check(L.hvr_render(ctx, img.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))), "render")
    _lib.check(L.hvr_sample_counts(ctx, cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))), "counts")
    clear, total, used = ctypes.c_ulonglong(), ctypes.c_ulonglong(), ctypes.c_int()
    _lib.check(L.hvr_skip_info(ctx, ctypes.byref(clear), ctypes.byref(total), ctypes.byref(used)), "info")
    return img, cnt, (clear.value, total.value, used.value)


def _prepared(sc):
    from pyvr_amd.hip_renderer import VolumeRenderer

    r = VolumeRenderer(sc.width, sc.height, config=sc.config, light=sc.light,
                       scalar_format="f16", normal_format="sn8x3")
    r.load_volume(sc.volume())
    r.set_camera(sc.camera)
    r.set_transfer_functions(sc.ctf, sc.otf, sc.lut_size)
    r.set_opacity_threshold(sc.tau)
    r.render()
    return r


@pytest.mark.parametrize("which", ["c0", "c1", "c2"])
def test_transparent_brick_skip_is_exact(hip, which):
    """Stepping over transparent bricks changes neither the image nor the
    counts, and does happen on these scenes (empty space is transparent)."""
    from pyvr_amd.hip_renderer import _lib

    sc = {"c0": lambda: scenes.c0(size=64, image=96), "c1": lambda: scenes.c1(size=96, image=128),
          "c2": lambda: scenes.c2(size=96, image=128)}[which]()
    r = _prepared(sc)
    try:
        a, ca, info_on = _render_raw(r, sc.lut(), _lib.SKIP_ON)
        b, cb, info_off = _render_raw(r, sc.lut(), _lib.SKIP_OFF)
    finally:
        r.close()
    print(f"[skip] {which}: transparent bricks {info_on[0]}/{info_on[1]}", flush=True)
    assert info_on[2] == 1 and info_on[0] > 0 and info_off[2] == 0
    assert np.array_equal(a, b) and np.array_equal(ca, cb)
    assert ca.sum() > 0 and a[..., 3].max() > 0


def test_transparent_brick_skip_with_nonfinite_lut(hip):
    """Alpha-0 LUT entries next to a NaN or infinite entry are not transparent
    (the filtered lookup reads the neighbour even at weight 0): skipping
    stays byte-identical to marching every sample."""
    from pyvr_amd.hip_renderer import _lib

    sc = scenes.c0(size=64, image=96)
    base = sc.lut()
    r = _prepared(sc)
    try:
        for entry, value in [(1, np.inf), (1, np.nan), (40, -np.inf)]:
            lut = base.copy()
            lut[:, 3] = np.where(np.arange(len(lut)) < 60, 0.0, lut[:, 3])
            lut[entry, entry % 3] = value
            a, ca, info_on = _render_raw(r, lut, _lib.SKIP_ON)
            b, cb, _ = _render_raw(r, lut, _lib.SKIP_OFF)
            print(f"[skip] entry {entry} = {value}: transparent bricks {info_on[0]}/{info_on[1]}", flush=True)
            assert np.array_equal(a, b) and np.array_equal(ca, cb), (entry, value)
    finally:
        r.close()


@pytest.mark.parametrize("name", ["hydrogen", "fuel"])
def test_vti_example_files(hip, name):
    """Volumes read from the reference's example .vti files
    (load_vtk_volume) render to the oracle's bar."""
    from pathlib import Path

    from pyvr_amd.dataloaders import load_vtk_volume

    v = load_vtk_volume(Path(__file__).parent / "golden" / "example_data" / f"{name}.vti")
    base = scenes.c0(size=8, image=128)
    sc = scenes.Scene(name=f"vti_{name}", data=v.data, normals=v.normals, camera=base.camera,
                      light=base.light, ctf=base.ctf, otf=scenes.seeded_otf(6, seed=5), config=base.config,
                      width=128, height=128, tau=0.99, bmin=v.min_bounds, bmax=v.max_bounds)
    img = render_hip(sc)
    ref = oracle_render(sc)
    assert_parity(img, ref["rgba8"], label=sc.name)
    assert ref["rgba8"][..., 3].max() > 0


@pytest.mark.parametrize("wh", [(200, 170), (168, 136)])
def test_non_square_ragged_frames(hip, wh):
    """Frames whose size is not a multiple of the 16-pixel block, with block
    rows both a multiple of 8 and not (the march deals whole groups of 8
    block rows to the XCDs and keeps the rest in plain order): every pixel
    is rendered, to the oracle's bar, with the oracle's sample counts."""
    import dataclasses

    w, h = wh
    sc = dataclasses.replace(scenes.c0(size=48, image=96), width=w, height=h)
    img, cnt = render_hip(sc, counts=True)
    ref = oracle_render(sc, mode="f32")
    assert_parity(img, oracle_render(sc)["rgba8"], label=f"{w}x{h}")
    diff = np.abs(cnt.astype(np.int64) - ref["counts"].astype(np.int64))
    assert (diff == 0).mean() > 0.999 and ref["counts"].sum() > 0


def test_pinned_read_equals_copy_read(hip):
    """render() fills its bytes through hvr_read (through pinned memory, in
    chunks of at least 1 MiB); they equal hvr_read_pinned's buffer and
    hvr_render's copy, frame after frame, at sizes taking one copy and
    two or three chunks (the last one ragged)."""
    import ctypes

    from pyvr_amd.hip_renderer import VolumeRenderer, _lib

    for w, h in ((64, 64), (200, 170), (13, 7), (1024, 600), (1000, 1000)):
        sc = scenes.c0(size=32, image=64)
        r = VolumeRenderer(w, h, config=sc.config, light=sc.light)
        try:
            r.load_volume(sc.volume())
            r.set_transfer_functions(sc.ctf, sc.otf)
            L, ctx = _lib.lib(), r.gl_manager.ctx
            for az in (0.3, 1.1, 2.0):
                r.set_camera(scenes.standard_camera(azimuth=az))
                a = r.render()
                ptr = ctypes.POINTER(ctypes.c_uint8)()
                _lib.check(L.hvr_read_pinned(ctx, ctypes.byref(ptr)), "hvr_read_pinned")
                pinned = ctypes.string_at(ptr, w * h * 4)
                out = (ctypes.c_uint8 * (w * h * 4))()
                _lib.check(L.hvr_render(ctx, out), "hvr_render")
                assert isinstance(a, bytes) and len(a) == w * h * 4
                assert a == pinned == bytes(out) and max(a) > 0, (w, h, az)
        finally:
            r.close()


def test_shaded_sample_counts(hip):
    """hvr_shaded_sample_counts: the samples that load texels.  Without
    skipping, every counted sample does; with skipping, those in transparent
    bricks do not, so per pixel 0 <= shaded <= counted, with fewer in total
    on a scene with empty space; the march itself is unchanged."""
    import ctypes

    from pyvr_amd.hip_renderer import _lib

    sc = scenes.c2(size=96, image=128)
    r = _prepared(sc)
    L, ctx = _lib.lib(), r.gl_manager.ctx
    try:
        out = {}
        for mode in (_lib.SKIP_OFF, _lib.SKIP_ON):
            img, cnt, _ = _render_raw(r, sc.lut(), mode)
            sh = np.zeros_like(cnt)
            _lib.check(L.hvr_shaded_sample_counts(ctx, sh.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))), "shaded")
            out[mode] = (img, cnt, sh)
    finally:
        r.close()
    img0, cnt0, sh0 = out[_lib.SKIP_OFF]
    img1, cnt1, sh1 = out[_lib.SKIP_ON]
    print(f"[shaded] samples {int(cnt1.sum())}, shaded {int(sh1.sum())} with skipping", flush=True)
    assert np.array_equal(sh0, cnt0) and np.array_equal(cnt0, cnt1) and np.array_equal(img0, img1)
    assert (sh1 <= cnt1).all() and 0 < sh1.sum() < cnt1.sum()


def test_touched_bricks(hip):
    """hvr_touched_bricks: the bricks a frame's samples load texels from.
    Skipping can only remove bricks (those it steps over); a LUT that is
    transparent everywhere leaves none with skipping, and without skipping
    at least as many as the opaque LUT (no ray terminates early); the count
    never exceeds the volume's bricks."""
    import ctypes

    from pyvr_amd.hip_renderer import _lib

    sc = scenes.c2(size=96, image=128)
    r = _prepared(sc)
    L, ctx = _lib.lib(), r.gl_manager.ctx

    def touched(lut, mode):
        _render_raw(r, lut, mode)
        t, n, b = ctypes.c_ulonglong(), ctypes.c_ulonglong(), ctypes.c_ulonglong()
        _lib.check(L.hvr_touched_bricks(ctx, ctypes.byref(t), ctypes.byref(n), ctypes.byref(b)), "touched")
        return t.value, n.value, b.value

    try:
        clear = sc.lut().copy()
        clear[:, 3] = 0.0
        on, nb, bb = touched(sc.lut(), _lib.SKIP_ON)
        off, _, _ = touched(sc.lut(), _lib.SKIP_OFF)
        clear_on, _, _ = touched(clear, _lib.SKIP_ON)
        clear_off, _, _ = touched(clear, _lib.SKIP_OFF)
    finally:
        r.close()
    print(f"[touched] {on} of {nb} bricks ({bb} B each) with skipping, {off} without; transparent LUT "
          f"{clear_on} / {clear_off}", flush=True)
    assert 0 < on <= off <= nb and bb > 0
    assert clear_on == 0 and off <= clear_off <= nb


def test_nan_densities_match_oracle(hip):
    """NaN voxels in an FP32 volume: a sample whose footprint holds one
    interpolates to NaN and reads LUT entry 0 (the clamp's max(NaN, 0) = 0);
    its normal is NaN and lights nothing.  With an opacity that is non-zero
    at density 0 such samples are visible.  Single-frame and batched renders
    against the FP32-mode oracle; the batch must equal the single render."""
    from pyvr_amd.datasets import compute_normal_volume
    from pyvr_amd.hip_renderer import BatchRenderer
    from pyvr_amd.transferfunctions import OpacityTransferFunction

    base = scenes.c0(size=48, image=96)
    rng = np.random.default_rng(7)
    data = base.data.astype(np.float32).copy()
    data[rng.random(data.shape) < 0.01] = np.nan
    with np.errstate(invalid="ignore"):
        normals = compute_normal_volume(data)
    sc = scenes.Scene(name="c0_nan", data=data, normals=normals, camera=base.camera, light=base.light,
                      ctf=base.ctf, otf=OpacityTransferFunction.linear(0.05, 0.9), config=base.config,
                      width=base.width, height=base.height, tau=base.tau)
    img, cnt = render_hip(sc, counts=True)
    r32 = oracle_render(sc, mode="f32")
    npix = sc.width * sc.height
    st = compare(img, r32["rgba8"])
    diff = int((cnt.astype(np.int64) != r32["counts"].astype(np.int64)).sum())
    print(f"[nan] vs FP32 oracle {st}; counts differ {diff}, samples {int(cnt.sum())}", flush=True)
    assert img[..., 3].max() > 0
    assert st["n_gt1"] <= max(1, int(1e-3 * npix)), st
    assert diff <= max(1, int(1e-3 * npix)), diff
    b = BatchRenderer(scalar_format="f32", normal_format="f32x3")
    try:
        i = b.add(sc.volume(), sc.camera, sc.ctf, sc.otf, sc.width, sc.height, sc.config, sc.light, sc.tau)
        b.run(count=True)
        got = np.frombuffer(b.result(i), np.uint8).reshape(sc.height, sc.width, 4)
        assert np.array_equal(got, img) and np.array_equal(b.counts(i), cnt)
    finally:
        b.close()
