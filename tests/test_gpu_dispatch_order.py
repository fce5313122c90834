"""Cost-ordered dispatch of single frames (hvr_set_dispatch_order): ray setup
files every 16x16 ray block under its rays' largest iteration bound and the
march starts the longest blocks first.  Only the order in which blocks run
changes, so images and sample counts must be byte-identical to row order,
on frames whose block rows are and are not a multiple of 8."""

import ctypes

import numpy as np
import pytest

import scenes

pytestmark = pytest.mark.gpu


def _frame(r, camera, mode):
    from pyvr_amd.hip_renderer import _lib

    L, ctx = _lib.lib(), r.gl_manager.ctx
    _lib.check(L.hvr_set_dispatch_order(ctx, mode), "order")
    r.set_camera(camera)
    img = np.frombuffer(r.render(), np.uint8).reshape(r.height, r.width, 4).copy()
    used = ctypes.c_int()
    _lib.check(L.hvr_dispatch_order_info(ctx, ctypes.byref(used)), "order info")
    cnt = r.sample_counts()
    _lib.check(L.hvr_dispatch_order_info(ctx, ctypes.byref(used)), "order info")
    return img, cnt, used.value


@pytest.mark.parametrize("wh", [(1024, 1024), (1000, 776), (200, 136)])
def test_cost_ordered_frames_equal_row_order(hip, wh):
    from pyvr_amd.hip_renderer import VolumeRenderer, _lib

    w, h = wh
    sc = scenes.c2(size=128, image=64)
    cams = scenes.c2_cameras(3)
    r = VolumeRenderer(w, h, config=sc.config, light=sc.light, scalar_format="f16", normal_format="f16x3")
    try:
        r.load_volume(sc.volume())
        r.set_transfer_functions(sc.ctf, sc.otf, sc.lut_size)
        r.set_opacity_threshold(sc.tau)
        for k in (1, 2):
            # another camera first, so that a block the ordered frame missed
            # would keep the other frame's pixels
            _frame(r, cams[0], _lib.ORDER_OFF)
            a, ca, used_on = _frame(r, cams[k], _lib.ORDER_ON)
            _frame(r, cams[0], _lib.ORDER_OFF)
            b, cb, used_off = _frame(r, cams[k], _lib.ORDER_OFF)
            print(f"[order] {w}x{h} camera {k}: ordered {used_on}, row order {used_off}, "
                  f"samples {int(cb.sum())}", flush=True)
            assert used_on == (1 if h >= 128 else 0) and used_off == 0
            assert np.array_equal(a, b) and np.array_equal(ca, cb)
            assert cb.sum() > 0 and b[..., 3].max() > 0
    finally:
        r.close()


def test_auto_orders_only_frames_larger_than_one_round(hip):
    """HVR_ORDER_AUTO (the default) orders 1024^2 frames (4096 blocks, more
    than the 8 per compute unit the device keeps resident) and leaves 256^2
    frames (256 blocks) in row order."""
    from pyvr_amd.hip_renderer import VolumeRenderer, _lib

    sc = scenes.c0(size=48, image=64)
    for size, expect in ((1024, 1), (256, 0)):
        r = VolumeRenderer(size, size, config=sc.config, light=sc.light)
        try:
            r.load_volume(sc.volume())
            r.set_transfer_functions(sc.ctf, sc.otf, sc.lut_size)
            _, _, used = _frame(r, sc.camera, _lib.ORDER_AUTO)
        finally:
            r.close()
        assert used == expect, (size, used)
