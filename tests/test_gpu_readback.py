"""Where a single frame goes (hvr_set_readback): the device framebuffer
copied out afterwards (HVR_READBACK_COPY), or the pinned host buffer the
march writes straight into (HVR_READBACK_DIRECT, the default).  The bytes
a frame reads back as, those hvr_read_pinned exposes and the sample counts
are the same either way, across a framebuffer that grows (its pinned
buffer reallocated) and shrinks."""

import numpy as np
import pytest

import scenes

pytestmark = pytest.mark.gpu


def test_direct_readback_equals_copy(hip):
    from pyvr_amd.hip_renderer import _lib
    from pyvr_amd.hip_renderer.plain import PlainRenderer

    sc = scenes.c2(size=96, image=64)
    vol = sc.volume()
    cams = scenes.c2_cameras(3)
    pr = PlainRenderer(scalar_format="f16", normal_format="f16x3")
    L = _lib.lib()
    try:
        for w, h in ((256, 192), (640, 480), (200, 136)):
            for k, cam in enumerate(cams):
                out = {}
                for mode in (_lib.READBACK_COPY, _lib.READBACK_DIRECT):
                    _lib.check(L.hvr_set_readback(pr.ctx, mode), "readback")
                    pr.prepare(vol, cam, sc.ctf, sc.otf, w, h, sc.config, sc.light, sc.tau, sc.lut_size)
                    img = pr.read()
                    pinned = bytes(pr.read_pinned())
                    out[mode] = (img, pinned, pr.sample_counts())
                a, b = out[_lib.READBACK_COPY], out[_lib.READBACK_DIRECT]
                print(f"[readback] {w}x{h} camera {k}: alpha max {max(b[0][3::4])}", flush=True)
                assert len(b[0]) == w * h * 4 and a[0] == b[0] and a[1] == b[1] == b[0]
                assert np.array_equal(a[2], b[2]) and a[2].sum() > 0
    finally:
        pr.close()
