"""The no-allocation entry point (hvr_render_async_to_noalloc), called
through the C ABI: its frames equal, byte for byte, those of
hvr_render_async_to for the same context state -- on two calls in a row, the
second with a smaller volume and image, and on two frames queued back to back
on one stream (the static scratch array is reused in stream order).  Sizes
beyond the array's limits are refused with a clear error.  Collected last
(its name), so that a failure here cannot hide the other GPU tests."""

import ctypes

import numpy as np
import pytest

import scenes

pytestmark = pytest.mark.gpu


def _fn(L):
    return L.hvr_render_async_to_noalloc


class _Frame:
    """A context of the request's image size with the request's volume,
    LUT and parameters bound (everything through the C ABI)."""

    def __init__(self, L, vol, req, sf, nf):
        from pyvr_amd.hip_renderer import _lib
        from pyvr_amd.hip_renderer.batch import rgba_lut
        from pyvr_amd.hip_renderer.manager import _DeviceVolume

        self.L, self.W, self.H = L, req.width, req.height
        self.ctx = ctypes.c_void_p()
        _lib.check(L.hvr_create(0, self.W, self.H, ctypes.byref(self.ctx)), "hvr_create")
        self.dv = _DeviceVolume(L, self.ctx, vol.data.shape, sf, nf if vol.has_normals else "none")
        self.dv.upload_scalar(vol.data)
        if vol.has_normals:
            self.dv.upload_normals(vol.normals)
        nrole = _lib.HVR_ROLE_NORMAL if vol.has_normals else _lib.HVR_ROLE_SCALAR
        _lib.check(L.hvr_bind(self.ctx, self.dv.handle, _lib.HVR_ROLE_SCALAR, self.dv.handle, nrole), "bind")
        lut = rgba_lut(req.ctf, req.otf)
        _lib.check(L.hvr_set_lut(self.ctx, lut.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), lut.shape[0]),
                   "lut")
        self.vol, self.req = vol, req
        self.set_camera(req.camera)
        self.bufs = []

    def set_camera(self, camera):
        from pyvr_amd.hip_renderer import _lib
        from pyvr_amd.hip_renderer.batch import request_params

        q = self.req
        self.prm = request_params(self.vol, camera, q.config, q.light, q.width, q.height, q.tau)
        _lib.check(self.L.hvr_set_params(self.ctx, ctypes.byref(self.prm)), "params")

    def buffer(self):
        from pyvr_amd.hip_renderer import _lib

        p = ctypes.c_void_p()
        _lib.check(self.L.hvr_device_alloc(self.ctx, self.W * self.H * 4, ctypes.byref(p)), "alloc")
        self.bufs.append(p)
        return p

    def host(self, p) -> bytes:
        from pyvr_amd.hip_renderer import _lib

        out = np.empty(self.W * self.H * 4, np.uint8)
        _lib.check(self.L.hvr_synchronize(self.ctx), "sync")
        _lib.check(self.L.hvr_copy_to_host(self.ctx, out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes), "copy")
        return out.tobytes()

    def close(self):
        for p in self.bufs:
            self.L.hvr_device_free(self.ctx, p)
        self.dv.destroy()
        self.L.hvr_destroy(self.ctx)


def _compare(L, fr: _Frame, label: str):
    from pyvr_amd.hip_renderer import _lib

    a, b = fr.buffer(), fr.buffer()
    _lib.check(L.hvr_render_async_to(fr.ctx, a), "hvr_render_async_to")
    want = fr.host(a)
    print(f"[noalloc] {label}: {fr.W}x{fr.H}, volume {fr.vol.data.shape}", flush=True)
    _lib.check(_fn(L)(fr.ctx, b), "hvr_render_async_to_noalloc")
    got = fr.host(b)
    ga, wa = np.frombuffer(got, np.uint8), np.frombuffer(want, np.uint8)
    assert got == want, f"{label}: {int((ga != wa).sum())} bytes differ"
    assert int(wa.reshape(-1, 4)[:, 3].astype(bool).sum()) > 0, f"{label}: empty frame"


def test_noalloc_equals_render_async_to_two_calls(hip):
    """Two calls in a row on one GPU, the second with a smaller volume (no
    normals, non-cubic, UNORM8) and a smaller image."""
    L = hip
    vols = scenes.c3_volumes(scale=2)                       # 32^3 .. 128^3, 48x80x112
    small = scenes.c3_volumes(scale=4, with_normals=False)[3]  # 24x40x56
    reqs = scenes.c3_requests(2, seed=61, image_scale=2)
    big = scenes.Request(2, reqs[0].camera, reqs[0].ctf, reqs[0].otf, 384, 256, reqs[0].config, reqs[0].light,
                         0.95)
    little = scenes.Request(0, reqs[1].camera, reqs[1].ctf, reqs[1].otf, 160, 120, reqs[1].config,
                            reqs[1].light, 0.99)
    f1 = _Frame(L, vols[2], big, "f16", "sn8x3")
    try:
        _compare(L, f1, "first call, 128^3 f16+sn8x3")
    finally:
        f1.close()
    f2 = _Frame(L, small, little, "u8", "sn8x3")
    try:
        _compare(L, f2, "second call, 24x40x56 u8, no normals")
    finally:
        f2.close()


def test_noalloc_frames_back_to_back_on_one_stream(hip):
    """Two frames of different cameras queued without a wait in between
    (the scratch array is reused in stream order), then the first frame's
    camera again after a larger brick count was staged: each equals
    hvr_render_async_to's frame."""
    from pyvr_amd.hip_renderer import _lib

    L = hip
    vol = scenes.c3_volumes(scale=2)[3]                     # 48x80x112
    reqs = scenes.c3_requests(2, seed=63, image_scale=4)
    q = scenes.Request(3, reqs[0].camera, reqs[0].ctf, reqs[0].otf, 200, 150, reqs[0].config, reqs[0].light,
                       0.99)
    fr = _Frame(L, vol, q, "f16", "f16x3")
    try:
        want = []
        for cam in (q.camera, reqs[1].camera):
            fr.set_camera(cam)
            a = fr.buffer()
            _lib.check(L.hvr_render_async_to(fr.ctx, a), "render")
            want.append(fr.host(a))
        got = []
        for cam in (q.camera, reqs[1].camera):
            fr.set_camera(cam)
            b = fr.buffer()
            _lib.check(_fn(L)(fr.ctx, b), "noalloc")
            got.append(b)
        assert fr.host(got[0]) == want[0]
        assert fr.host(got[1]) == want[1]
    finally:
        fr.close()


def test_noalloc_rejects_what_the_array_cannot_hold(hip):
    from pyvr_amd.hip_renderer import _lib

    L = hip
    vol = scenes.c3_volumes(scale=8)[0]
    q = scenes.c3_requests(1, seed=65, image_scale=8)[0]
    fr = _Frame(L, vol, q, "f16", "sn8x3")
    try:
        b = fr.buffer()
        big = np.zeros((4096, 4), np.float32)
        big[:, 3] = 0.5
        _lib.check(L.hvr_set_lut(fr.ctx, big.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 4096), "lut")
        rc = _fn(L)(fr.ctx, b)
        assert rc == -1 and b"2048" in L.hvr_last_error()
        _lib.check(L.hvr_bind(fr.ctx, None, _lib.HVR_ROLE_SCALAR, None, _lib.HVR_ROLE_SCALAR), "unbind")
        rc = _fn(L)(fr.ctx, b)
        assert rc == -4 and b"scalar" in L.hvr_last_error()
        assert _fn(L)(fr.ctx, None) == -1
    finally:
        fr.close()
