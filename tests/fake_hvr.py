"""A Python stand-in for libhipvr.so used by the CPU tests of the host logic
(texture units, sampler resolution, uniforms, resource release), in the way
the reference tests its renderer against a mocked ModernGL.  It records the
calls and renders nothing."""

import ctypes


def _set(ref, value):
    ref._obj.value = value


def _hv(h):
    if h is None:
        return None
    return h.value if hasattr(h, "value") else h


class FakeHvr:
    def __init__(self):
        self.next = 1000
        self.contexts = {}
        self.volumes = {}
        self.destroyed = []
        self.binds = []
        self.params = []
        self.luts = []
        self.frames = 0

    def _new(self):
        self.next += 1
        return self.next

    def hvr_abi_version(self):
        return 1

    def hvr_last_error(self):
        return b""

    def hvr_device_count(self, ref):
        _set(ref, 1)
        return 0

    def hvr_create(self, device, w, h, ref):
        k = self._new()
        self.contexts[k] = (w, h)
        _set(ref, k)
        return 0

    def hvr_destroy(self, ctx):
        self.contexts.pop(_hv(ctx), None)
        return 0

    def hvr_volume_create(self, ctx, s0, s1, s2, sf, nf, ref):
        k = self._new()
        self.volumes[k] = {"shape": (s0, s1, s2), "sf": sf, "nf": nf, "scalar": False, "normals": False}
        _set(ref, k)
        return 0

    def hvr_volume_upload_scalar(self, v, ptr):
        self.volumes[_hv(v)]["scalar"] = True
        return 0

    def hvr_volume_upload_normals(self, v, ptr):
        self.volumes[_hv(v)]["normals"] = True
        return 0

    def hvr_volume_bytes(self, v, ref):
        s = self.volumes[_hv(v)]["shape"]
        _set(ref, s[0] * s[1] * s[2] * 4)
        return 0

    def hvr_volume_destroy(self, v):
        self.destroyed.append(_hv(v))
        self.volumes.pop(_hv(v), None)
        return 0

    def hvr_set_lut(self, ctx, ptr, n):
        self.luts.append(n)
        return 0

    def hvr_set_params(self, ctx, ref):
        p = ref._obj
        self.params.append({f: (list(getattr(p, f)) if hasattr(getattr(p, f), "__len__") else getattr(p, f))
                            for f, _ in p._fields_})
        return 0

    def hvr_bind(self, ctx, s, srole, n, nrole):
        self.binds.append((_hv(s), srole, _hv(n), nrole))
        return 0

    def hvr_render_async(self, ctx):
        self.frames += 1
        return 0

    def hvr_read(self, ctx, buf):
        w, h = self.contexts[_hv(ctx)]
        # row y (bottom-up) is filled with the byte value y
        data = b"".join(bytes([y % 256]) * (4 * w) for y in range(h))
        ctypes.memmove(buf, data, len(data))
        return 0

    def hvr_read_pinned(self, ctx, ref):
        w, h = self.contexts[_hv(ctx)]
        buf = (ctypes.c_uint8 * (4 * w * h))()
        self.hvr_read(ctx, buf)
        self._pinned = buf                       # kept alive like the library's buffer
        ref._obj.contents = ctypes.c_uint8.from_buffer(buf)
        return 0

    def hvr_sample_counts(self, ctx, ptr):
        return 0

    def hvr_last_kernel_ms(self, ctx, a, b):
        _set(a, 1.0)
        _set(b, 0.1)
        return 0

    def hvr_set_brick_log2(self, ctx, lb):
        return 0
