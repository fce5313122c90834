"""ctypes binding of ``libhipvr.so`` (declarations in ``include/hipvr.h``).

The library is built in-tree by ``pyvr_amd.build`` (hipcc, gfx950).  There is
no fallback: if the library is missing or cannot be loaded, every renderer
call raises ``RuntimeError``.
"""

import ctypes
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# HVR_LIBRARY names another build of the same ABI (for comparing builds)
LIB_PATH = os.environ.get("HVR_LIBRARY") or os.path.join(_PKG, "libhipvr.so")

HVR_OK = 0
HVR_SCALAR = {"none": 0, "f32": 1, "f16": 2, "u8": 3}
HVR_NORMAL = {"none": 0, "f32x3": 1, "sn8x3": 2, "f16x3": 3}


def check_formats(scalar_format: str, normal_format: str) -> None:
    """``ValueError`` unless the library stores this pair (canonical names;
    "none" for an absent field): FP16x3 normals pair with FP16 scalars or
    none, the other normal formats with every scalar format."""
    sf, nf = scalar_format, normal_format
    if (sf not in HVR_SCALAR or nf not in HVR_NORMAL or (sf == "none" and nf == "none")
            or (nf == "f16x3" and sf not in ("none", "f16"))):
        raise ValueError(f"unsupported formats scalar={sf!r} normal={nf!r}"
                         + (" (f16x3 normals pair with f16 scalars)" if nf == "f16x3" else ""))
HVR_ROLE_SCALAR = 0
HVR_ROLE_NORMAL = 1

# exported symbols, as declared in include/hipvr.h
EXPORTS = (
    "hvr_abi_version", "hvr_last_error", "hvr_device_count", "hvr_create", "hvr_destroy",
    "hvr_volume_create", "hvr_volume_upload_scalar", "hvr_volume_upload_normals",
    "hvr_volume_bytes", "hvr_volume_destroy", "hvr_set_lut", "hvr_set_params", "hvr_bind",
    "hvr_render", "hvr_render_async", "hvr_read", "hvr_synchronize", "hvr_framebuffer",
    "hvr_sample_counts", "hvr_last_kernel_ms", "hvr_set_brick_log2",
    "hvr_render_async_to", "hvr_device_alloc", "hvr_device_free", "hvr_copy_to_host",
    "hvr_timing_begin", "hvr_timing_end", "hvr_device_copy", "hvr_set_empty_skip",
    "hvr_skip_info", "hvr_batch_create", "hvr_batch_destroy", "hvr_batch_clear", "hvr_batch_set_tile_shape",
    "hvr_batch_add", "hvr_batch_run_async", "hvr_batch_wait", "hvr_batch_run", "hvr_batch_output",
    "hvr_batch_counts", "hvr_batch_last_ms", "hvr_volume_upload_scalar_normals", "hvr_volume_generate_blobs",
    "hvr_batch_query", "hvr_device_memory", "hvr_read_pinned", "hvr_volume_download",
    "hvr_resize", "hvr_debug_inject_fault", "hvr_render_async_to_noalloc",
    "hvr_batch_set_kernel", "hvr_set_dispatch_order", "hvr_dispatch_order_info", "hvr_set_readback",
    "hvr_shaded_sample_counts", "hvr_touched_bricks", "hvr_batch_reserve", "hvr_batch_set_readback",
)

# batch ray-tile shapes (hvr_batch_set_tile_shape)
TILE_8X8, TILE_16X4 = 0, 1

# transparent-brick skipping modes (hvr_set_empty_skip)
SKIP_OFF, SKIP_ON, SKIP_AUTO = 0, 1, 2

# cost-ordered dispatch of single frames (hvr_set_dispatch_order)
ORDER_OFF, ORDER_ON, ORDER_AUTO = 0, 1, 2

# where single frames go (hvr_set_readback)
READBACK_COPY, READBACK_DIRECT = 0, 1


class HvrParams(ctypes.Structure):
    _fields_ = [
        ("view", ctypes.c_float * 16),
        ("proj", ctypes.c_float * 16),
        ("cam_pos", ctypes.c_float * 3),
        ("bmin", ctypes.c_float * 3),
        ("bmax", ctypes.c_float * 3),
        ("step", ctypes.c_float),
        ("ref_step", ctypes.c_float),
        ("max_steps", ctypes.c_int32),
        ("ambient", ctypes.c_float),
        ("diffuse", ctypes.c_float),
        ("light_pos", ctypes.c_float * 3),
        ("light_target", ctypes.c_float * 3),
        ("et_threshold", ctypes.c_float),
        ("camera_set", ctypes.c_int32),
        ("clear_rgba", ctypes.c_uint8 * 4),
        ("blend", ctypes.c_int32),
    ]


_lib = None
_load_error = None


def _declare(L):
    vp, i, u8p = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint8)
    fp = ctypes.POINTER(ctypes.c_float)
    sig = {
        "hvr_abi_version": ([], i),
        "hvr_last_error": ([], ctypes.c_char_p),
        "hvr_device_count": ([ctypes.POINTER(i)], i),
        "hvr_create": ([i, i, i, ctypes.POINTER(vp)], i),
        "hvr_destroy": ([vp], i),
        "hvr_resize": ([vp, i, i], i),
        "hvr_debug_inject_fault": ([ctypes.c_char_p, i, i], i),
        "hvr_volume_create": ([vp, i, i, i, i, i, ctypes.POINTER(vp)], i),
        "hvr_volume_upload_scalar": ([vp, fp], i),
        "hvr_volume_upload_normals": ([vp, fp], i),
        "hvr_volume_bytes": ([vp, ctypes.POINTER(ctypes.c_uint64)], i),
        "hvr_volume_download": ([vp, fp, fp], i),
        "hvr_volume_destroy": ([vp], i),
        "hvr_set_lut": ([vp, fp, i], i),
        "hvr_set_params": ([vp, ctypes.POINTER(HvrParams)], i),
        "hvr_bind": ([vp, vp, i, vp, i], i),
        "hvr_render": ([vp, u8p], i),
        "hvr_render_async": ([vp], i),
        "hvr_read": ([vp, u8p], i),
        "hvr_read_pinned": ([vp, ctypes.POINTER(u8p)], i),
        "hvr_synchronize": ([vp], i),
        "hvr_framebuffer": ([vp, ctypes.POINTER(vp)], i),
        "hvr_sample_counts": ([vp, ctypes.POINTER(ctypes.c_uint32)], i),
        "hvr_shaded_sample_counts": ([vp, ctypes.POINTER(ctypes.c_uint32)], i),
        "hvr_touched_bricks": ([vp, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong),
                                ctypes.POINTER(ctypes.c_ulonglong)], i),
        "hvr_last_kernel_ms": ([vp, fp, fp], i),
        "hvr_set_brick_log2": ([vp, i], i),
        "hvr_set_empty_skip": ([vp, i], i),
        "hvr_set_dispatch_order": ([vp, i], i),
        "hvr_set_readback": ([vp, i], i),
        "hvr_dispatch_order_info": ([vp, ctypes.POINTER(i)], i),
        "hvr_skip_info": ([vp, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong),
                           ctypes.POINTER(i)], i),
        "hvr_batch_create": ([vp, ctypes.POINTER(vp)], i),
        "hvr_batch_destroy": ([vp], i),
        "hvr_batch_clear": ([vp], i),
        "hvr_batch_set_tile_shape": ([vp, i], i),
        "hvr_batch_set_kernel": ([vp, i, i, i], i),
        "hvr_batch_set_readback": ([vp, i], i),
        "hvr_batch_add": ([vp, vp, i, vp, i, fp, i, vp, i, i, ctypes.POINTER(i)], i),
        "hvr_batch_run_async": ([vp, i], i),
        "hvr_batch_reserve": ([vp, i, ctypes.c_ulonglong, ctypes.c_ulonglong], i),
        "hvr_batch_wait": ([vp], i),
        "hvr_batch_run": ([vp, i], i),
        "hvr_batch_output": ([vp, i, ctypes.POINTER(u8p)], i),
        "hvr_batch_counts": ([vp, i, ctypes.POINTER(ctypes.POINTER(ctypes.c_uint32))], i),
        "hvr_batch_last_ms": ([vp, fp], i),
        "hvr_volume_upload_scalar_normals": ([vp, fp], i),
        "hvr_batch_query": ([vp, ctypes.POINTER(i)], i),
        "hvr_device_memory": ([i, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong)], i),
        "hvr_volume_generate_blobs": ([vp, i, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double), ctypes.c_double, ctypes.c_ulonglong], i),
        "hvr_render_async_to": ([vp, vp], i),
        "hvr_render_async_to_noalloc": ([vp, vp], i),
        "hvr_device_alloc": ([vp, ctypes.c_uint64, ctypes.POINTER(vp)], i),
        "hvr_device_free": ([vp, vp], i),
        "hvr_copy_to_host": ([vp, vp, vp, ctypes.c_uint64], i),
        "hvr_device_copy": ([vp, vp, vp, ctypes.c_uint64], i),
        "hvr_timing_begin": ([vp, i], i),
        "hvr_timing_end": ([vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(i)], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res


def lib():
    """The loaded library; raises RuntimeError when it is not available."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libhipvr.so not found at {LIB_PATH}; build it with "
            "`python -c 'import pyvr_amd; pyvr_amd.build()'` (needs hipcc)"
        )
    try:
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
    except OSError as e:  # missing HIP runtime etc.
        _load_error = e
        raise RuntimeError(f"cannot load {LIB_PATH}: {e}") from e
    _lib = L
    return _lib


def device_memory(device: int = 0):
    """(free, total) device memory in bytes."""
    f, t = ctypes.c_ulonglong(), ctypes.c_ulonglong()
    check(lib().hvr_device_memory(int(device), ctypes.byref(f), ctypes.byref(t)), "hvr_device_memory")
    return int(f.value), int(t.value)


def check(rc: int, what: str = "") -> None:
    """Raise RuntimeError carrying the library's message for a failed call."""
    if rc != HVR_OK:
        msg = lib().hvr_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}" if what else f"hipvr error {rc}: {msg}")


def inject_fault(function, code: int = -2, after: int = 0) -> None:
    """Make the library entry point ``function`` fail with ``code`` on its
    ``after``-th call from now, once (``hvr_debug_inject_fault``); ``None``
    disarms.  A testing aid: the failure surfaces as the RuntimeError that
    ``check`` raises, with the message "injected fault in <function>"."""
    fn = None if function is None else function.encode()
    check(lib().hvr_debug_inject_fault(fn, int(code), int(after)), "hvr_debug_inject_fault")
