"""CPU oracle for the HIP volume renderer -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker.  The product
(``pyvr_amd``) never imports it.

Two independent restatements of the reference's render path (the fragment
shader ``pyvr/shaders/volume.frag.glsl:70-122`` and the GL state of
``pyvr/moderngl_renderer/manager.py``):

* ``render`` -- the C oracle (``oracle_vr.c``), FP64 (canonical) or FP32
  (the shader's own arithmetic; also the CPU baseline), OpenMP over pixels;
* ``numpy_oracle.render_numpy`` -- a NumPy FP64 restatement written with a
  different structure (vectorised over rays, GL-spec weighted sums, numpy
  matrix inverse), used to cross-check the C code on small frames.

Pins of each function (closed forms, special cases, brute force, library
equivalents) live in ``tests/test_oracle_*.py``.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_vr.so")
_lib = None


class OvrParams(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int),
        ("height", ctypes.c_int),
        ("view", ctypes.c_float * 16),
        ("proj", ctypes.c_float * 16),
        ("cam_pos", ctypes.c_float * 3),
        ("bmin", ctypes.c_float * 3),
        ("bmax", ctypes.c_float * 3),
        ("step", ctypes.c_float),
        ("ref_step", ctypes.c_float),
        ("max_steps", ctypes.c_int),
        ("ambient", ctypes.c_float),
        ("diffuse", ctypes.c_float),
        ("light_pos", ctypes.c_float * 3),
        ("light_target", ctypes.c_float * 3),
        ("et_threshold", ctypes.c_float),
        ("full_loop", ctypes.c_int),
        ("direct_positions", ctypes.c_int),
    ]


class OvrTex3(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.POINTER(ctypes.c_float)),
        ("s0", ctypes.c_int),
        ("s1", ctypes.c_int),
        ("s2", ctypes.c_int),
        ("channels", ctypes.c_int),
    ]


def build(force: bool = False) -> str:
    """Compile ``liboracle_vr.so`` with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        args = [ctypes.POINTER(OvrParams), ctypes.POINTER(OvrTex3), ctypes.POINTER(OvrTex3),
                ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_long, ctypes.c_long,
                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        for name in ("ovr_render_f64", "ovr_render_f32"):
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.ovr_tex3_f64.argtypes = [ctypes.POINTER(OvrTex3), ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.ovr_lut_f64.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_double,
                                  ctypes.POINTER(ctypes.c_double)]
        L.ovr_inv4.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.ovr_inv4.restype = ctypes.c_int
        L.ovr_blend8.argtypes = [ctypes.c_double, ctypes.c_double]
        L.ovr_blend8.restype = ctypes.c_uint
        _lib = L
    return _lib


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def make_tex(arr):
    """(OvrTex3, keepalive) for a float32 (s0,s1,s2) or (s0,s1,s2,3) array."""
    if arr is None:
        return None, None
    a = np.ascontiguousarray(arr, dtype=np.float32)
    ch = 1 if a.ndim == 3 else a.shape[3]
    t = OvrTex3(_fptr(a), a.shape[0], a.shape[1], a.shape[2], ch)
    return t, a


def make_params(u: dict, full_loop: bool = False) -> OvrParams:
    """Uniform block from a dict with keys width, height, view, proj, cam_pos,
    bmin, bmax, step, ref_step, max_steps, ambient, diffuse, light_pos,
    light_target, et_threshold (matrices: numpy float32 4x4 as produced by
    Camera.get_view_matrix / get_projection_matrix)."""
    p = OvrParams()
    p.width, p.height = int(u["width"]), int(u["height"])
    p.view[:] = np.asarray(u["view"], np.float32).ravel().tolist()
    p.proj[:] = np.asarray(u["proj"], np.float32).ravel().tolist()
    p.cam_pos[:] = np.asarray(u["cam_pos"], np.float32).tolist()
    p.bmin[:] = np.asarray(u["bmin"], np.float32).tolist()
    p.bmax[:] = np.asarray(u["bmax"], np.float32).tolist()
    p.step = u["step"]
    p.ref_step = u["ref_step"]
    p.max_steps = int(u["max_steps"])
    p.ambient = u["ambient"]
    p.diffuse = u["diffuse"]
    p.light_pos[:] = np.asarray(u["light_pos"], np.float32).tolist()
    p.light_target[:] = np.asarray(u["light_target"], np.float32).tolist()
    p.et_threshold = u.get("et_threshold", 0.99)
    p.full_loop = 1 if full_loop else 0
    return p


def render(u: dict, scalar, normal, lut, mode: str = "f64", full_loop: bool = False,
           pixels=None, nthreads: int = 0, want=("rgba8", "rgba_f", "counts")):
    """Render a frame (or the linear pixel range ``pixels=(begin, end)``).

    ``scalar``: float32 (s0,s1,s2) array or None (unbound sampler reads 0).
    ``normal``: float32 (s0,s1,s2,3) array; the scalar array itself to model a
    renderer without normals (sampled as (rho, 0, 0)); or None (reads 0).
    ``lut``: float32 (N, 4) RGBA LUT or None (reads (0, 0, 0, 1)).
    Returns a dict with the requested arrays: rgba8 (H, W, 4) uint8 bottom
    row first, rgba_f (H, W, 4) float64 shader output, counts (H, W) uint32.
    """
    L = lib()
    W, H = int(u["width"]), int(u["height"])
    p = make_params(u, full_loop)
    ts, ks = make_tex(scalar)
    tn, kn = make_tex(normal)
    lut_a = None if lut is None else np.ascontiguousarray(lut, dtype=np.float32).reshape(-1, 4)
    begin, end = (0, W * H) if pixels is None else pixels
    out = {}
    rgba8 = np.zeros((H, W, 4), np.uint8) if "rgba8" in want else None
    rgba_f = np.zeros((H, W, 4), np.float64) if "rgba_f" in want else None
    counts = np.zeros((H, W), np.uint32) if "counts" in want else None
    fn = L.ovr_render_f64 if mode == "f64" else L.ovr_render_f32
    rc = fn(ctypes.byref(p),
            ctypes.byref(ts) if ts is not None else None,
            ctypes.byref(tn) if tn is not None else None,
            _fptr(lut_a) if lut_a is not None else None,
            0 if lut_a is None else lut_a.shape[0],
            begin, end,
            rgba8.ctypes.data if rgba8 is not None else None,
            rgba_f.ctypes.data if rgba_f is not None else None,
            counts.ctypes.data if counts is not None else None,
            nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle render failed ({rc})")
    del ks, kn
    if rgba8 is not None:
        out["rgba8"] = rgba8
    if rgba_f is not None:
        out["rgba_f"] = rgba_f
    if counts is not None:
        out["counts"] = counts
    return out


def tex3(arr, c):
    """GL_LINEAR sample of a texture (oracle's C implementation) at normalised
    coordinate c = (cx, cy, cz); returns 3 channels."""
    t, keep = make_tex(arr)
    out = (ctypes.c_double * 3)()
    lib().ovr_tex3_f64(ctypes.byref(t), float(c[0]), float(c[1]), float(c[2]), out)
    return np.array(out[:])


def lut_sample(lut, rho):
    a = np.ascontiguousarray(lut, dtype=np.float32).reshape(-1, 4)
    out = (ctypes.c_double * 4)()
    lib().ovr_lut_f64(_fptr(a), a.shape[0], float(rho), out)
    return np.array(out[:])


def inv4(m):
    a = np.ascontiguousarray(m, dtype=np.float64).ravel()
    out = np.zeros(16, np.float64)
    ok = lib().ovr_inv4(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out.reshape(4, 4) if ok else None


def blend8(c, a):
    """Stored (channel, alpha) bytes for shader output colour c, alpha a."""
    v = lib().ovr_blend8(float(c), float(a))
    return v & 0xFF, (v >> 8) & 0xFF
