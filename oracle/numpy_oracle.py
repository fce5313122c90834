"""NumPy FP64 restatement of the render path -- TEST INFRASTRUCTURE ONLY.

A second, independently structured transcription of the reference's
fragment shader (``pyvr/shaders/volume.frag.glsl:47-121``) and GL state
(``pyvr/moderngl_renderer/manager.py:29-31, 95-101, 123-129, 163-181,
212-230``), used to cross-check the C oracle on small frames:

* matrices are inverted with ``numpy.linalg.inv`` (the C oracle uses its own
  Gauss-Jordan);
* textures are filtered with the GL specification's eight-term weighted sum
  (the C oracle uses nested linear interpolation);
* rays are processed as arrays, one ray-step per loop iteration, and every
  iteration ``i < max_steps`` is evaluated (the shader's loop, no shortcut).

Cost is O(W * H * max_steps) array work: use it for frames up to ~64x64.
"""

import numpy as np

FLT_MIN = 2.0 ** -126


def _texel_table(arr):
    a = np.asarray(arr, dtype=np.float64)
    ch = 1 if a.ndim == 3 else a.shape[3]
    s0, s1, s2 = a.shape[:3]
    # GL texel (x, y, z) = array bytes at x + s0*(y + s1*z): C-order flat index
    return a.reshape(-1, ch), (s0, s1, s2), ch


def tex_gl(arr, c):
    """GL_LINEAR + CLAMP_TO_EDGE sample of a 3-D texture at coordinates c
    (..., 3); returns (..., 3) with missing channels as 0."""
    flat, (s0, s1, s2), ch = _texel_table(arr)
    c = np.asarray(c, dtype=np.float64)
    sizes = np.array([s0, s1, s2], dtype=np.float64)
    u = c * sizes - 0.5
    base = np.floor(u)
    frac = u - base
    base = base.astype(np.int64)
    out = np.zeros(c.shape[:-1] + (3,), dtype=np.float64)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = np.ones(c.shape[:-1])
                idx = []
                for axis, dd in enumerate((dx, dy, dz)):
                    f = frac[..., axis]
                    w = w * (f if dd else 1.0 - f)
                    idx.append(np.clip(base[..., axis] + dd, 0, int(sizes[axis]) - 1))
                lin = idx[0] + s0 * (idx[1] + s1 * idx[2])
                out[..., :ch] += w[..., None] * flat[lin]
    return out


def lut_gl(lut, rho):
    """Linear, clamp-to-edge sample of an (N, 4) LUT at u = rho*N - 0.5."""
    q = np.asarray(lut, dtype=np.float64).reshape(-1, 4)
    n = q.shape[0]
    u = np.clip(np.asarray(rho, dtype=np.float64) * n - 0.5, -1.0, float(n))
    i0 = np.floor(u)
    a = (u - i0)[..., None]
    i0 = i0.astype(np.int64)
    return (1.0 - a) * q[np.clip(i0, 0, n - 1)] + a * q[np.clip(i0 + 1, 0, n - 1)]


def quantise_blend(rgba_f):
    """RGBA8 bytes stored by the blend (SRC_ALPHA, ONE_MINUS_SRC_ALPHA) over
    a zero-cleared RGBA8 target."""
    x = np.clip(np.nan_to_num(rgba_f, nan=0.0), 0.0, 1.0)
    q = np.floor(x * 255.0 + 0.5).astype(np.int64)
    aq = q[..., 3:4]
    out = np.empty_like(q)
    out[..., :3] = (q[..., :3] * aq + 127) // 255
    out[..., 3:] = (aq * aq + 127) // 255
    return out.astype(np.uint8)


def ray_setup(u):
    """Per-pixel (origin, direction, t_near, t_far, hit) in float64."""
    W, H = int(u["width"]), int(u["height"])
    p_gl = np.asarray(u["proj"], np.float32).astype(np.float64).T
    v_gl = np.asarray(u["view"], np.float32).astype(np.float64).T
    pinv = np.linalg.inv(p_gl)
    vinv = np.linalg.inv(v_gl)
    py, px = np.mgrid[0:H, 0:W]
    ndc_x = (px + 0.5) / W * 2.0 - 1.0
    ndc_y = (py + 0.5) / H * 2.0 - 1.0
    clip = np.stack([ndc_x, ndc_y, -np.ones_like(ndc_x), np.ones_like(ndc_x)], axis=-1)
    eye = clip @ pinv.T
    eye[..., 2] = -1.0
    eye[..., 3] = 0.0
    world = (eye @ vinv.T)[..., :3]
    d = world / np.linalg.norm(world, axis=-1, keepdims=True)
    o = np.asarray(u["cam_pos"], np.float32).astype(np.float64)
    bmin = np.asarray(u["bmin"], np.float32).astype(np.float64)
    bmax = np.asarray(u["bmax"], np.float32).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        t0 = (bmin - o) * inv
        t1 = (bmax - o) * inv
    lo = np.where(np.isnan(t0) | np.isnan(t1), -np.inf, np.minimum(t0, t1))
    hi = np.where(np.isnan(t0) | np.isnan(t1), np.inf, np.maximum(t0, t1))
    tn = lo.max(axis=-1)
    tf = hi.min(axis=-1)
    hit = (tn <= tf) & (tf > 0)
    return o, d, np.maximum(tn, 0.0), tf, hit


def render_numpy(u, scalar, normal, lut):
    """Shader output (H, W, 4) float64, RGBA8 bytes and per-pixel sample counts.
    Arguments as ``oracle.render``."""
    o, d, tn, tf, hit = ray_setup(u)
    H, W = hit.shape
    bmin = np.asarray(u["bmin"], np.float32).astype(np.float64)
    bmax = np.asarray(u["bmax"], np.float32).astype(np.float64)
    ext = bmax - bmin
    s = float(np.float32(u["step"]))
    r = float(np.float32(u["ref_step"]))
    tau = float(np.float32(u.get("et_threshold", 0.99)))
    ka = float(np.float32(u["ambient"]))
    kd = float(np.float32(u["diffuse"]))
    lvec = (np.asarray(u["light_target"], np.float32).astype(np.float64)
            - np.asarray(u["light_pos"], np.float32).astype(np.float64))
    lnorm = np.sqrt(lvec @ lvec)
    L = lvec / lnorm if lnorm > 0 else None

    C = np.zeros((H, W, 3))
    A = np.zeros((H, W))
    count = np.zeros((H, W), np.uint32)
    for i in range(int(u["max_steps"])):
        live = hit & (A < tau)
        if not live.any():
            break
        pos = o + d * (tn + i * s)[..., None]
        q = (pos - bmin) / ext
        c = q[..., ::-1]   # swizzle (z, y, x)
        inside = np.all((c >= 0.0) & (c <= 1.0), axis=-1) & live
        if not inside.any():
            continue
        cs = c[inside]
        rho = tex_gl(scalar, cs)[:, 0] if scalar is not None else np.zeros(len(cs))
        rgba = lut_gl(lut, rho) if lut is not None else np.tile([0.0, 0.0, 0.0, 1.0], (len(cs), 1))
        a = 1.0 - np.exp(-rgba[:, 3] * s / r)
        diff = np.zeros(len(cs))
        if normal is not None and L is not None:
            nv = tex_gl(normal, cs)
            nn = np.sum(nv * nv, axis=-1)
            ok = nn >= FLT_MIN
            dots = np.zeros(len(cs))
            dots[ok] = (nv[ok] @ L) / np.sqrt(nn[ok])
            diff = np.maximum(dots, 0.0)
        lam = ka + kd * diff
        tr = 1.0 - A[inside]
        C[inside] += tr[:, None] * rgba[:, :3] * (lam * a)[:, None]
        A[inside] += tr * a
        count[inside] += 1
    rgba_f = np.concatenate([C, A[..., None]], axis=-1)
    return {"rgba_f": rgba_f, "rgba8": quantise_blend(rgba_f), "counts": count}
