"""The march's direction-normalised walk, brute-forced in float32 on the CPU.

The march kernel (march_ray in pyvr_amd/csrc/hvr_device.h) walks the in-box
run of a ray with the positions negated on every axis whose step is
negative, q = s*p with s = -1 there (else +1), steps s*ds >= 0 and offsets
a = fl(q - s*bmin), and tests only a <= (s > 0 ? ext : 0) per axis.  Its
claim: a == s * fl(p - bmin) exactly at every sample (round-to-nearest
commutes with negation), and within the run the one-sided test gives the
same answer as the shader's two-sided test 0 <= fl(p - bmin) <= ext
(volume.frag.glsl:88-92), because a never decreases and its lower bound is
met at the run's first sample.  Here both walks run side by side in numpy
float32 (each operation correctly rounded, as on the GPU) over rays chosen
to stress the argument: negative, zero and negative-zero steps, steps below
half an ulp of the position, starts on the faces, asymmetric boxes.  The
GPU parity tests check the kernel itself against the oracle; this pins the
arithmetic fact it rests on.
"""

import numpy as np
import pytest

f32 = np.float32


def shader_walk(p0, ds, bmin, ext, n):
    """Positions p_k = fl(p_{k-1} + ds), offsets w_k = fl(p_k - bmin) and the
    two-sided in-box test, for k < n."""
    p = p0.copy()
    ws, ins = [], []
    for _ in range(n):
        w = (p - bmin).astype(f32)
        ws.append(w)
        ins.append(np.all((w >= 0) & (w <= ext), axis=1))
        p = (p + ds).astype(f32)
    return np.stack(ws), np.stack(ins)


def kernel_run(p0, ds, bmin, ext, n, ins_ref):
    """The kernel's walk from each ray's first in-box sample (found with the
    two-sided test, as the kernel's leading loop does): (a per sample, the
    one-sided test per sample, the first index of the run)."""
    nray = p0.shape[0]
    start = np.where(ins_ref.any(axis=0), ins_ref.argmax(axis=0), n)
    s = np.where(ds < 0, f32(-1), f32(1)).astype(f32)
    dq = (s * ds).astype(f32)
    bq = (-(s * bmin)).astype(f32)
    lim = np.where(s > 0, ext, f32(0)).astype(f32)
    # advance each ray to its start with the shader's own steps, then negate
    p = p0.copy()
    q = np.zeros_like(p)
    a_all = np.full((n, nray, 3), np.nan, f32)
    one_sided = np.zeros((n, nray), bool)
    for k in range(n):
        at_start = start == k
        q[at_start] = s[at_start] * p[at_start]
        active = start <= k
        a = (q + bq).astype(f32)
        a_all[k, active] = a[active]
        one_sided[k] = active & np.all(a <= lim, axis=1)
        q = np.where(active[:, None], (q + dq).astype(f32), q)
        p = (p + ds).astype(f32)
    return a_all, one_sided, start, s


def make_rays(rng, n):
    bmin = rng.uniform(-1.0, 0.2, 3).astype(f32)
    ext = rng.uniform(0.3, 1.8, 3).astype(f32)
    bmax = (bmin + ext).astype(f32)
    ext = (bmax - bmin).astype(f32)              # the kernel's ext = fl(bmax - bmin)
    # starts: inside, on faces, slightly outside
    p0 = rng.uniform(bmin - 0.1, bmax + 0.1, (n, 3)).astype(f32)
    face = rng.integers(0, 7, (n, 3))
    p0 = np.where(face == 1, bmin, p0)
    p0 = np.where(face == 2, bmax, p0).astype(f32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    kind = rng.integers(0, 8, (n, 3))
    d = np.where(kind == 1, 0.0, d)              # +0 step on an axis
    d = np.where(kind == 2, -0.0, d)             # -0 step: s = +1
    d = np.where(kind == 3, np.sign(d) * 1e-9, d)   # below half an ulp: the axis stays put
    step = rng.choice([0.005, 0.01, 0.05, 0.13])
    ds = (d * step).astype(f32)
    return p0, ds, bmin, ext


@pytest.mark.parametrize("seed", range(6))
def test_negated_walk_is_exact_and_one_sided_test_agrees(seed):
    rng = np.random.default_rng(seed)
    p0, ds, bmin, ext = make_rays(rng, 4000)
    n = 160
    ws, ins_ref = shader_walk(p0, ds, bmin, ext, n)
    a_all, one_sided, start, s = kernel_run(p0, ds, bmin, ext, n, ins_ref)
    assert (start < n).sum() > 1000              # most rays have a run
    for r in range(p0.shape[0]):
        k0 = start[r]
        if k0 >= n:
            continue
        a = a_all[k0:, r]
        w = ws[k0:, r]
        # a == s * w exactly (as numbers: the sign of a zero may differ)
        assert np.array_equal(a, s[r] * w), r
        # the run: the shader's in-box samples from k0 up to the first outside one
        out = np.flatnonzero(~ins_ref[k0:, r])
        end = out[0] if out.size else n - k0
        assert ins_ref[k0:k0 + end, r].all()
        assert one_sided[k0:k0 + end, r].all(), r
        if end < n - k0:
            assert not one_sided[k0 + end, r], r
        # in the box the footprint takes |a| == w
        assert np.array_equal(np.abs(a[:end]), w[:end])


def test_a_wrong_sign_is_caught():
    """The same comparison with the offset's sign not flipped (a plausible
    slip: bq = s*bmin) fails on rays with a negative step."""
    rng = np.random.default_rng(11)
    p0, ds, bmin, ext = make_rays(rng, 500)
    ds = -np.abs(ds)                             # every axis steps down
    s = np.where(ds < 0, f32(-1), f32(1)).astype(f32)
    w = (p0 - bmin).astype(f32)
    a_wrong = ((s * p0) + (s * bmin)).astype(f32)
    assert not np.array_equal(a_wrong, s * w)
