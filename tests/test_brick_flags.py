"""The transparent-brick rule of the march kernel, emulated on the CPU with
float32 arithmetic (pyvr_amd/csrc/hvr_device.h: k_brick_range, k_brick_flags,
the LUT prefix counts of hvr_set_lut, tri_weights and lut_sample).

Skipping a flagged brick is exact only if no density that trilinear
filtering can produce inside it weights a LUT entry with alpha != 0 or a
non-finite colour.  These tests draw bricks and LUTs built to sit on the
rule's edges (densities on entry boundaries, zero-alpha runs ending next to
the range, large and mixed magnitudes) and check every flagged brick by
brute force over many filter weights; and that the rule is not vacuous.
"""

import numpy as np
import pytest

f32 = np.float32


def nz_prefix(lut):
    """Prefix counts of entries that are not exactly transparent (hvr_set_lut):
    alpha 0, and the colours of the entry and its successor and the
    successor's alpha bounded by 1e38 (the delta form reads the successor
    even with weight 0, and 0 * inf is NaN)."""
    nxt = np.concatenate([lut[1:], lut[-1:]])
    ok = lambda x: (np.abs(x) <= 1e38).all(axis=1)
    clear = (lut[:, 3] == 0) & ok(lut[:, :3]) & ok(nxt)
    return np.concatenate([[0], np.cumsum(~clear)])


def brick_flag(lo, hi, nz, n):
    """k_brick_flags for one brick, float32 like the device."""
    lo, hi, fn, fn1 = f32(lo), f32(hi), f32(n), f32(n - 1)
    m = f32(f32(max(abs(lo), abs(hi)) * fn) * f32(1e-5)) + f32(1e-3)
    ulo = f32(f32(lo * fn) - f32(0.5)) - m
    uhi = f32(f32(hi * fn) - f32(0.5)) + m
    i0 = int(np.floor(min(max(ulo, f32(0)), fn1)))
    i1 = int(np.ceil(min(max(uhi, f32(0)), fn1)))
    return nz[i1 + 1] - nz[i0] == 0


def fma32(a, b, c):
    return f32(np.float64(a) * np.float64(b) + np.float64(c))


def trilinear(taps, a, b, g):
    """tri_weights + the fma chain of sample1, float32."""
    a, b, g = f32(a), f32(b), f32(g)
    a1, b1, g1 = f32(1) - a, f32(1) - b, f32(1) - g
    w00, w10, w01, w11 = b1 * g1, b * g1, b1 * g, b * g
    w = [a1 * w00, a * w00, a1 * w10, a * w10, a1 * w01, a * w01, a1 * w11, a * w11]
    r = f32(w[0] * taps[0])
    for k in range(1, 8):
        r = fma32(w[k], taps[k], r)
    return r


def lut_sample(lut, rho):
    """lut_sample with the LDS value/delta table, float32."""
    n = len(lut)
    val = lut.astype(f32)
    dlt = np.zeros_like(val)
    dlt[:-1] = val[1:] - val[:-1]
    u = min(max(fma32(rho, f32(n), f32(-0.5)), f32(0)), f32(n - 1))
    i = int(np.floor(u))
    a = f32(u - f32(np.floor(u)))
    return np.array([fma32(a, dlt[i, c], val[i, c]) for c in range(4)], dtype=f32)


def make_lut(rng, n):
    lut = rng.uniform(0, 1, (n, 4)).astype(f32)
    # transparent runs of random length, some ending at the table's edges
    for _ in range(rng.integers(1, 5)):
        s = int(rng.integers(0, n))
        e = min(n, s + int(rng.integers(1, max(2, n // 3))))
        lut[s:e, 3] = 0
    if rng.random() < 0.5:
        lut[: int(rng.integers(1, n // 2 + 2)), 3] = 0
    return lut


def brick_values(rng, n, lut):
    """Eight taps whose range hugs a transparent run of the LUT."""
    kind = rng.integers(0, 5)
    if kind == 0:                      # all zero (empty space)
        return np.zeros(8, f32)
    if kind == 1:                      # densities on entry boundaries
        k = rng.integers(0, n, 8)
        return ((k + 0.5) / n).astype(f32)
    if kind == 2:                      # tiny noise near zero
        return np.abs(rng.normal(0, 1.0 / n, 8)).astype(f32)
    if kind == 3:                      # inside a zero-alpha run
        zs = np.flatnonzero(lut[:, 3] == 0)
        k = rng.choice(zs) if len(zs) else 0
        return ((k + rng.uniform(-0.5, 1.5, 8)) / n).astype(f32)
    big = f32(10.0 ** rng.uniform(0, 6))      # large / mixed magnitudes
    return (rng.uniform(-1, 1, 8) * big).astype(f32)


@pytest.mark.parametrize("seed", range(6))
def test_flagged_bricks_read_only_transparent_entries(seed):
    rng = np.random.default_rng(1000 + seed)
    flagged = 0
    for _ in range(150):
        n = int(rng.choice([1, 2, 3, 16, 256, 2048]))
        lut = make_lut(rng, n)
        if rng.random() < 0.3:                       # non-finite or huge values anywhere
            lut[rng.integers(0, n), rng.integers(0, 4)] = rng.choice([np.nan, np.inf, -np.inf, 3e38])
        nz = nz_prefix(lut)
        taps = brick_values(rng, n, lut)
        if not brick_flag(taps.min(), taps.max(), nz, n):
            continue
        flagged += 1
        pts = [(0, 0, 0), (1 - 2 ** -24, 0.5, 0.25)] + [tuple(rng.random(3)) for _ in range(60)]
        for a, b, g in pts:
            q = lut_sample(lut, trilinear(taps, a, b, g))
            assert q[3] == 0 and np.isfinite(q[:3]).all(), (n, taps, (a, b, g), q)
    assert flagged > 10


def test_rule_flags_empty_space_and_spares_visible_bricks():
    n = 256
    lut = np.zeros((n, 4), f32)
    lut[:, :3] = 0.5
    lut[1:, 3] = np.linspace(0.01, 1, n - 1)       # only density ~0 is transparent
    nz = nz_prefix(lut)
    assert brick_flag(0.0, 0.0, nz, n)                 # empty space: u clamps to 0, weight of entry 1 is 0
    assert brick_flag(0.0, 0.4 / n, nz, n)             # u < 0 everywhere
    assert not brick_flag(0.0, 0.6 / n, nz, n)         # reaches entry 1
    assert not brick_flag(0.5, 0.5, nz, n)
    lut[7, :3] = np.inf                                # an alpha-0 entry with infinite colour
    lut[7, 3] = 0
    lut[8:11, 3] = 0
    nz = nz_prefix(lut)
    assert not brick_flag(7.5 / n, 7.5 / n, nz, n)
    assert not brick_flag(6.5 / n, 6.5 / n, nz, n)     # weight 0 on entry 7 still reads it
    assert brick_flag(9.5 / n, 9.5 / n, nz, n)         # u = 9 +- margin: entries 8..10
    assert not brick_flag(8.5 / n, 8.5 / n, nz, n)     # u = 8 +- margin reaches entry 7
    lut[1, 3] = np.nan                                 # the successor of empty space
    assert not brick_flag(0.0, 0.0, nz_prefix(lut), n)
    assert not brick_flag(-np.inf, np.inf, nz, n)      # the range a NaN texel produces
