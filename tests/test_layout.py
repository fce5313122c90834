"""The device's apron-brick texture layout, emulated on the CPU with the same
integer arithmetic as pyvr_amd/csrc/hvr_device.h (axis_slots / k_upload for
the scatter, footprint() for the reads).  Checks, for awkward shapes and
every brick size: every slot written and every tap read lies inside the
allocation, every tap holds the texel GL's CLAMP_TO_EDGE selects, and the
filtered value equals the oracle's texture sample."""

import numpy as np
import pytest

import oracle


def bricks_along(s, lb):
    return (s >> lb) + 1


def axis_slots(c, s, lb, nb):
    B, M = 1 << lb, (1 << lb) - 1
    X = [c + 1] + ([0] if c == 0 else []) + ([s + 1] if c == s - 1 else [])
    out = []
    for x in X:
        if (x >> lb) < nb:
            out.append((x >> lb, x & M))
        if (x & M) == 0 and x > 0:
            out.append(((x >> lb) - 1, B))
    return out


def build_storage(tex, lb):
    s0, s1, s2 = tex.shape
    E = (1 << lb) + 1
    nb = [bricks_along(s, lb) for s in (s0, s1, s2)]
    n_slots = nb[0] * nb[1] * nb[2] * E ** 3
    store = np.full(n_slots, np.nan)
    flat = tex.ravel()                      # GL texel (x, y, z) = flat[x + s0 (y + s1 z)]
    for z in range(s2):
        for y in range(s1):
            for x in range(s0):
                v = flat[x + s0 * (y + s1 * z)]
                for bz, lz in axis_slots(z, s2, lb, nb[2]):
                    for by, ly in axis_slots(y, s1, lb, nb[1]):
                        for bx, lx in axis_slots(x, s0, lb, nb[0]):
                            brick = (bz * nb[1] + by) * nb[0] + bx
                            local = (lz * E + ly) * E + lx
                            idx = brick * E ** 3 + local
                            assert 0 <= idx < n_slots
                            assert np.isnan(store[idx]) or store[idx] == v  # one source per slot
                            store[idx] = v
    return store, nb


def sample(store, nb, lb, shape, w, ext):
    """footprint() + trilinear weights, float32 arithmetic like the kernel."""
    E, M = (1 << lb) + 1, (1 << lb) - 1
    k = [np.float32(shape[a] / np.float64(ext[a])) for a in range(3)]
    u = [np.float32(np.float32(w[a]) * k[a] + np.float32(0.5)) for a in range(3)]
    X = [int(np.floor(ua)) for ua in u]
    frac = [float(u[a] - np.floor(u[a])) for a in range(3)]
    for a in range(3):
        assert 0 <= X[a] <= shape[a]
    brick = ((X[2] >> lb) * nb[1] + (X[1] >> lb)) * nb[0] + (X[0] >> lb)
    local = ((X[2] & M) * E + (X[1] & M)) * E + (X[0] & M)
    base = brick * E ** 3 + local
    acc, taps = 0.0, []
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                idx = base + dx + dy * E + dz * E * E
                assert 0 <= idx < len(store)
                val = store[idx]
                assert not np.isnan(val)
                wgt = ((frac[0] if dx else 1 - frac[0]) * (frac[1] if dy else 1 - frac[1])
                       * (frac[2] if dz else 1 - frac[2]))
                acc += wgt * val
                taps.append((X[0] + dx, X[1] + dy, X[2] + dz, val))
    return acc, taps


@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 7, 7), (8, 5, 3), (15, 16, 17), (9, 2, 31), (23, 24, 6)])
@pytest.mark.parametrize("lb", [2, 3, 4])
def test_apron_layout_reads_clamped_texels(shape, lb):
    rng = np.random.default_rng(sum(shape) + lb)
    tex = rng.random(shape).astype(np.float32)
    store, nb = build_storage(tex, lb)
    ext = (2.0, 1.5, 0.75)          # world extents of the axes sampled by texture x, y, z
    flat = tex.ravel()
    s0, s1, s2 = shape
    pts = [(0.0, 0.0, 0.0), ext, (ext[0], 0.0, ext[2])] + [tuple(rng.uniform(0, e) for e in ext)
                                                            for _ in range(150)]
    for w in pts:
        got, taps = sample(store, nb, lb, shape, w, ext)
        for X, Y, Z, val in taps:   # extended coordinate X holds texel clamp(X - 1)
            x, y, z = (min(max(X - 1, 0), s0 - 1), min(max(Y - 1, 0), s1 - 1), min(max(Z - 1, 0), s2 - 1))
            assert val == flat[x + s0 * (y + s1 * z)]
        ref = oracle.tex3(tex, tuple(w[a] / ext[a] for a in range(3)))[0]
        assert got == pytest.approx(ref, abs=1e-5)
