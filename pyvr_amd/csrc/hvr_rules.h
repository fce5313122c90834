// hvr_rules.h -- the texel-layout and transparency rules of libhipvr as
// plain C++17 (no HIP headers): constexpr functions, which HIP compiles for
// both the host and the device, so the kernels (hvr_device.h), the host side
// (hipvr.hip) and a native test driver built with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/native/rules_check.cpp) run the same
// code.
#pragma once

#include <stdint.h>

namespace hvr {

// Bricks along an axis of GL size s with 2^lb interior texels per brick: the
// extended texture holds s + 2 texels (the edge copies), X in [0, s + 1].
constexpr uint32_t bricks_along(int s, int lb) { return ((uint32_t)s >> lb) + 1u; }

// Slots along one axis that hold texel coordinate c of a texture of size s
// stored in nb bricks: (brick, local) pairs; returns how many (1..4).  In
// extended coordinates the texel sits at X = c + 1, and also at X = 0 when
// c = 0 and at X = s + 1 when c = s - 1 (the CLAMP_TO_EDGE copies).  Each X
// lives in brick X >> lb at local X & (B - 1), and, when it is the first
// texel of a brick, in the previous brick's apron at local B.  A position
// whose own brick would be brick nb (X = s + 1 with s + 1 a multiple of B)
// is only ever read through the apron of brick nb - 1.
constexpr int axis_slots(int c, int s, int lb, uint32_t nb, uint32_t br[4], uint32_t lo[4])
{
    const uint32_t B = 1u << lb, M = B - 1u;
    int n = 0;
    uint32_t X[3] = {(uint32_t)c + 1u, 0u, 0u};
    int nx = 1;
    if (c == 0) X[nx++] = 0u;
    if (c == s - 1) X[nx++] = (uint32_t)s + 1u;
    for (int k = 0; k < nx; ++k) {
        if ((X[k] >> lb) < nb) {
            br[n] = X[k] >> lb;
            lo[n] = X[k] & M;
            ++n;
        }
        if ((X[k] & M) == 0u && X[k] > 0u) {  // apron copy in the previous brick
            br[n] = (X[k] >> lb) - 1u;
            lo[n] = B;
            ++n;
        }
    }
    return n;
}

// Whether a brick whose density range is [lo, hi] can only ever sample LUT
// entries that are exactly transparent.  A sample of density rho reads
// u = clamp(rho*N - 1/2, 0, N-1) and weights entry floor(u) by 1 - frac(u)
// and entry floor(u) + 1 by frac(u), so the entries with non-zero weight over
// densities in [lo, hi] are floor(u(lo)) .. ceil(u(hi)); the range is widened
// by a margin for the rounding of the interpolated density (a few ulps of
// the largest tap).  `nz` holds the prefix counts (N + 1 values) of entries
// that are not exactly transparent.  Ranges of bricks holding a NaN are
// (-inf, +inf).  The clamps follow IEEE maxNum/minNum (fmaxf/fminf); the
// clamped values are non-negative, so truncation is floor.
constexpr bool brick_clear_range(float lo, float hi, const int *nz, int n)
{
    const float fn = (float)n, fn1 = (float)(n - 1);
    const float alo = lo < 0.f ? -lo : lo, ahi = hi < 0.f ? -hi : hi;
    const float m = (alo > ahi ? alo : ahi) * fn * 1e-5f + 1e-3f;
    float ulo = lo * fn - 0.5f - m;
    float uhi = hi * fn - 0.5f + m;
    ulo = ulo > 0.f ? (ulo < fn1 ? ulo : fn1) : 0.f;
    uhi = uhi > 0.f ? (uhi < fn1 ? uhi : fn1) : 0.f;
    const int i0 = (int)ulo;
    const int t1 = (int)uhi;
    const int i1 = (float)t1 < uhi ? t1 + 1 : t1;
    return nz[i1 + 1] - nz[i0] == 0;
}

// Prefix counts (n + 1 values) of LUT entries that are not exactly
// transparent.  The march reads entry i as v[i] + a*(v[i+1] - v[i]) with a
// in [0, 1), so entry i yields alpha 0 and a finite colour for every a (the
// reach of a > 0 into entry i+1 is accounted for by the brick rule) only
// when its alpha is 0 and the colours of it and of its successor, and the
// successor's alpha, are bounded (|x| <= 1e38 also excludes NaN and
// overflow of the difference: 0 * inf is NaN).
constexpr bool lut_bounded(float x) { return (x < 0.f ? -x : x) <= 1e38f; }

constexpr void lut_prefix_counts(const float *src, int cnt, int *nz)
{
    nz[0] = 0;
    for (int i = 0; i < cnt; ++i) {
        const float *e = src + 4 * i, *f = i + 1 < cnt ? e + 4 : e;
        const bool clear = e[3] == 0.f && lut_bounded(e[0]) && lut_bounded(e[1]) && lut_bounded(e[2]) &&
                           lut_bounded(f[0]) && lut_bounded(f[1]) && lut_bounded(f[2]) && lut_bounded(f[3]);
        nz[i + 1] = nz[i] + (clear ? 0 : 1);
    }
}

// Cost-ordered dispatch of a frame's 16x16 ray blocks.  A frame ends when
// its longest waves do; dispatched in row order, long blocks that start
// late (after the first round of resident blocks) set the frame's end.
// Ray setup gives each block a cost key, the largest iteration bound of its
// rays, and files the block under a key bucket; the march then takes the
// blocks of the largest keys first.  Buckets split [0, max_steps] evenly,
// bucket 0 holding the largest keys.
constexpr int kOrderBuckets = 16;
constexpr int order_bucket(int key, int max_steps)
{
    if (key <= 0 || max_steps <= 0) return kOrderBuckets - 1;
    const long long k = (long long)key * kOrderBuckets / ((long long)max_steps + 1);
    return kOrderBuckets - 1 - (int)(k < kOrderBuckets - 1 ? k : kOrderBuckets - 1);
}

// Dispatch position b of a cost-ordered frame -> logical block (row-major,
// `ntx` blocks per row).  The first `head` positions (the blocks of the
// first multiple-of-8 block rows) are dealt to 8 classes in turn: class
// b & 7 holds block rows class, class + 8, ... (blocks are dispatched
// round-robin over the 8 XCDs, so one class runs on one XCD and its L2 sees
// whole rows), and within a class the blocks are taken bucket by bucket
// from `slots` (`cap` entries per class and bucket, cap = the class's
// block count) in the numbers `cnt` holds (kOrderBuckets per class).
// Later positions keep the row-major order.  Returns b itself when the
// counts do not cover the position (never with counts ray setup wrote).
template <typename U>
constexpr int ordered_block(int b, int head, const U *cnt, const U *slots, int cap)
{
    if (b >= head) return b;
    const int x = b & 7;
    uint32_t j = (uint32_t)(b >> 3);
    for (int k = 0; k < kOrderBuckets; ++k) {
        const uint32_t c = (uint32_t)cnt[x * kOrderBuckets + k];
        if (j < c)
            return j < (uint32_t)cap ? (int)slots[(uint64_t)(x * kOrderBuckets + k) * (uint64_t)cap + j] : b;
        j -= c;
    }
    return b;
}

// Inverse of a 4x4 matrix (row-major, m[4*r + c]) by Gauss-Jordan
// elimination with partial pivoting.  Returns false when singular.
constexpr bool inverse4(const double *m, double *inv)
{
    double a[4][8] = {};
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 8; ++c) a[r][c] = c < 4 ? m[4 * r + c] : (c - 4 == r ? 1.0 : 0.0);
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int r = col + 1; r < 4; ++r) {
            const double x = a[r][col] < 0 ? -a[r][col] : a[r][col];
            const double y = a[piv][col] < 0 ? -a[piv][col] : a[piv][col];
            if (x > y) piv = r;
        }
        if (a[piv][col] == 0.0) return false;
        if (piv != col)
            for (int c = 0; c < 8; ++c) {
                const double t = a[col][c];
                a[col][c] = a[piv][c];
                a[piv][c] = t;
            }
        const double d = a[col][col];
        for (int c = 0; c < 8; ++c) a[col][c] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == col) continue;
            const double f = a[r][col];
            if (f != 0.0)
                for (int c = 0; c < 8; ++c) a[r][c] -= f * a[col][c];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[4 * r + c] = a[r][c + 4];
    return true;
}

}  // namespace hvr
