// Native check of the layout and transparency rules the kernels use
// (pyvr_amd/csrc/hvr_rules.h), built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_sanitizers.py:
//  * axis_slots: scattering every texel of an axis to its slots and reading
//    every trilinear footprint back (lower corner X0 in [0, s], taps X0 and
//    X0 + 1 of brick X0 >> lb) gives CLAMP_TO_EDGE's texels clamp(X0 - 1)
//    and clamp(X0), every slot read was written, and every slot is in range;
//  * brick_clear_range: a brick declared transparent gives, for densities
//    across its range (ends and their float neighbours included), an exact
//    alpha 0 and finite colours through the march's LUT read
//    (v[i] + a * (v[i+1] - v[i]) at u = clamp(rho*N - 1/2, 0, N-1));
//  * lut_prefix_counts against a direct count;
//  * inverse4: M * inverse4(M) = I on random matrices, singular rejected;
//  * cost-ordered dispatch: blocks filed as the ordered ray setup files them
//    (in a random order, as blocks run) and read back through
//    ordered_block give every block exactly once, keep each dispatch class
//    on its own block rows (class = position & 7 -> rows = class mod 8), and
//    take a class's blocks in non-decreasing bucket order (decreasing keys).
// Exit status 0 on success.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "hvr_rules.h"

using namespace hvr;

static uint64_t g_state = 0x2545f4914f6cdd1dull;
static double urand()
{
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return (double)(g_state >> 11) * 0x1p-53;
}
static int irand(int n) { return (int)(urand() * n); }

static int fails = 0;
#define CHECK(cond, ...)                                 \
    do {                                                 \
        if (!(cond)) {                                   \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                \
            fprintf(stderr, "\n");                       \
            if (++fails > 20) exit(1);                   \
        }                                                \
    } while (0)

static void check_layout(int s, int lb)
{
    const uint32_t nb = bricks_along(s, lb), B = 1u << lb, E = B + 1u;
    std::vector<int> slot((size_t)nb * E, -1);
    for (int c = 0; c < s; ++c) {
        uint32_t br[4], lo[4];
        const int n = axis_slots(c, s, lb, nb, br, lo);
        CHECK(n >= 1 && n <= 4, "s=%d lb=%d c=%d: %d slots", s, lb, c, n);
        for (int k = 0; k < n; ++k) {
            CHECK(br[k] < nb && lo[k] <= B, "s=%d lb=%d c=%d: slot (%u, %u) out of range", s, lb, c, br[k], lo[k]);
            if (br[k] < nb && lo[k] <= B) {
                int &v = slot[(size_t)br[k] * E + lo[k]];
                CHECK(v == -1 || v == c, "s=%d lb=%d: slot written twice", s, lb);
                v = c;
            }
        }
    }
    for (int X0 = 0; X0 <= s; ++X0) {
        const uint32_t b = (uint32_t)X0 >> lb, l = (uint32_t)X0 & (B - 1u);
        CHECK(b < nb, "s=%d lb=%d X0=%d: brick %u of %u", s, lb, X0, b, nb);
        if (b >= nb) continue;
        const int t0 = slot[(size_t)b * E + l], t1 = slot[(size_t)b * E + l + 1];
        const int w0 = X0 - 1 < 0 ? 0 : (X0 - 1 > s - 1 ? s - 1 : X0 - 1);
        const int w1 = X0 > s - 1 ? s - 1 : X0;
        CHECK(t0 == w0 && t1 == w1, "s=%d lb=%d X0=%d: taps %d %d want %d %d", s, lb, X0, t0, t1, w0, w1);
    }
}

// the march's LUT read (lut_sample in hvr_device.h): entry/difference pairs
static void lut_read(const std::vector<float> &lut, int n, float rho, float out[4])
{
    const float fn = (float)n, fn1 = (float)(n - 1);
    float u = fmaf(rho, fn, -0.5f);
    u = fminf(fmaxf(u, 0.f), fn1);
    const int i = (int)u;
    const float a = u - floorf(u);
    for (int c = 0; c < 4; ++c) {
        const float v = lut[4 * i + c];
        const float d = i + 1 < n ? lut[4 * (i + 1) + c] - v : 0.f;
        out[c] = fmaf(a, d, v);
    }
}

static void check_brick_rule()
{
    for (int it = 0; it < 20000; ++it) {
        const int n = 1 + irand(40);
        std::vector<float> lut(4 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            for (int c = 0; c < 3; ++c) lut[4 * i + c] = (float)urand();
            lut[4 * i + 3] = irand(3) ? 0.f : (float)urand();
        }
        if (irand(6) == 0) lut[4 * irand(n) + irand(4)] = irand(2) ? NAN : INFINITY;
        if (irand(8) == 0) lut[4 * irand(n) + irand(4)] = 3e38f;
        std::vector<int> nz((size_t)n + 1);
        lut_prefix_counts(lut.data(), n, nz.data());
        // the prefix counts against a direct count
        int cnt = 0;
        for (int i = 0; i < n; ++i) {
            const int j = i + 1 < n ? i + 1 : i;
            bool clear = lut[4 * i + 3] == 0.f;
            for (int c = 0; c < 3; ++c) clear = clear && fabsf(lut[4 * i + c]) <= 1e38f && fabsf(lut[4 * j + c]) <= 1e38f;
            clear = clear && fabsf(lut[4 * j + 3]) <= 1e38f;
            cnt += clear ? 0 : 1;
            CHECK(nz[i + 1] == cnt, "prefix count %d: %d want %d", i, nz[i + 1], cnt);
        }
        float lo = (float)(1.4 * urand() - 0.2), hi = lo + (float)(0.3 * urand() * urand());
        if (irand(20) == 0) { lo = -INFINITY; hi = INFINITY; }
        if (!brick_clear_range(lo, hi, nz.data(), n)) continue;
        for (int k = 0; k < 64; ++k) {
            float rho = k == 0 ? lo : k == 1 ? hi : k == 2 ? nextafterf(lo, -INFINITY)
                      : k == 3 ? nextafterf(hi, INFINITY) : (float)(lo + (hi - lo) * urand());
            if (!isfinite(rho)) rho = (float)(2 * urand() - 0.5);
            float q[4];
            lut_read(lut, n, rho, q);
            CHECK(q[3] == 0.f && isfinite(q[0]) && isfinite(q[1]) && isfinite(q[2]),
                  "clear brick [%g, %g], N=%d, rho=%g reads alpha %g", lo, hi, n, rho, q[3]);
        }
    }
}

static void check_inverse()
{
    for (int it = 0; it < 2000; ++it) {
        double m[16], inv[16];
        for (double &x : m) x = 4 * urand() - 2;
        if (!inverse4(m, inv)) continue;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) {
                double s = 0;
                for (int k = 0; k < 4; ++k) s += m[4 * r + k] * inv[4 * k + c];
                CHECK(fabs(s - (r == c)) < 1e-8, "M * inv(M) [%d][%d] = %g", r, c, s);
            }
    }
    double z[16] = {0}, inv[16];
    CHECK(!inverse4(z, inv), "zero matrix inverted");
    double rank3[16] = {1, 2, 3, 4, 2, 4, 6, 8, 0, 1, 0, 1, 1, 0, 0, 1};
    CHECK(!inverse4(rank3, inv), "singular matrix inverted");
}

static void check_order(int ntx, int nty, int max_steps)
{
    const int nblocks = ntx * nty, head_rows = nty >> 3 << 3, cap = (head_rows >> 3) * ntx;
    const int head = head_rows * ntx;
    std::vector<uint32_t> cnt(8 * kOrderBuckets, 0u), slots((size_t)8 * kOrderBuckets * (cap > 0 ? cap : 1), ~0u);
    std::vector<int> key(nblocks), perm(nblocks);
    for (int t = 0; t < nblocks; ++t) {
        key[t] = irand(4) == 0 ? 0 : irand(max_steps + 1);
        perm[t] = t;
    }
    for (int t = nblocks - 1; t > 0; --t) {   // blocks file themselves in any order
        const int u = irand(t + 1), x = perm[t];
        perm[t] = perm[u];
        perm[u] = x;
    }
    for (int t : perm) {
        const int ty = t / ntx;
        if (ty >= head_rows) continue;
        const int q = (ty & 7) * kOrderBuckets + order_bucket(key[t], max_steps);
        const uint32_t s = cnt[q]++;
        CHECK(s < (uint32_t)cap, "ntx=%d nty=%d: class bucket over capacity", ntx, nty);
        if (s < (uint32_t)cap) slots[(size_t)q * cap + s] = (uint32_t)t;
    }
    std::vector<int> seen(nblocks, 0);
    std::vector<int> last_bucket(8, -1);
    for (int b = 0; b < nblocks; ++b) {
        const int l = ordered_block(b, head, cnt.data(), slots.data(), cap);
        CHECK(l >= 0 && l < nblocks, "ntx=%d nty=%d b=%d: block %d out of range", ntx, nty, b, l);
        if (l < 0 || l >= nblocks) continue;
        ++seen[l];
        if (b < head) {
            CHECK((l / ntx) % 8 == (b & 7), "ntx=%d nty=%d b=%d: row %d not in class %d", ntx, nty, b, l / ntx, b & 7);
            const int k = order_bucket(key[l], max_steps);
            CHECK(k >= last_bucket[b & 7], "ntx=%d nty=%d b=%d: bucket order", ntx, nty, b);
            last_bucket[b & 7] = k;
        } else {
            CHECK(l == b, "ntx=%d nty=%d b=%d: tail block moved", ntx, nty, b);
        }
    }
    for (int t = 0; t < nblocks; ++t) CHECK(seen[t] == 1, "ntx=%d nty=%d: block %d taken %d times", ntx, nty, t, seen[t]);
    // buckets: 0 holds the largest keys; keys <= 0 the last bucket
    CHECK(order_bucket(0, max_steps) == kOrderBuckets - 1 &&
              (max_steps < kOrderBuckets - 1 || order_bucket(max_steps, max_steps) == 0),
          "bucket ends (max_steps %d)", max_steps);
    for (int k = 1; k <= max_steps; k += 1 + max_steps / 97)
        CHECK(order_bucket(k, max_steps) <= order_bucket(k - 1, max_steps), "bucket monotonicity at %d", k);
}

int main()
{
    for (int i = 0; i < 300; ++i) check_order(1 + irand(40), 1 + irand(40), 1 + irand(2000));
    check_order(64, 64, 1000);
    check_order(128, 128, 1000);
    for (int lb = 2; lb <= 4; ++lb)
        for (int s = 1; s <= 70; ++s) check_layout(s, lb);
    check_brick_rule();
    check_inverse();
    if (fails) return 1;
    printf("rules ok\n");
    return 0;
}
