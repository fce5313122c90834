// hvr_device.h -- device code of libhipvr: texel formats, GL-style sampling,
// the ray-setup kernel and the march-and-composite kernel (gfx950, wave64).
//
// Semantics follow the reference fragment shader pyvr/shaders/volume.frag.glsl
// and the GL state of pyvr/moderngl_renderer/manager.py (paths relative to the
// reference repository); each block names the lines it implements.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hvr {

// ---------------------------------------------------------------------------
// Texel formats.  A volume stores one texel per voxel; a texel holds the
// scalar field, the normal field, or both (interleaved so that one load per
// trilinear tap returns everything a sample needs).
//   S_*  scalar only         N_*  normals only        C_*  both
// Decoding yields (nx, ny, nz, rho).  For scalar-only texels the normal part
// reads (rho, 0, 0), which is what a 3-channel GL sampler returns from a
// 1-channel texture; for normal-only texels rho reads nx (the .r channel).
// SNORM8 normals are kept as raw integers in the sampling fast paths: the
// normal is normalised afterwards, so the 1/127 scale cancels.
enum TexFmt : int {
    TF_NONE = 0,
    TF_S_F32 = 1,      // float                                    4 B
    TF_S_F16 = 2,      // half                                     2 B
    TF_S_U8 = 3,       // unorm8                                   1 B
    TF_N_F32 = 4,      // float4 (nx, ny, nz, 0)                  16 B
    TF_N_SN8 = 5,      // u32 nx | ny<<8 | nz<<16                  4 B
    TF_C_F32_F32 = 6,  // float4 (nx, ny, nz, rho)                16 B
    TF_C_F16_F32 = 7,  // float4 (nx, ny, nz, bits: half rho)     16 B
    TF_C_U8_F32 = 8,   // float4 (nx, ny, nz, bits: u8 rho)       16 B
    TF_C_F32_SN8 = 9,  // u32x2 (float rho, sn8 xyz)               8 B
    TF_C_F16_SN8 = 10, // u32x2 (half rho, sn8 xyz)                8 B
    TF_C_U8_SN8 = 11,  // u32 sn8 xyz | u8 rho << 24               4 B
    TF_COUNT = 12
};

__host__ __device__ constexpr int texel_bytes(int f)
{
    return f == TF_S_F32 ? 4 : f == TF_S_F16 ? 2 : f == TF_S_U8 ? 1
         : f == TF_N_F32 ? 16 : f == TF_N_SN8 ? 4
         : (f == TF_C_F32_F32 || f == TF_C_F16_F32 || f == TF_C_U8_F32) ? 16
         : (f == TF_C_F32_SN8 || f == TF_C_F16_SN8) ? 8 : f == TF_C_U8_SN8 ? 4 : 0;
}
__host__ __device__ constexpr bool fmt_has_scalar(int f)
{
    return f == TF_S_F32 || f == TF_S_F16 || f == TF_S_U8 || f >= TF_C_F32_F32;
}
__host__ __device__ constexpr bool fmt_has_normal(int f)
{
    return f == TF_N_F32 || f == TF_N_SN8 || f >= TF_C_F32_F32;
}
__host__ __device__ constexpr bool fmt_sn8_normal(int f)
{
    return f == TF_N_SN8 || f == TF_C_F32_SN8 || f == TF_C_F16_SN8 || f == TF_C_U8_SN8;
}
__host__ __device__ constexpr bool fmt_u8_scalar(int f)
{
    return f == TF_S_U8 || f == TF_C_U8_F32 || f == TF_C_U8_SN8;
}

// Layout of a device texture: GL size (s0, s1, s2) (x fastest), stored in
// cubic bricks of edge 2^lb (lb = 0: plain linear layout).  Texel (x, y, z)
// lives at offx(x) + offy(y) + offz(z), see brick_off_*.
struct TexDesc {
    const void *base;
    int s0, s1, s2;
    float fs0, fs1, fs2;
    int lb;
    uint32_t bsy, bsz;  // texels per brick row (nbx*B^3) and per brick slab
    int fmt;
    int role;           // field a sampler reads: 0 scalar, 1 normal
};

__device__ __forceinline__ uint32_t brick_off_x(int i, int lb)
{
    const int m = (1 << lb) - 1;
    return ((uint32_t)(i >> lb) << (3 * lb)) + (uint32_t)(i & m);
}
__device__ __forceinline__ uint32_t brick_off_y(int i, int lb, uint32_t bsy)
{
    const int m = (1 << lb) - 1;
    return (uint32_t)(i >> lb) * bsy + ((uint32_t)(i & m) << lb);
}
__device__ __forceinline__ uint32_t brick_off_z(int i, int lb, uint32_t bsz)
{
    const int m = (1 << lb) - 1;
    return (uint32_t)(i >> lb) * bsz + ((uint32_t)(i & m) << (2 * lb));
}

__device__ __forceinline__ float sn8f(uint32_t w, int sh) { return (float)(int8_t)((w >> sh) & 0xffu); }

// Decode one texel to (nx, ny, nz, rho); SNORM8 normals as raw integers,
// UNORM8 scalars as raw integers (0..255).
template <int F>
__device__ __forceinline__ float4 fetch4(const void *base, uint32_t idx)
{
    if constexpr (F == TF_S_F32) {
        float v = ((const float *)base)[idx];
        return make_float4(v, 0.f, 0.f, v);
    } else if constexpr (F == TF_S_F16) {
        float v = __half2float(((const __half *)base)[idx]);
        return make_float4(v, 0.f, 0.f, v);
    } else if constexpr (F == TF_S_U8) {
        float v = (float)((const uint8_t *)base)[idx];
        return make_float4(v, 0.f, 0.f, v);
    } else if constexpr (F == TF_N_F32) {
        float4 t = ((const float4 *)base)[idx];
        return make_float4(t.x, t.y, t.z, t.x);
    } else if constexpr (F == TF_N_SN8) {
        uint32_t w = ((const uint32_t *)base)[idx];
        float x = sn8f(w, 0);
        return make_float4(x, sn8f(w, 8), sn8f(w, 16), x);
    } else if constexpr (F == TF_C_F32_F32) {
        return ((const float4 *)base)[idx];
    } else if constexpr (F == TF_C_F16_F32) {
        float4 t = ((const float4 *)base)[idx];
        __half h = __ushort_as_half((unsigned short)(__float_as_uint(t.w) & 0xffffu));
        return make_float4(t.x, t.y, t.z, __half2float(h));
    } else if constexpr (F == TF_C_U8_F32) {
        float4 t = ((const float4 *)base)[idx];
        return make_float4(t.x, t.y, t.z, (float)(__float_as_uint(t.w) & 0xffu));
    } else if constexpr (F == TF_C_F32_SN8) {
        uint2 t = ((const uint2 *)base)[idx];
        return make_float4(sn8f(t.y, 0), sn8f(t.y, 8), sn8f(t.y, 16), __uint_as_float(t.x));
    } else if constexpr (F == TF_C_F16_SN8) {
        uint2 t = ((const uint2 *)base)[idx];
        __half h = __ushort_as_half((unsigned short)(t.x & 0xffffu));
        return make_float4(sn8f(t.y, 0), sn8f(t.y, 8), sn8f(t.y, 16), __half2float(h));
    } else if constexpr (F == TF_C_U8_SN8) {
        uint32_t w = ((const uint32_t *)base)[idx];
        return make_float4(sn8f(w, 0), sn8f(w, 8), sn8f(w, 16), (float)(w >> 24));
    } else {
        return make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Scalar-only decode (avoids interpolating unused channels).
template <int F>
__device__ __forceinline__ float fetch1(const void *base, uint32_t idx)
{
    if constexpr (F == TF_S_F32) return ((const float *)base)[idx];
    else if constexpr (F == TF_S_F16) return __half2float(((const __half *)base)[idx]);
    else if constexpr (F == TF_S_U8) return (float)((const uint8_t *)base)[idx];
    else return fetch4<F>(base, idx).w;
}

// Runtime-format decode with exact GL scaling (generic path).
__device__ __forceinline__ float4 fetch_any(int fmt, const void *base, uint32_t idx)
{
    float4 t;
    switch (fmt) {
    case TF_S_F32: t = fetch4<TF_S_F32>(base, idx); break;
    case TF_S_F16: t = fetch4<TF_S_F16>(base, idx); break;
    case TF_S_U8: t = fetch4<TF_S_U8>(base, idx); break;
    case TF_N_F32: t = fetch4<TF_N_F32>(base, idx); break;
    case TF_N_SN8: t = fetch4<TF_N_SN8>(base, idx); break;
    case TF_C_F32_F32: t = fetch4<TF_C_F32_F32>(base, idx); break;
    case TF_C_F16_F32: t = fetch4<TF_C_F16_F32>(base, idx); break;
    case TF_C_U8_F32: t = fetch4<TF_C_U8_F32>(base, idx); break;
    case TF_C_F32_SN8: t = fetch4<TF_C_F32_SN8>(base, idx); break;
    case TF_C_F16_SN8: t = fetch4<TF_C_F16_SN8>(base, idx); break;
    case TF_C_U8_SN8: t = fetch4<TF_C_U8_SN8>(base, idx); break;
    default: return make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const float k127 = 1.0f / 127.0f, k255 = 1.0f / 255.0f;
    if (fmt_sn8_normal(fmt)) {
        t.x = fmaxf(t.x * k127, -1.f);
        t.y = fmaxf(t.y * k127, -1.f);
        t.z = fmaxf(t.z * k127, -1.f);
        if (fmt == TF_N_SN8) t.w = t.x;
    }
    if (fmt_u8_scalar(fmt)) {
        t.w *= k255;
        if (fmt == TF_S_U8) t.x = t.w;
    }
    return t;
}

// ---------------------------------------------------------------------------
// Trilinear sampling (GL_LINEAR, CLAMP_TO_EDGE, no mipmaps;
// manager.py:95-101, 123-129).  Coordinate (cx, cy, cz) in [0, 1]:
// u = c*size - 1/2 per axis, taps floor(u) and floor(u)+1 clamped to
// [0, size-1], weights frac(u).
struct Taps {
    uint32_t i000, i100, i010, i110, i001, i101, i011, i111;
    float a, b, g;
};

__device__ __forceinline__ Taps make_taps(const TexDesc &t, float cx, float cy, float cz)
{
    float u = fmaf(cx, t.fs0, -0.5f), v = fmaf(cy, t.fs1, -0.5f), w = fmaf(cz, t.fs2, -0.5f);
    float fu = floorf(u), fv = floorf(v), fw = floorf(w);
    Taps k;
    k.a = u - fu;
    k.b = v - fv;
    k.g = w - fw;
    int x0 = (int)fu, y0 = (int)fv, z0 = (int)fw;
    int x1 = min(x0 + 1, t.s0 - 1), y1 = min(y0 + 1, t.s1 - 1), z1 = min(z0 + 1, t.s2 - 1);
    x0 = max(x0, 0);
    y0 = max(y0, 0);
    z0 = max(z0, 0);
    uint32_t ox0 = brick_off_x(x0, t.lb), ox1 = brick_off_x(x1, t.lb);
    uint32_t oy0 = brick_off_y(y0, t.lb, t.bsy), oy1 = brick_off_y(y1, t.lb, t.bsy);
    uint32_t oz0 = brick_off_z(z0, t.lb, t.bsz), oz1 = brick_off_z(z1, t.lb, t.bsz);
    uint32_t y0z0 = oy0 + oz0, y1z0 = oy1 + oz0, y0z1 = oy0 + oz1, y1z1 = oy1 + oz1;
    k.i000 = ox0 + y0z0;
    k.i100 = ox1 + y0z0;
    k.i010 = ox0 + y1z0;
    k.i110 = ox1 + y1z0;
    k.i001 = ox0 + y0z1;
    k.i101 = ox1 + y0z1;
    k.i011 = ox0 + y1z1;
    k.i111 = ox1 + y1z1;
    return k;
}

__device__ __forceinline__ float lerpf(float x, float y, float t) { return fmaf(t, y - x, x); }
__device__ __forceinline__ float4 lerp4(float4 x, float4 y, float t)
{
    return make_float4(lerpf(x.x, y.x, t), lerpf(x.y, y.y, t), lerpf(x.z, y.z, t), lerpf(x.w, y.w, t));
}

template <int F>
__device__ __forceinline__ float4 sample4(const TexDesc &t, const Taps &k)
{
    float4 c000 = fetch4<F>(t.base, k.i000), c100 = fetch4<F>(t.base, k.i100);
    float4 c010 = fetch4<F>(t.base, k.i010), c110 = fetch4<F>(t.base, k.i110);
    float4 c001 = fetch4<F>(t.base, k.i001), c101 = fetch4<F>(t.base, k.i101);
    float4 c011 = fetch4<F>(t.base, k.i011), c111 = fetch4<F>(t.base, k.i111);
    float4 e0 = lerp4(c000, c100, k.a), e1 = lerp4(c010, c110, k.a);
    float4 e2 = lerp4(c001, c101, k.a), e3 = lerp4(c011, c111, k.a);
    return lerp4(lerp4(e0, e1, k.b), lerp4(e2, e3, k.b), k.g);
}

template <int F>
__device__ __forceinline__ float sample1(const TexDesc &t, const Taps &k)
{
    float c000 = fetch1<F>(t.base, k.i000), c100 = fetch1<F>(t.base, k.i100);
    float c010 = fetch1<F>(t.base, k.i010), c110 = fetch1<F>(t.base, k.i110);
    float c001 = fetch1<F>(t.base, k.i001), c101 = fetch1<F>(t.base, k.i101);
    float c011 = fetch1<F>(t.base, k.i011), c111 = fetch1<F>(t.base, k.i111);
    float e0 = lerpf(c000, c100, k.a), e1 = lerpf(c010, c110, k.a);
    float e2 = lerpf(c001, c101, k.a), e3 = lerpf(c011, c111, k.a);
    return lerpf(lerpf(e0, e1, k.b), lerpf(e2, e3, k.b), k.g);
}

__device__ __forceinline__ float4 sample_any(const TexDesc &t, const Taps &k)
{
    float4 c000 = fetch_any(t.fmt, t.base, k.i000), c100 = fetch_any(t.fmt, t.base, k.i100);
    float4 c010 = fetch_any(t.fmt, t.base, k.i010), c110 = fetch_any(t.fmt, t.base, k.i110);
    float4 c001 = fetch_any(t.fmt, t.base, k.i001), c101 = fetch_any(t.fmt, t.base, k.i101);
    float4 c011 = fetch_any(t.fmt, t.base, k.i011), c111 = fetch_any(t.fmt, t.base, k.i111);
    float4 e0 = lerp4(c000, c100, k.a), e1 = lerp4(c010, c110, k.a);
    float4 e2 = lerp4(c001, c101, k.a), e3 = lerp4(c011, c111, k.a);
    return lerp4(lerp4(e0, e1, k.b), lerp4(e2, e3, k.b), k.g);
}

// ---------------------------------------------------------------------------
// Per-frame constants (uniforms after host-side FP64 preparation).
struct FrameConst {
    float prow[8];    // rows 0-1 of inverse(P_gl)          (volume.frag.glsl:50)
    float rot[9];     // 3x3 block of inverse(V_gl)         (volume.frag.glsl:52)
    float o[3];       // camera_pos
    float bmin[3], bmax[3], ext[3];   // ext = bmax - bmin in float (:26)
    float step;       // step_size
    float alpha_k;    // step_size / reference_step_size    (:101)
    float ka, kd;     // ambient_light, diffuse_light       (:110)
    float L[3];       // normalize(light_target - light_position) (:107)
    float tau;        // early-termination alpha            (:87)
    float fW, fH;
    int W, H;
    int max_steps;
    int ok;           // 0: camera unset or singular matrices -> every pixel 0
    int L_ok;         // 0: zero-length light vector -> diffuse term 0
    uint32_t clear[4];  // framebuffer clear colour (8-bit per channel)
    int blend;          // 1: SRC_ALPHA, ONE_MINUS_SRC_ALPHA over the clear colour
};

// One ray per pixel, produced by the setup kernel.
struct __align__(16) RayRec {
    float p0x, p0y, p0z;  // first sample position (entry point or camera)
    int nlim;             // iterations to run (0: the ray misses the box)
    float dsx, dsy, dsz;  // ray_dir * step_size
    int pad;
};

// K1 -- ray setup (volume.vert.glsl:8; volume.frag.glsl:47-68, 71-82).
// Pixel (px, py), py = 0 the bottom row.  Positions use separately rounded
// products and sums (no fused multiply-add), as the shader states them.
__global__ void __launch_bounds__(256) k_ray_setup(const FrameConst fc, RayRec *__restrict__ rays)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= fc.W * fc.H) return;
    const int px = idx % fc.W, py = idx / fc.W;
    RayRec r;
    r.p0x = r.p0y = r.p0z = 0.f;
    r.dsx = r.dsy = r.dsz = 0.f;
    r.nlim = 0;
    r.pad = 0;
    if (fc.ok) {
        // A1: pixel centre -> NDC
        const float ux = ((float)px + 0.5f) / fc.fW, uy = ((float)py + 0.5f) / fc.fH;
        const float nx = ux * 2.f - 1.f, ny = uy * 2.f - 1.f;
        // A2: eye direction: inverse(P) * (ndc, -1, 1), then (x, y, -1, 0)
        const float ex = fc.prow[0] * nx + fc.prow[1] * ny - fc.prow[2] + fc.prow[3];
        const float ey = fc.prow[4] * nx + fc.prow[5] * ny - fc.prow[6] + fc.prow[7];
        // A3: world direction
        const float w0 = fc.rot[0] * ex + fc.rot[1] * ey - fc.rot[2];
        const float w1 = fc.rot[3] * ex + fc.rot[4] * ey - fc.rot[5];
        const float w2 = fc.rot[6] * ex + fc.rot[7] * ey - fc.rot[8];
        const float wn = 1.0f / sqrtf(w0 * w0 + w1 * w1 + w2 * w2);
        const float d[3] = {w0 * wn, w1 * wn, w2 * wn};
        // A4: slab test
        float tn = -INFINITY, tf = INFINITY;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float inv = 1.0f / d[k];
            const float t0 = (fc.bmin[k] - fc.o[k]) * inv, t1 = (fc.bmax[k] - fc.o[k]) * inv;
            float lo = fminf(t0, t1), hi = fmaxf(t0, t1);
            if (t0 != t0 || t1 != t1) { lo = -INFINITY; hi = INFINITY; }
            tn = fmaxf(tn, lo);
            tf = fminf(tf, hi);
        }
        if (tn <= tf && tf > 0.f) {
            tn = fmaxf(tn, 0.f);
            // A5: entry point and step vector
            r.p0x = __fadd_rn(fc.o[0], __fmul_rn(d[0], tn));
            r.p0y = __fadd_rn(fc.o[1], __fmul_rn(d[1], tn));
            r.p0z = __fadd_rn(fc.o[2], __fmul_rn(d[2], tn));
            r.dsx = __fmul_rn(d[0], fc.step);
            r.dsy = __fmul_rn(d[1], fc.step);
            r.dsz = __fmul_rn(d[2], fc.step);
            // Iterations: every in-box sample lies within 1.0625*(tf-tn)/step
            // + 16 steps of the entry point (margin for FP32 stepping drift);
            // later iterations of the shader's loop are out-of-box no-ops.
            const float span = (tf - tn) / fc.step * 1.0625f + 16.f;
            r.nlim = span < (float)fc.max_steps ? (int)span : fc.max_steps;
        }
    }
    rays[idx] = r;
}

// Stored RGBA8 (manager.py:29-31, 212-220): the source colour is clamped
// and converted to 8 bits (Cq, Aq), then blended with (SRC_ALPHA,
// ONE_MINUS_SRC_ALPHA) over the 8-bit clear colour D in fixed point:
// out = round((S*Aq + D*(255 - Aq)) / 255) per channel.  With the
// reference's clear colour 0 this is rgb = Cq*Aq/255, a = Aq*Aq/255.
__device__ __forceinline__ uint32_t unorm8(float x)
{
    x = fminf(fmaxf(x, 0.f), 1.f);  // NaN -> 0
    return (uint32_t)floorf(fmaf(x, 255.f, 0.5f));
}
__device__ __forceinline__ uchar4 blend_store(const FrameConst &fc, float c0, float c1, float c2, float a)
{
    const uint32_t aq = unorm8(a);
    const uint32_t q0 = unorm8(c0), q1 = unorm8(c1), q2 = unorm8(c2);
    if (!fc.blend) return make_uchar4((unsigned char)q0, (unsigned char)q1, (unsigned char)q2, (unsigned char)aq);
    const uint32_t ia = 255u - aq;
    return make_uchar4((unsigned char)((q0 * aq + fc.clear[0] * ia + 127u) / 255u),
                       (unsigned char)((q1 * aq + fc.clear[1] * ia + 127u) / 255u),
                       (unsigned char)((q2 * aq + fc.clear[2] * ia + 127u) / 255u),
                       (unsigned char)((aq * aq + fc.clear[3] * ia + 127u) / 255u));
}

// How the normal sampler is served in a march kernel instance.
enum NormalMode : int {
    NM_SAME = 0,   // normals interleaved in the scalar volume's texels
    NM_ALIAS = 1,  // scalar-only volume: normal = (rho, 0, 0)
    NM_ZERO = 2,   // unbound: normal = 0 -> diffuse term 0
    NM_ANY = 3     // generic: scalar and normal from any volumes/formats
};

// LUT lookup at (rho, 0.5) with u = rho*N - 1/2, clamp to edge
// (manager.py:163-181; volume.frag.glsl:96).
template <typename LutPtr>
__device__ __forceinline__ float4 lut_sample(LutPtr lut, int n, float fn, float rho)
{
    float u = fminf(fmaxf(fmaf(rho, fn, -0.5f), -1.f), fn);
    const float fu = floorf(u);
    const float a = u - fu;
    const int i0 = (int)fu;
    const int j0 = min(max(i0, 0), n - 1), j1 = min(max(i0 + 1, 0), n - 1);
    return lerp4(lut[j0], lut[j1], a);
}

// K2 -- march and composite (volume.frag.glsl:84-121), one thread per ray.
// A 256-thread block covers 16x16 pixels; each wave64 an 8x8 tile.  Blocks
// are spread so that each of the 8 XCDs gets a contiguous band of tiles (its
// own L2 sees a compact part of the volume).  A wave runs while any of its
// rays is live; a ray stops at max_steps / its iteration bound, when its
// alpha reaches tau, or when it leaves the box.
template <int SF, int NM, bool COUNT>
__global__ void __launch_bounds__(256)
k_march(const FrameConst fc, const TexDesc ts, const TexDesc tn, const float4 *__restrict__ lut_g,
        int lut_n, const RayRec *__restrict__ rays, uchar4 *__restrict__ out,
        uint32_t *__restrict__ counts, int ntx, int nblocks)
{
    extern __shared__ float4 s_lut[];
    if constexpr (NM != NM_ANY) {
        for (int i = threadIdx.x; i < lut_n; i += 256) s_lut[i] = lut_g[i];
        __syncthreads();
    }
    const int b = blockIdx.x;
    const int per = nblocks >> 3, full = per << 3;
    const int logical = b < full ? (b & 7) * per + (b >> 3) : b;
    const int tx = logical % ntx, ty = logical / ntx;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int px = tx * 16 + (wv & 1) * 8 + (lane & 7);
    const int py = ty * 16 + (wv >> 1) * 8 + (lane >> 3);
    if (px >= fc.W || py >= fc.H) return;
    const int pix = py * fc.W + px;
    const RayRec r = rays[pix];

    float x = r.p0x, y = r.p0y, z = r.p0z;
    float C0 = 0.f, C1 = 0.f, C2 = 0.f, A = 0.f;
    uint32_t ns = 0;
    bool entered = false;
    const float fn = (float)lut_n;
    const float u8k = fmt_u8_scalar(SF) ? 1.0f / 255.0f : 1.0f;
    for (int i = 0; i < r.nlim; ++i) {
        if (!(A < fc.tau)) break;
        // A6: normalised coordinate and inclusive [0,1] test on (z, y, x)
        const float qx = (x - fc.bmin[0]) / fc.ext[0];
        const float qy = (y - fc.bmin[1]) / fc.ext[1];
        const float qz = (z - fc.bmin[2]) / fc.ext[2];
        const bool valid = qz >= 0.f && qz <= 1.f && qy >= 0.f && qy <= 1.f && qx >= 0.f && qx <= 1.f;
        if (valid) {
            entered = true;
            if constexpr (COUNT) ++ns;
            // texture x <- q.z (size s0), y <- q.y (s1), z <- q.x (s2)
            float rho, n0, n1, n2;
            if constexpr (NM == NM_ANY) {
                rho = 0.f;
                if (ts.base) {
                    const float4 t = sample_any(ts, make_taps(ts, qz, qy, qx));
                    rho = ts.role == 0 ? t.w : t.x;
                }
                n0 = n1 = n2 = 0.f;
                if (tn.base) {
                    const float4 t = sample_any(tn, make_taps(tn, qz, qy, qx));
                    if (tn.role == 1) { n0 = t.x; n1 = t.y; n2 = t.z; }
                    else n0 = t.w;
                }
            } else if constexpr (SF == TF_NONE) {
                rho = 0.f;
                n0 = n1 = n2 = 0.f;
            } else {
                const Taps k = make_taps(ts, qz, qy, qx);
                if constexpr (NM == NM_SAME) {
                    const float4 t = sample4<SF>(ts, k);
                    rho = t.w * u8k;
                    n0 = t.x; n1 = t.y; n2 = t.z;
                } else {
                    rho = sample1<SF>(ts, k) * u8k;
                    n0 = NM == NM_ALIAS ? rho : 0.f;
                    n1 = n2 = 0.f;
                }
            }
            float4 q;
            if constexpr (NM == NM_ANY) q = lut_sample(lut_g, lut_n, fn, rho);
            else q = lut_sample(s_lut, lut_n, fn, rho);
            // Beer-Lambert opacity correction (:101)
            const float a = 1.f - __expf(-(q.w * fc.alpha_k));
            // Lambert term with the zero-vector rule: a normal of zero (or
            // subnormal) length has no direction and gives no diffuse light.
            float diff = 0.f;
            if constexpr (NM != NM_ZERO) {
                const float nn = n0 * n0 + n1 * n1 + n2 * n2;
                if (fc.L_ok && nn >= 1.17549435e-38f) {
                    const float rs = rsqrtf(nn);
                    const float dt = (n0 * fc.L[0] + n1 * fc.L[1] + n2 * fc.L[2]) * rs;
                    diff = fmaxf(dt, 0.f);
                }
            }
            const float lam = fmaf(fc.kd, diff, fc.ka);
            // front-to-back composite (:112-115)
            const float ta = (1.f - A) * a;
            C0 = fmaf(ta, q.x * lam, C0);
            C1 = fmaf(ta, q.y * lam, C1);
            C2 = fmaf(ta, q.z * lam, C2);
            A = A + ta;
        } else if (entered) {
            break;
        }
        // A5: incremental FP32 stepping (:118)
        x = __fadd_rn(x, r.dsx);
        y = __fadd_rn(y, r.dsy);
        z = __fadd_rn(z, r.dsz);
    }
    out[pix] = blend_store(fc, C0, C1, C2, A);
    if constexpr (COUNT) counts[pix] = ns;
}

// K3 -- upload: convert a slab of host float32 data (numpy C order = GL
// texel order, x fastest) into the volume's texel format and brick layout.
// part: 1 = scalar field, 2 = normal field (3 floats per voxel).  The other
// part of an interleaved texel is preserved.
__device__ __forceinline__ uint32_t enc_sn8(float v)
{
    v = fminf(fmaxf(v, -1.f), 1.f);  // NaN -> -1 .. kept finite
    const int q = (int)rintf(v * 127.f);
    return (uint32_t)(uint8_t)(int8_t)q;
}
__device__ __forceinline__ uint32_t enc_u8(float v)
{
    v = fminf(fmaxf(v, 0.f), 1.f);
    return (uint32_t)rintf(v * 255.f);
}

__global__ void __launch_bounds__(256)
k_upload(void *__restrict__ dst, int fmt, int s0, int s1, int lb, uint32_t bsy, uint32_t bsz,
         const float *__restrict__ src, int part, int z0, uint64_t nvox)
{
    const uint64_t l = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (l >= nvox) return;
    const uint64_t plane = (uint64_t)s0 * (uint64_t)s1;
    const int z = (int)(l / plane) + z0;
    const uint64_t rem = l % plane;
    const int y = (int)(rem / (uint64_t)s0), x = (int)(rem % (uint64_t)s0);
    const uint32_t idx = brick_off_x(x, lb) + brick_off_y(y, lb, bsy) + brick_off_z(z, lb, bsz);
    if (part == 1) {
        const float v = src[l];
        switch (fmt) {
        case TF_S_F32: ((float *)dst)[idx] = v; break;
        case TF_S_F16: ((__half *)dst)[idx] = __float2half_rn(v); break;
        case TF_S_U8: ((uint8_t *)dst)[idx] = (uint8_t)enc_u8(v); break;
        case TF_C_F32_F32: ((float *)dst)[4 * (uint64_t)idx + 3] = v; break;
        case TF_C_F16_F32:
            ((uint32_t *)dst)[4 * (uint64_t)idx + 3] = (uint32_t)__half_as_ushort(__float2half_rn(v));
            break;
        case TF_C_U8_F32: ((uint32_t *)dst)[4 * (uint64_t)idx + 3] = enc_u8(v); break;
        case TF_C_F32_SN8: ((float *)dst)[2 * (uint64_t)idx] = v; break;
        case TF_C_F16_SN8:
            ((uint32_t *)dst)[2 * (uint64_t)idx] = (uint32_t)__half_as_ushort(__float2half_rn(v));
            break;
        case TF_C_U8_SN8: ((uint8_t *)dst)[4 * (uint64_t)idx + 3] = (uint8_t)enc_u8(v); break;
        default: break;
        }
    } else {
        const float nx = src[3 * l], ny = src[3 * l + 1], nz = src[3 * l + 2];
        switch (fmt) {
        case TF_N_F32:
        case TF_C_F32_F32:
        case TF_C_F16_F32:
        case TF_C_U8_F32: {
            float *p = (float *)dst + 4 * (uint64_t)idx;
            p[0] = nx; p[1] = ny; p[2] = nz;
            if (fmt == TF_N_F32) p[3] = 0.f;
            break;
        }
        case TF_N_SN8:
            ((uint32_t *)dst)[idx] = enc_sn8(nx) | (enc_sn8(ny) << 8) | (enc_sn8(nz) << 16);
            break;
        case TF_C_F32_SN8:
        case TF_C_F16_SN8:
            ((uint32_t *)dst)[2 * (uint64_t)idx + 1] = enc_sn8(nx) | (enc_sn8(ny) << 8) | (enc_sn8(nz) << 16);
            break;
        case TF_C_U8_SN8: {
            uint8_t *p = (uint8_t *)dst + 4 * (uint64_t)idx;
            p[0] = (uint8_t)enc_sn8(nx); p[1] = (uint8_t)enc_sn8(ny); p[2] = (uint8_t)enc_sn8(nz);
            break;
        }
        default: break;
        }
    }
}

}  // namespace hvr
