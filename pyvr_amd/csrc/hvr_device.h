// hvr_device.h -- device code of libhipvr: texel formats, GL-style sampling,
// the ray-setup kernel, the march-and-composite kernel and the upload kernel
// (gfx950, wave64).
//
// Semantics follow the reference fragment shader pyvr/shaders/volume.frag.glsl
// and the GL state of pyvr/moderngl_renderer/manager.py (paths relative to the
// reference repository); each block names the lines it implements.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hvr_rules.h"

namespace hvr {

// ---------------------------------------------------------------------------
// Texel formats.  A volume stores one texel per voxel; a texel holds the
// scalar field, the normal field, or both (interleaved so that one load per
// trilinear tap returns everything a sample needs).
//   S_*  scalar only         N_*  normals only        C_*  both
// Decoding yields (nx, ny, nz, rho).  For scalar-only texels the normal part
// reads (rho, 0, 0), which is what a 3-channel GL sampler returns from a
// 1-channel texture; for normal-only texels rho reads nx (the .r channel).
// In the fast paths SNORM8 normals stay raw integers (the normal is
// normalised afterwards, so the 1/127 scale cancels) and UNORM8 scalars are
// scaled by 1/255 after interpolation (filtering is linear).  FP16 scalars
// and normals are interpolated straight from their halves (v_fma_mix_f32:
// float32 accumulation of exactly converted halves), with no conversion
// instructions.
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
    TF_C_F16_SN8 = 10, // u32x2 (half rho | half qx << 16, half qy | half qz << 16), q the
                       // SNORM8 integers in [-127, 127], exact as halves  8 B
    TF_C_U8_SN8 = 11,  // u32 sn8 xyz | u8 rho << 24               4 B
    TF_C_F16_F16 = 12, // u32x2 (half rho | half nx << 16, half ny | half nz << 16)  8 B
    TF_N_F16 = 13,     // u32x2 (half nx | half ny << 16, half nz)  8 B
    TF_COUNT = 14
};

__host__ __device__ constexpr int texel_bytes(int f)
{
    return f == TF_S_F32 ? 4 : f == TF_S_F16 ? 2 : f == TF_S_U8 ? 1
         : f == TF_N_F32 ? 16 : f == TF_N_SN8 ? 4
         : (f == TF_C_F32_F32 || f == TF_C_F16_F32 || f == TF_C_U8_F32) ? 16
         : (f == TF_C_F32_SN8 || f == TF_C_F16_SN8) ? 8 : f == TF_C_U8_SN8 ? 4
         : (f == TF_C_F16_F16 || f == TF_N_F16) ? 8 : 0;
}
__host__ __device__ constexpr bool fmt_has_scalar(int f)
{
    return f == TF_S_F32 || f == TF_S_F16 || f == TF_S_U8 || (f >= TF_C_F32_F32 && f <= TF_C_F16_F16);
}
__host__ __device__ constexpr bool fmt_has_normal(int f)
{
    return f == TF_N_F32 || f == TF_N_SN8 || f == TF_N_F16 || f >= TF_C_F32_F32;
}
__host__ __device__ constexpr bool fmt_sn8_normal(int f)
{
    return f == TF_N_SN8 || f == TF_C_F32_SN8 || f == TF_C_F16_SN8 || f == TF_C_U8_SN8;
}
__host__ __device__ constexpr bool fmt_u8_scalar(int f)
{
    return f == TF_S_U8 || f == TF_C_U8_F32 || f == TF_C_U8_SN8;
}

// ---------------------------------------------------------------------------
// Brick layout with apron.  A texture of GL size (s0, s1, s2) is addressed in
// extended coordinates X = x + 1 in [0, s+1] per axis, where X = 0 and
// X = s+1 hold copies of the edge texels (x = 0 and x = s-1): sampling the
// extended texture without clamping is exactly GL's CLAMP_TO_EDGE.  The
// extended texture is cut into bricks of B = 2^lb interior texels per axis,
// each stored with one more texel per axis (its neighbour's first), so a
// brick holds E^3 texels, E = B + 1, x fastest.  A trilinear footprint with
// lower corner X0 lies in brick X0 >> lb at local X0 & (B-1) .. +1: one base
// address per sample, the eight taps at fixed offsets.
struct TexDesc {
    const void *base;
    int s0, s1, s2;        // GL size
    float k0, k1, k2;      // s_axis / world extent of the axis it samples
    int lb;                // log2 of the brick interior edge
    uint32_t nbx, nby;     // bricks along x and y
    uint32_t brick_bytes;  // E^3 * texel bytes
    uint32_t nbytes32;     // allocation bytes when below 4 GiB with fewer than 2^24
                           // bricks (buffer-load path), else 0
    int fmt;
    int role;              // field a sampler reads: 0 scalar, 1 normal
};


// Loads and stores through pointers the compiler cannot place (fetched from
// descriptors in memory): they address device global memory, so issue
// global (not flat) instructions.
template <typename T>
__device__ __forceinline__ T ld_global(const T *p)
{
    return *(const __attribute__((address_space(1))) T *)p;
}
__device__ __forceinline__ float4 ld_global(const float4 *p)
{
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 v = *(const __attribute__((address_space(1))) v4 *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 ld_global(const float2 *p)
{
    typedef float v2 __attribute__((ext_vector_type(2)));
    const v2 v = *(const __attribute__((address_space(1))) v2 *)p;
    return make_float2(v.x, v.y);
}
template <typename T>
__device__ __forceinline__ void st_global(T *p, const T &v)
{
    *(__attribute__((address_space(1))) T *)p = v;
}
__device__ __forceinline__ void st_global(uchar4 *p, const uchar4 &v)
{
    const uint32_t w = (uint32_t)v.x | ((uint32_t)v.y << 8) | ((uint32_t)v.z << 16) | ((uint32_t)v.w << 24);
    *(__attribute__((address_space(1))) uint32_t *)p = w;
}

__device__ __forceinline__ float sn8f(uint32_t w, int sh) { return (float)(int8_t)((w >> sh) & 0xffu); }
// the half in the low 16 bits of w
__device__ __forceinline__ float h16f(uint32_t w) { return __half2float(__ushort_as_half((unsigned short)(w & 0xffffu))); }

// Decode the texel at byte address p to (nx, ny, nz, rho) (raw SNORM8/UNORM8).
template <int F>
__device__ __forceinline__ float4 fetch4(const char *p)
{
    if constexpr (F == TF_S_F32) {
        const float v = *(const float *)p;
        return make_float4(v, 0.f, 0.f, v);
    } else if constexpr (F == TF_S_F16) {
        const float v = __half2float(*(const __half *)p);
        return make_float4(v, 0.f, 0.f, v);
    } else if constexpr (F == TF_S_U8) {
        const float v = (float)*(const uint8_t *)p;
        return make_float4(v, 0.f, 0.f, v);
    } else if constexpr (F == TF_N_F32) {
        const float4 t = *(const float4 *)p;
        return make_float4(t.x, t.y, t.z, t.x);
    } else if constexpr (F == TF_N_SN8) {
        const uint32_t w = *(const uint32_t *)p;
        const float x = sn8f(w, 0);
        return make_float4(x, sn8f(w, 8), sn8f(w, 16), x);
    } else if constexpr (F == TF_C_F32_F32) {
        return *(const float4 *)p;
    } else if constexpr (F == TF_C_F16_F32) {
        const float4 t = *(const float4 *)p;
        const __half h = __ushort_as_half((unsigned short)(__float_as_uint(t.w) & 0xffffu));
        return make_float4(t.x, t.y, t.z, __half2float(h));
    } else if constexpr (F == TF_C_U8_F32) {
        const float4 t = *(const float4 *)p;
        return make_float4(t.x, t.y, t.z, (float)(__float_as_uint(t.w) & 0xffu));
    } else if constexpr (F == TF_C_F32_SN8) {
        const uint2 t = *(const uint2 *)p;
        return make_float4(sn8f(t.y, 0), sn8f(t.y, 8), sn8f(t.y, 16), __uint_as_float(t.x));
    } else if constexpr (F == TF_C_U8_SN8) {
        const uint32_t w = *(const uint32_t *)p;
        return make_float4(sn8f(w, 0), sn8f(w, 8), sn8f(w, 16), (float)(w >> 24));
    } else if constexpr (F == TF_C_F16_F16 || F == TF_C_F16_SN8) {
        const uint2 t = *(const uint2 *)p;
        return make_float4(h16f(t.x >> 16), h16f(t.y), h16f(t.y >> 16), h16f(t.x));
    } else if constexpr (F == TF_N_F16) {
        const uint2 t = *(const uint2 *)p;
        const float x = h16f(t.x);
        return make_float4(x, h16f(t.x >> 16), h16f(t.y), x);
    } else {
        return make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Scalar-only decode (avoids interpolating unused channels).
template <int F>
__device__ __forceinline__ float fetch1(const char *p)
{
    if constexpr (F == TF_S_F32) return *(const float *)p;
    else if constexpr (F == TF_S_F16) return __half2float(*(const __half *)p);
    else if constexpr (F == TF_S_U8) return (float)*(const uint8_t *)p;
    else return fetch4<F>(p).w;
}

// Runtime-format decode with exact GL scaling (generic path).
__device__ __forceinline__ float4 fetch_any(int fmt, const char *p)
{
    float4 t;
    switch (fmt) {
    case TF_S_F32: t = fetch4<TF_S_F32>(p); break;
    case TF_S_F16: t = fetch4<TF_S_F16>(p); break;
    case TF_S_U8: t = fetch4<TF_S_U8>(p); break;
    case TF_N_F32: t = fetch4<TF_N_F32>(p); break;
    case TF_N_SN8: t = fetch4<TF_N_SN8>(p); break;
    case TF_C_F32_F32: t = fetch4<TF_C_F32_F32>(p); break;
    case TF_C_F16_F32: t = fetch4<TF_C_F16_F32>(p); break;
    case TF_C_U8_F32: t = fetch4<TF_C_U8_F32>(p); break;
    case TF_C_F32_SN8: t = fetch4<TF_C_F32_SN8>(p); break;
    case TF_C_F16_SN8: t = fetch4<TF_C_F16_SN8>(p); break;
    case TF_C_U8_SN8: t = fetch4<TF_C_U8_SN8>(p); break;
    case TF_C_F16_F16: t = fetch4<TF_C_F16_F16>(p); break;
    case TF_N_F16: t = fetch4<TF_N_F16>(p); break;
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
// manager.py:95-101, 123-129).  Per axis the GL rule is u = c*size - 1/2,
// taps floor(u), floor(u)+1 (clamped), weight frac(u).  With the extended
// coordinates of the brick layout the lower tap is X0 = floor(c*size + 1/2)
// in [0, size] and no clamp is needed.  (w0, w1, w2) is the position in
// world units from the box minimum along the axes sampled by texture x, y, z
// (world z, y, x: the shader's swizzle, volume.frag.glsl:90).
struct Footprint {
    const char *p;    // byte address of the lower-corner texel (global-load path)
    uint32_t off;     // its byte offset in the allocation (buffer-load path)
    float u, v, w;    // extended texel coordinates along texture x, y, z (>= 0 in
                      // the box; the weights are their fractional parts)
    uint32_t X, Y, Z; // the lower-corner texel (floor of u, v, w)
    uint32_t brick;   // brick holding the whole footprint
};

// A footprint in two steps: the coordinates and the brick (all that the
// transparent-brick test needs), then the texel address (what the loads
// need), so that samples in transparent bricks skip the second.
template <int LB, bool BUF>
__device__ __forceinline__ Footprint footprint_brick(const TexDesc &t, float w0, float w1, float w2)
{
    // u, v, w >= 0 for in-box positions (and the 0 the callers pass for
    // others), so truncation is floor
    Footprint f;
    f.u = fmaf(w0, t.k0, 0.5f);
    f.v = fmaf(w1, t.k1, 0.5f);
    f.w = fmaf(w2, t.k2, 0.5f);
    f.X = (uint32_t)f.u;
    f.Y = (uint32_t)f.v;
    f.Z = (uint32_t)f.w;
    if constexpr (BUF) {
        // The buffer path is only used with nbricks < 2^24 (and the
        // allocation below 4 GiB): every partial brick index and
        // brick_bytes (<= 17^3 * 16) fit 24 bits and the byte offset 32, so
        // single-pass 24-bit multiplies are exact.
        f.brick = __umul24(__umul24(f.Z >> LB, t.nby) + (f.Y >> LB), t.nbx) + (f.X >> LB);
    } else {
        f.brick = ((f.Z >> LB) * t.nby + (f.Y >> LB)) * t.nbx + (f.X >> LB);
    }
    f.p = nullptr;
    f.off = 0;
    return f;
}

template <int T, int LB, bool BUF>
__device__ __forceinline__ void footprint_addr(const TexDesc &t, Footprint &f)
{
    constexpr uint32_t M = (1u << LB) - 1u, E = (1u << LB) + 1u;
    const uint32_t local = ((f.Z & M) * E + (f.Y & M)) * E + (f.X & M);
    if constexpr (BUF)
        f.off = __umul24(f.brick, t.brick_bytes) + local * (uint32_t)T;   // < nbytes32 < 2^32
    else
        f.p = (const char *)t.base + (uint64_t)f.brick * t.brick_bytes + (uint64_t)(local * (uint32_t)T);
}


// Where the march reads texels from: a buffer resource (base and size in
// scalar registers, 32-bit per-lane offsets, reads past the allocation
// return 0) for allocations below 4 GiB, else a 64-bit base pointer.
template <bool BUF>
struct Src;
template <>
struct Src<true> {
    __amdgpu_buffer_rsrc_t r;
    __device__ explicit Src(const TexDesc &t)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(t.base), (short)0, (int)t.nbytes32,
                                              0x00020000)) {}
};
template <>
struct Src<false> {
    __device__ explicit Src(const TexDesc &) {}
};

typedef uint32_t hvr_u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t hvr_u32x4 __attribute__((ext_vector_type(4)));

// NB dwords at byte offset o from the footprint's lower corner (o a
// compile-time constant; the buffer path folds it into the instruction).
template <int NB, bool BUF>
__device__ __forceinline__ void load_words(const Src<BUF> &s, const Footprint &f, uint32_t o, uint32_t *w)
{
    if constexpr (BUF) {
        if constexpr (NB == 1) {
            w[0] = __builtin_amdgcn_raw_buffer_load_b32(s.r, f.off + o, 0, 0);
        } else if constexpr (NB == 2) {
            const hvr_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(s.r, f.off + o, 0, 0);
            w[0] = v.x; w[1] = v.y;
        } else {
#pragma unroll
            for (int q = 0; q < NB / 4; ++q) {
                const hvr_u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(s.r, f.off + o + 16u * q, 0, 0);
                w[4 * q] = v.x; w[4 * q + 1] = v.y; w[4 * q + 2] = v.z; w[4 * q + 3] = v.w;
            }
        }
    } else {
        // texels are 4-byte aligned: load with that alignment
        typedef hvr_u32x2 u2a __attribute__((aligned(4)));
        typedef hvr_u32x4 u4a __attribute__((aligned(4)));
        const char *p = f.p + o;
        if constexpr (NB == 1) {
            w[0] = *(const uint32_t *)p;
        } else if constexpr (NB == 2) {
            const hvr_u32x2 v = *(const u2a *)p;
            w[0] = v.x; w[1] = v.y;
        } else {
#pragma unroll
            for (int q = 0; q < NB / 4; ++q) {
                const hvr_u32x4 v = *(const u4a *)(p + 16 * q);
                w[4 * q] = v.x; w[4 * q + 1] = v.y; w[4 * q + 2] = v.z; w[4 * q + 3] = v.w;
            }
        }
    }
}

// Decode one texel held in dwords w[0 .. T/4-1] (formats of 4 bytes and
// more) to (nx, ny, nz, rho), raw SNORM8/UNORM8 integers as fetch4 gives.
template <int F>
__device__ __forceinline__ float4 decode_words(const uint32_t *w)
{
    if constexpr (F == TF_S_F32) {
        const float v = __uint_as_float(w[0]);
        return make_float4(v, 0.f, 0.f, v);
    } else if constexpr (F == TF_N_F32) {
        const float x = __uint_as_float(w[0]);
        return make_float4(x, __uint_as_float(w[1]), __uint_as_float(w[2]), x);
    } else if constexpr (F == TF_N_SN8) {
        const float x = sn8f(w[0], 0);
        return make_float4(x, sn8f(w[0], 8), sn8f(w[0], 16), x);
    } else if constexpr (F == TF_C_F32_F32) {
        return make_float4(__uint_as_float(w[0]), __uint_as_float(w[1]), __uint_as_float(w[2]),
                           __uint_as_float(w[3]));
    } else if constexpr (F == TF_C_F16_F32) {
        return make_float4(__uint_as_float(w[0]), __uint_as_float(w[1]), __uint_as_float(w[2]),
                           __half2float(__ushort_as_half((unsigned short)(w[3] & 0xffffu))));
    } else if constexpr (F == TF_C_U8_F32) {
        return make_float4(__uint_as_float(w[0]), __uint_as_float(w[1]), __uint_as_float(w[2]),
                           (float)(w[3] & 0xffu));
    } else if constexpr (F == TF_C_F32_SN8) {
        return make_float4(sn8f(w[1], 0), sn8f(w[1], 8), sn8f(w[1], 16), __uint_as_float(w[0]));
    } else if constexpr (F == TF_C_U8_SN8) {
        return make_float4(sn8f(w[0], 0), sn8f(w[0], 8), sn8f(w[0], 16), (float)(w[0] >> 24));
    } else if constexpr (F == TF_C_F16_F16 || F == TF_C_F16_SN8) {
        return make_float4(h16f(w[0] >> 16), h16f(w[1]), h16f(w[1] >> 16), h16f(w[0]));
    } else if constexpr (F == TF_N_F16) {
        const float x = h16f(w[0]);
        return make_float4(x, h16f(w[0] >> 16), h16f(w[1]), x);
    } else {
        return make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Scalar (raw, as decode_words' w) and normal of a combined-format texel.
template <int F>
__device__ __forceinline__ float decode_scalar(const uint32_t *w)
{
    if constexpr (F == TF_C_F32_F32) return __uint_as_float(w[3]);
    else if constexpr (F == TF_C_F16_F32) return __half2float(__ushort_as_half((unsigned short)(w[3] & 0xffffu)));
    else if constexpr (F == TF_C_U8_F32) return (float)(w[3] & 0xffu);
    else if constexpr (F == TF_C_F32_SN8) return __uint_as_float(w[0]);
    else if constexpr (F == TF_C_U8_SN8) return (float)(w[0] >> 24);
    else if constexpr (F == TF_C_F16_F16 || F == TF_C_F16_SN8) return h16f(w[0]);
    else return 0.f;
}
template <int F>
__device__ __forceinline__ float3 decode_normal(const uint32_t *w)
{
    if constexpr (F == TF_C_F32_F32 || F == TF_C_F16_F32 || F == TF_C_U8_F32)
        return make_float3(__uint_as_float(w[0]), __uint_as_float(w[1]), __uint_as_float(w[2]));
    else if constexpr (F == TF_C_F32_SN8)
        return make_float3(sn8f(w[1], 0), sn8f(w[1], 8), sn8f(w[1], 16));
    else if constexpr (F == TF_C_U8_SN8)
        return make_float3(sn8f(w[0], 0), sn8f(w[0], 8), sn8f(w[0], 16));
    else if constexpr (F == TF_C_F16_F16 || F == TF_C_F16_SN8)
        return make_float3(h16f(w[0] >> 16), h16f(w[1]), h16f(w[1] >> 16));
    else return make_float3(0.f, 0.f, 0.f);
}

// The raw dwords of a footprint's eight texels of a combined format (4
// bytes and more): pair k = taps 2k, 2k+1 (y = k & 1, z = k >> 1).
template <int F, int LB, bool BUF>
__device__ __forceinline__ void load_pairs(const Src<BUF> &s, const Footprint &f, uint32_t raw[4][texel_bytes(F) / 2])
{
    constexpr int T = texel_bytes(F);
    constexpr uint32_t E = (1u << LB) + 1u;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        load_words<T / 2, BUF>(s, f, ((uint32_t)(k & 1) * E + (uint32_t)(k >> 1) * E * E) * (uint32_t)T, raw[k]);
}

// Trilinear scalar of a combined-format footprint from its raw pairs.
template <int F>
__device__ __forceinline__ float interp_scalar(const uint32_t (*raw)[texel_bytes(F) / 2], const float w[8])
{
    constexpr int H = texel_bytes(F) / 4;
    // an fma from 0 (not a product) lets a half operand fold into v_fma_mix
    float r = fmaf(w[0], decode_scalar<F>(raw[0]), 0.f);
#pragma unroll
    for (int k = 1; k < 8; ++k) r = fmaf(w[k], decode_scalar<F>(raw[k >> 1] + (k & 1) * H), r);
    return r;
}
template <int F>
__device__ __forceinline__ float3 interp_normal(const uint32_t (*raw)[texel_bytes(F) / 2], const float w[8])
{
    constexpr int H = texel_bytes(F) / 4;
    const float3 t0 = decode_normal<F>(raw[0]);
    float n0 = fmaf(w[0], t0.x, 0.f), n1 = fmaf(w[0], t0.y, 0.f), n2 = fmaf(w[0], t0.z, 0.f);
#pragma unroll
    for (int k = 1; k < 8; ++k) {
        const float3 t = decode_normal<F>(raw[k >> 1] + (k & 1) * H);
        n0 = fmaf(w[k], t.x, n0);
        n1 = fmaf(w[k], t.y, n1);
        n2 = fmaf(w[k], t.z, n2);
    }
    return make_float3(n0, n1, n2);
}

// The eight taps of a footprint, x fastest: tap i at (i&1, (i>>1)&1, i>>2).
// Formats of 4 bytes and more load each x-adjacent pair with one load of
// 2T bytes (texels are dword aligned); 1- and 2-byte formats load per tap.
template <int F, int LB, bool BUF>
__device__ __forceinline__ void load_taps(const Src<BUF> &s, const Footprint &f, float4 t[8])
{
    constexpr int T = texel_bytes(F);
    constexpr uint32_t E = (1u << LB) + 1u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t o = ((uint32_t)(k & 1) * E + (uint32_t)(k >> 1) * E * E) * (uint32_t)T;
        if constexpr (T >= 4) {
            uint32_t w[T / 2];
            load_words<T / 2, BUF>(s, f, o, w);
            t[2 * k] = decode_words<F>(w);
            t[2 * k + 1] = decode_words<F>(w + T / 4);
        } else {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                float v;
                if constexpr (BUF) {
                    if constexpr (T == 2)
                        v = __half2float(__ushort_as_half(
                            __builtin_amdgcn_raw_buffer_load_b16(s.r, f.off + o + 2u * j, 0, 0)));
                    else
                        v = (float)__builtin_amdgcn_raw_buffer_load_b8(s.r, f.off + o + (uint32_t)j, 0, 0);
                } else {
                    v = fetch1<F>(f.p + o + T * j);
                }
                t[2 * k + j] = make_float4(v, 0.f, 0.f, v);
            }
        }
    }
}

// The eight trilinear weights, x fastest: w[i] for tap (i&1, (i>>1)&1, i>>2).
__device__ __forceinline__ void tri_weights(float a, float b, float g, float w[8])
{
    const float a1 = 1.f - a, b1 = 1.f - b, g1 = 1.f - g;
    const float w00 = b1 * g1, w10 = b * g1, w01 = b1 * g, w11 = b * g;
    w[0] = a1 * w00; w[1] = a * w00;
    w[2] = a1 * w10; w[3] = a * w10;
    w[4] = a1 * w01; w[5] = a * w01;
    w[6] = a1 * w11; w[7] = a * w11;
}

// Weights of a footprint: frac(u) = u - floor(u), exact for u >= 0 (the
// box); v_fract_f32 computes it in one instruction.
__device__ __forceinline__ void fp_weights(const Footprint &f, float w[8])
{
    tri_weights(__builtin_amdgcn_fractf(f.u), __builtin_amdgcn_fractf(f.v), __builtin_amdgcn_fractf(f.w), w);
}

template <int F, int LB, bool BUF>
__device__ __forceinline__ float sample1(const Src<BUF> &s, const Footprint &f)
{
    float w[8];
    float4 t[8];
    load_taps<F, LB, BUF>(s, f, t);
    fp_weights(f, w);
    float r = w[0] * t[0].w;
#pragma unroll
    for (int k = 1; k < 8; ++k) r = fmaf(w[k], t[k].w, r);
    return r;
}

// Generic path: runtime format and brick size.
__device__ __forceinline__ float4 sample_any(const TexDesc &t, float w0, float w1, float w2)
{
    const float u = fmaf(w0, t.k0, 0.5f), v = fmaf(w1, t.k1, 0.5f), w = fmaf(w2, t.k2, 0.5f);
    const float fu = floorf(u), fv = floorf(v), fw = floorf(w);
    const uint32_t X = (uint32_t)(int)fu, Y = (uint32_t)(int)fv, Z = (uint32_t)(int)fw;
    const int lb = t.lb;
    const uint32_t M = (1u << lb) - 1u, E = (1u << lb) + 1u, T = (uint32_t)texel_bytes(t.fmt);
    const uint32_t brick = ((Z >> lb) * t.nby + (Y >> lb)) * t.nbx + (X >> lb);
    const uint32_t local = ((Z & M) * E + (Y & M)) * E + (X & M);
    const char *p = (const char *)t.base + (uint64_t)brick * t.brick_bytes + (uint64_t)local * T;
    float wt[8];
    tri_weights(u - fu, v - fv, w - fw, wt);
    float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int k = 0; k < 8; ++k) {
        const uint32_t off = ((k & 1) + ((k >> 1) & 1) * E + (k >> 2) * E * E) * T;
        const float4 c = fetch_any(t.fmt, p + off);
        r.x = fmaf(wt[k], c.x, r.x);
        r.y = fmaf(wt[k], c.y, r.y);
        r.z = fmaf(wt[k], c.z, r.z);
        r.w = fmaf(wt[k], c.w, r.w);
    }
    return r;
}

// ---------------------------------------------------------------------------
// Per-frame constants (uniforms after host-side FP64 preparation).
struct FrameConst {
    float prow[8];    // rows 0-1 of inverse(P_gl)          (volume.frag.glsl:50)
    float rot[9];     // 3x3 block of inverse(V_gl)         (volume.frag.glsl:52)
    float o[3];       // camera_pos
    float bmin[3], bmax[3], ext[3];   // ext = bmax - bmin in float (:26)
    float step;       // step_size
    float alpha_k;    // log2(e) * step_size / reference_step_size  (:101)
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
    int count_shaded;   // counting instances: count only samples that load texels
    uint8_t *touched;   // counting instances: set to 1 for every brick a sample loads from (or null)
};

// One ray per pixel, produced by the setup kernel.
struct __align__(16) RayRec {
    float p0x, p0y, p0z;  // first sample position (entry point or camera)
    int nlim;             // iterations to run (0: the ray misses the box)
    float dsx, dsy, dsz;  // ray_dir * step_size
    int pad;
};

// Ray setup (volume.vert.glsl:8; volume.frag.glsl:47-68, 71-82), run inline
// by the march kernels (once per ray) and by the block-ordering pass.
// Pixel (px, py), py = 0 the bottom row.  Positions use separately rounded
// products and sums (no fused multiply-add), as the shader states them.
__device__ __forceinline__ RayRec ray_setup(const FrameConst &fc, int px, int py)
{
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
    return r;
}

// K1 -- the dispatch order of a cost-ordered frame: one block per 16x16 ray
// block (row-major, ntx per row) sets up its rays (ray_setup, nothing
// stored) and files the block for the march (ordered_block) under the
// largest iteration bound of its rays.
// `cnt` holds 8 x kOrderBuckets counters (zero on entry), `slots` cap
// entries per class and bucket; only the blocks of the first `head_rows`
// block rows (a multiple of 8) are filed.  Block 0 also clears `clear`, the
// other set of counters, which the previous ordered frame's march read and
// the next ordered frame fills (counters alternate between frames, so no
// separate clearing launch is needed).
static __global__ void __launch_bounds__(256)
k_order_blocks(const FrameConst fc, uint32_t *__restrict__ cnt, uint32_t *__restrict__ clear,
               uint32_t *__restrict__ slots, int ntx, int head_rows, int cap)
{
    __shared__ int s_key[4];
    const int t = blockIdx.x, tx = t % ntx, ty = t / ntx;
    if (t == 0 && threadIdx.x < 8 * kOrderBuckets) clear[threadIdx.x] = 0u;
    if (ty >= head_rows) return;   // the whole block
    const int px = tx * 16 + (threadIdx.x & 15), py = ty * 16 + (threadIdx.x >> 4);
    int key = 0;
    if (px < fc.W && py < fc.H) key = ray_setup(fc, px, py).nlim;
    for (int off = 32; off > 0; off >>= 1) key = max(key, __shfl_xor(key, off, 64));
    if ((threadIdx.x & 63) == 0) s_key[threadIdx.x >> 6] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int m = max(max(s_key[0], s_key[1]), max(s_key[2], s_key[3]));
        const int q = (ty & 7) * kOrderBuckets + order_bucket(m, fc.max_steps);
        const uint32_t s = atomicAdd(cnt + q, 1u);
        if (s < (uint32_t)cap) slots[(uint64_t)q * (uint64_t)cap + s] = (uint32_t)t;
    }
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
    NM_ALIAS = 1,  // the scalar volume read as (rho, 0, 0)
    NM_ZERO = 2,   // unbound: normal = 0 -> diffuse term 0
    NM_ANY = 3     // generic: scalar and normal from any volumes/formats
};

// LUT lookup at (rho, 0.5) with u = rho*N - 1/2, clamp to edge
// (manager.py:163-181; volume.frag.glsl:96).  `lut` holds, per entry i, the
// pair (entry i, entry i+1 - entry i) (difference 0 after the last), so that
// clamping u to [0, N-1] gives both edge cases and one fma per channel.  The
// clamp is one v_med3_f32 (which maps a NaN to min(0, N-1) = 0, as
// fminf(fmaxf(NaN, 0), N-1) does), so u >= 0: truncation is floor and
// v_fract_f32 gives u - floor(u) exactly.
__device__ __forceinline__ float4 lut_sample(const float4 *lut, float fn1, float fn, float rho)
{
    const float u = __builtin_amdgcn_fmed3f(fmaf(rho, fn, -0.5f), 0.f, fn1);
    const float a = __builtin_amdgcn_fractf(u);
    const int i = (int)u;
    const float4 v = lut[2 * i], d = lut[2 * i + 1];
    return make_float4(fmaf(a, d.x, v.x), fmaf(a, d.y, v.y), fmaf(a, d.z, v.z), fmaf(a, d.w, v.w));
}

static constexpr int kLutLdsMax = 2048;  // entries; 2 x 16 B each in LDS

// Clamp-to-edge LUT lookup from global memory (generic path).
__device__ __forceinline__ float4 lut_global(const float4 *__restrict__ lut, int n, float fn1, float fn, float rho)
{
    const float u = fminf(fmaxf(fmaf(rho, fn, -0.5f), 0.f), fn1);
    const float fu = floorf(u);
    const int i0 = (int)fu, i1 = min(i0 + 1, n - 1);
    const float a = u - fu;
    const float4 v0 = lut[i0], v1 = lut[i1];
    return make_float4(fmaf(a, v1.x - v0.x, v0.x), fmaf(a, v1.y - v0.y, v0.y), fmaf(a, v1.z - v0.z, v0.z),
                       fmaf(a, v1.w - v0.w, v0.w));
}

// One sample's contribution (volume.frag.glsl:96-115): Beer-Lambert opacity
// correction 1 - exp(-alpha*step/ref), ambient + Lambert with the zero-vector
// rule (a normal of zero or subnormal length has no direction and gives no
// diffuse light), front-to-back composite.
__device__ __forceinline__ float lambert(const FrameConst &fc, float n0, float n1, float n2)
{
    // v_rsq_f32 without the subnormal-input rescaling of rsqrtf: the result
    // is only used for nn >= FLT_MIN, where the two agree
    const float nn = fmaf(n0, n0, fmaf(n1, n1, n2 * n2));
    const float dt = fmaf(n0, fc.L[0], fmaf(n1, fc.L[1], n2 * fc.L[2])) * __builtin_amdgcn_rsqf(nn);
    return (fc.L_ok && nn >= 1.17549435e-38f) ? fmaxf(dt, 0.f) : 0.f;
}

// The composite with the light factor lam = ka + kd * diffuse given.
__device__ __forceinline__ void composite_lam(const FrameConst &fc, const float4 &q, float lam, float &C0,
                                              float &C1, float &C2, float &A)
{
    const float a = 1.f - __builtin_amdgcn_exp2f(-(q.w * fc.alpha_k));
    const float ta = (1.f - A) * a;
    const float tl = ta * lam;
    C0 = fmaf(tl, q.x, C0);
    C1 = fmaf(tl, q.y, C1);
    C2 = fmaf(tl, q.z, C2);
    A = A + ta;
}

__device__ __forceinline__ void composite_diff(const FrameConst &fc, const float4 &q, float diff, float &C0,
                                               float &C1, float &C2, float &A)
{
    // v_exp_f32 without the subnormal-result rescaling of exp2f: when
    // exp2(-x) is subnormal, 1 - exp2(-x) rounds to 1 either way
    const float a = 1.f - __builtin_amdgcn_exp2f(-(q.w * fc.alpha_k));
    const float lam = fmaf(fc.kd, diff, fc.ka);
    const float ta = (1.f - A) * a;
    const float tl = ta * lam;
    C0 = fmaf(tl, q.x, C0);
    C1 = fmaf(tl, q.y, C1);
    C2 = fmaf(tl, q.z, C2);
    A = A + ta;
}

template <int NM>
__device__ __forceinline__ void composite(const FrameConst &fc, const float4 &q, float n0, float n1, float n2,
                                          float &C0, float &C1, float &C2, float &A)
{
    composite_diff(fc, q, NM != NM_ZERO ? lambert(fc, n0, n1, n2) : 0.f, C0, C1, C2, A);
}

// K2 -- march and composite (volume.frag.glsl:84-121), one thread per ray.
// A 256-thread block covers 16x16 pixels; each wave64 an 8x8 tile.  Blocks
// are dispatched round-robin over the 8 XCDs (block b on XCD (b + c) mod 8,
// c varying from launch to launch), and each XCD runs its blocks in order.
// The block -> tile mapping deals 16-pixel block rows to the XCDs in turn
// (XCD of b & 7 takes rows b & 7, 8 + (b & 7), ...): the work per row varies
// strongly over the image (where the rays cross dense matter), so
// contiguous bands per XCD left some XCDs idle for half the frame, while
// interleaved rows give every XCD an even share, and its L2 still sees whole
// rows.  Rows past the last multiple of 8 keep the plain order.  With
// `order` (a cost-ordered frame, k_order_blocks), each XCD's class of
// rows is taken in decreasing order of its blocks' iteration bounds
// instead (ordered_block), so that the long blocks start in the first round
// of resident blocks rather than after it.  A wave runs
// while any of its rays is live; a ray stops at its iteration bound
// (max_steps at most), when its alpha reaches tau, or when it leaves the box.
//
// In-box test (:88-92): fl((p - bmin)/ext) lies in [0, 1] exactly when
// fl(p - bmin) lies in [0, ext] (for ext > 0 the quotient of a larger
// difference rounds above 1), so the test needs no division.  The in-box
// samples of a ray are one contiguous run (each coordinate moves
// monotonically), so the loop skips the leading outside samples (the entry
// point may round outside) and ends at the first outside sample after them.
//
// Transparent bricks (SKIP instances): `skip[brick]` is set when every LUT
// entry that a density of the brick can weight is transparent (alpha 0 and
// a finite colour).  Such a sample contributes exactly nothing
// (C += (1-A)*lambda*rgb*0, A += (1-A)*0; lambda is finite because the
// Lambert term is 0 or max(dot, 0) with fmaxf dropping a NaN), so it is
// counted and stepped over without loading texels.
__device__ __forceinline__ bool inbox(const FrameConst &fc, float wx, float wy, float wz)
{
    return (wz >= 0.f) & (wz <= fc.ext[2]) & (wy >= 0.f) & (wy <= fc.ext[1]) & (wx >= 0.f) & (wx <= fc.ext[0]);
}

// The march of one ray (volume.frag.glsl:84-121) from its setup record:
// accumulated colour (C0, C1, C2), alpha A, and the number of composited
// samples ns (COUNT).  s_lut: the LUT with its differences in LDS (lut_sample).
// GATE: interpolate and light the normal only for samples of non-zero
// alpha (fewer instructions where the march is instruction-bound; the
// result is identical either way).
template <int SF, int NM, int LB, bool COUNT, bool SKIP, bool BUF, bool GATE = false>
__device__ __forceinline__ void march_ray(const FrameConst &fc, const TexDesc &ts, const TexDesc &tn,
                                          const float4 *__restrict__ lut_g, int lut_n, const float4 *s_lut,
                                          const RayRec &r, const uint8_t *__restrict__ skip,
                                          float &C0, float &C1, float &C2, float &A, uint32_t &ns)
{
    float x = r.p0x, y = r.p0y, z = r.p0z;
    const float fn = (float)lut_n, fn1 = (float)(lut_n - 1);
    constexpr float u8k = fmt_u8_scalar(SF) ? 1.0f / 255.0f : 1.0f;
    constexpr int T = texel_bytes(SF);
    const int nlim = r.nlim;
    int i = 0;
    // leading samples outside the box
    for (; i < nlim; ++i) {
        if (inbox(fc, x - fc.bmin[0], y - fc.bmin[1], z - fc.bmin[2])) break;
        x = __fadd_rn(x, r.dsx);
        y = __fadd_rn(y, r.dsy);
        z = __fadd_rn(z, r.dsz);
    }
    if constexpr (NM == NM_ANY || SF == TF_NONE) {
        // the in-box run, generic texture access
        for (; i < nlim; ++i) {
            if (!(A < fc.tau)) break;
            const float wx = x - fc.bmin[0], wy = y - fc.bmin[1], wz = z - fc.bmin[2];
            if (!inbox(fc, wx, wy, wz)) break;
            if constexpr (COUNT) ++ns;
            // texture x <- world z (size s0), y <- world y (s1), z <- world x (s2)
            float rho = 0.f, n0 = 0.f, n1 = 0.f, n2 = 0.f;
            if constexpr (NM == NM_ANY) {
                if (ts.base) {
                    const float4 t = sample_any(ts, wz, wy, wx);
                    rho = ts.role == 0 ? t.w : t.x;
                }
                if (tn.base) {
                    const float4 t = sample_any(tn, wz, wy, wx);
                    if (tn.role == 1) { n0 = t.x; n1 = t.y; n2 = t.z; }
                    else n0 = t.w;
                }
            }
            const float4 q = NM == NM_ANY ? lut_global(lut_g, lut_n, fn1, fn, rho)
                                          : lut_sample(s_lut, fn1, fn, rho);
            composite<NM>(fc, q, n0, n1, n2, C0, C1, C2, A);
            x = __fadd_rn(x, r.dsx);
            y = __fadd_rn(y, r.dsy);
            z = __fadd_rn(z, r.dsz);
        }
    } else {
        const Src<BUF> src(ts);
        // one sample's contribution from its footprint (coordinates and
        // brick; the address is completed here)
        auto shade = [&](const Footprint &fb) {
            Footprint f = fb;
            footprint_addr<T, LB, BUF>(ts, f);
            if constexpr (COUNT) {
                if (fc.touched) fc.touched[f.brick] = 1;   // the bricks the frame reads
            }
            if constexpr (NM == NM_SAME) {
                constexpr int T = texel_bytes(SF);
                uint32_t raw[4][T / 2];
                load_pairs<SF, LB, BUF>(src, f, raw);
                float w[8];
                fp_weights(f, w);
                const float rho = interp_scalar<SF>(raw, w) * u8k;
                if constexpr (GATE) {
                    // Normals only for a non-zero alpha: with alpha 0 the
                    // sample adds (1-A)*0*lambda = 0 for any finite lambda
                    // (the Lambert term is 0 or fmaxf(dot, 0)), so skipping
                    // the normal changes nothing; the texel loads are shared.
                    const float4 q = lut_sample(s_lut, fn1, fn, rho);
                    float lam = fc.ka;   // fma(kd, 0, ka) = ka exactly (kd finite)
                    if (q.w != 0.f) {
                        const float3 n = interp_normal<SF>(raw, w);
                        lam = fmaf(fc.kd, lambert(fc, n.x, n.y, n.z), fc.ka);
                    }
                    composite_lam(fc, q, lam, C0, C1, C2, A);
                } else {
                    const float3 n = interp_normal<SF>(raw, w);
                    const float4 q = lut_sample(s_lut, fn1, fn, rho);
                    composite_diff(fc, q, lambert(fc, n.x, n.y, n.z), C0, C1, C2, A);
                }
            } else {
                const float rho = sample1<SF, LB, BUF>(src, f) * u8k;
                const float4 q = lut_sample(s_lut, fn1, fn, rho);
                composite<NM>(fc, q, NM == NM_ALIAS ? rho : 0.f, 0.f, 0.f, C0, C1, C2, A);
            }
        };
        // The in-box run tests one bound per axis.  Per axis, with s = -1
        // where the step is negative (else +1), the run walks q = s*p by
        // s*ds >= 0 and forms a = fl(q - s*bmin): round-to-nearest commutes
        // with negation, so q is exactly s times the shader's position and
        // a = s*fl(p - bmin).  fl is monotonic, so a never decreases; the
        // run starts at a sample that passed the two-sided test, so a's
        // lower bound stays met and only a <= (s > 0 ? ext : 0) is tested.
        // In the box, p - bmin >= 0, so the footprint takes |a| (a free
        // source modifier) as fl(p - bmin).
        const float sx = r.dsx < 0.f ? -1.f : 1.f, sy = r.dsy < 0.f ? -1.f : 1.f,
                    sz = r.dsz < 0.f ? -1.f : 1.f;
        x *= sx;
        y *= sy;
        z *= sz;
        const float dqx = sx * r.dsx, dqy = sy * r.dsy, dqz = sz * r.dsz;
        const float bqx = -(sx * fc.bmin[0]), bqy = -(sy * fc.bmin[1]), bqz = -(sz * fc.bmin[2]);
        const float lx = sx > 0.f ? fc.ext[0] : 0.f, ly = sy > 0.f ? fc.ext[1] : 0.f,
                    lz = sz > 0.f ? fc.ext[2] : 0.f;
        if constexpr (!SKIP) {
            // the in-box run
            for (; i < nlim; ++i) {
                if (!(A < fc.tau)) break;
                const float ax = __fadd_rn(x, bqx), ay = __fadd_rn(y, bqy), az = __fadd_rn(z, bqz);
                if (!((az <= lz) & (ay <= ly) & (ax <= lx))) break;
                if constexpr (COUNT) ++ns;
                shade(footprint_brick<LB, BUF>(ts, fabsf(az), fabsf(ay), fabsf(ax)));
                x = __fadd_rn(x, dqx);
                y = __fadd_rn(y, dqy);
                z = __fadd_rn(z, dqz);
            }
        } else {
            // The in-box run with transparent bricks stepped over: a sample
            // in a transparent brick is counted without loading texels.  The
            // flag is loaded again only when the sample's brick changes (a
            // brick spans several steps), which saves most flag loads.  Each
            // iteration first advances to the next sample (its position, box
            // test and brick: none depends on the composite) and issues that
            // sample's flag load, then shades the current sample: the flag
            // load is in flight together with the current sample's texel
            // loads, so a wave waits for memory once per iteration, not
            // twice (some lane of a wave changes brick on most iterations).
            // A footprint outside the box is never used.
            float ax = __fadd_rn(x, bqx), ay = __fadd_rn(y, bqy), az = __fadd_rn(z, bqz);
            bool in = i < nlim && (az <= lz) & (ay <= ly) & (ax <= lx);
            // Unrolled by two, two footprints trading roles (no register
            // copies between iterations).  The flag held is always that
            // of the current sample's brick.
            Footprint fa = footprint_brick<LB, BUF>(ts, in ? fabsf(az) : 0.f, in ? fabsf(ay) : 0.f,
                                                    in ? fabsf(ax) : 0.f),
                      fb;
            uint32_t flag = ld_global(skip + fa.brick);
            bool clear = in && flag != 0;
            // one sample: f the current sample's footprint, g receives the next one's
            auto step = [&](const Footprint &f, Footprint &g) {
                ++i;
                x = __fadd_rn(x, dqx);
                y = __fadd_rn(y, dqy);
                z = __fadd_rn(z, dqz);
                ax = __fadd_rn(x, bqx);
                ay = __fadd_rn(y, bqy);
                az = __fadd_rn(z, bqz);
                const bool in_next = i < nlim && (az <= lz) & (ay <= ly) & (ax <= lx);
                g = footprint_brick<LB, BUF>(ts, fabsf(az), fabsf(ay), fabsf(ax));
                if (in_next && g.brick != f.brick) flag = ld_global(skip + g.brick);
                if (!clear) shade(f);
                if constexpr (COUNT) ns += (fc.count_shaded && clear) ? 0u : 1u;
                in = in_next;
                clear = in && flag != 0;
            };
            while (in && A < fc.tau) {
                step(fa, fb);
                if (!(in && A < fc.tau)) break;
                step(fb, fa);
            }
        }
    }
}

// Load the LUT of n entries, each followed by its difference to the next
// entry (0 after the last), into LDS.
__device__ __forceinline__ void load_lut_lds(const float4 *__restrict__ lut_g, int n, float4 *s_lut)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const float4 v = ld_global(lut_g + i);
        float4 d = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i + 1 < n) {
            const float4 nx = ld_global(lut_g + i + 1);
            d = make_float4(nx.x - v.x, nx.y - v.y, nx.z - v.z, nx.w - v.w);
        }
        s_lut[2 * i] = v;
        s_lut[2 * i + 1] = d;
    }
}

template <int SF, int NM, int LB, bool COUNT, bool SKIP, bool BUF>
__global__ void __launch_bounds__(256)
k_march(const FrameConst fc, const TexDesc ts, const TexDesc tn, const float4 *__restrict__ lut_g,
        int lut_n, uchar4 *__restrict__ out,
        uint32_t *__restrict__ counts, const uint8_t *__restrict__ skip, int ntx, int nblocks,
        const uint32_t *__restrict__ order_cnt, const uint32_t *__restrict__ order_slots, int order_cap)
{
    extern __shared__ float4 s_lut[];
    if constexpr (NM != NM_ANY) {
        load_lut_lds(lut_g, lut_n, s_lut);
        __syncthreads();
    }
    const int b = blockIdx.x;
    const int head = ((nblocks / ntx) >> 3 << 3) * ntx;   // blocks of the first multiple-of-8 rows
    int logical = b;
    if (order_cnt) {
        // cost-ordered (k_order_blocks filed the blocks)
        logical = ordered_block(b, head, order_cnt, order_slots, order_cap);
        if (logical < 0 || logical >= nblocks) logical = b;
    } else if (b < head) {
        const int j = b >> 3;   // j-th block of its XCD
        logical = ((j / ntx) * 8 + (b & 7)) * ntx + j % ntx;
    }
    const int tx = logical % ntx, ty = logical / ntx;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int px = tx * 16 + (wv & 1) * 8 + (lane & 7);
    const int py = ty * 16 + (wv >> 1) * 8 + (lane >> 3);
    if (px >= fc.W || py >= fc.H) return;
    const int pix = py * fc.W + px;
    const RayRec r = ray_setup(fc, px, py);

    float C0 = 0.f, C1 = 0.f, C2 = 0.f, A = 0.f;
    uint32_t ns = 0;
    march_ray<SF, NM, LB, COUNT, SKIP, BUF>(fc, ts, tn, lut_g, lut_n, s_lut, r, skip, C0, C1, C2, A, ns);
    out[pix] = blend_store(fc, C0, C1, C2, A);
    if constexpr (COUNT) counts[pix] = ns;
}

// K4 -- per-brick density range of a volume's scalar field: one wave per
// brick reduces the decoded scalar of its E^3 texels (apron included, so
// the range covers every footprint the brick serves).  A NaN widens the
// range to everything.  UNORM8 ranges are in decoded units (raw / 255).
template <int F, int LB>
__global__ void __launch_bounds__(256)
k_brick_range(const char *__restrict__ texels, uint32_t brick_bytes, uint32_t nbricks,
              float2 *__restrict__ range)
{
    const uint32_t brick = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (brick >= nbricks) return;
    constexpr int E = (1 << LB) + 1, N = E * E * E, T = texel_bytes(F);
    const char *base = texels + (uint64_t)brick * brick_bytes;
    float lo = INFINITY, hi = -INFINITY;
    bool nan = false;
    for (int k = lane; k < N; k += 64) {
        float v = fetch1<F>(base + (uint64_t)k * T);
        if constexpr (fmt_u8_scalar(F)) v *= 1.0f / 255.0f;
        nan |= v != v;
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, off, 64));
        hi = fmaxf(hi, __shfl_xor(hi, off, 64));
        nan |= __shfl_xor((int)nan, off, 64) != 0;
    }
    if (lane == 0) range[brick] = nan ? make_float2(-INFINITY, INFINITY) : make_float2(lo, hi);
}

// K5 -- transparent-brick flags for one LUT.  A sample of density rho reads
// u = clamp(rho*N - 1/2, 0, N-1) and weights entry floor(u) by 1 - frac(u)
// and entry floor(u) + 1 by frac(u), so the entries with non-zero weight
// over densities in [lo, hi] are floor(u(lo)) .. ceil(u(hi)); the range is
// widened by a margin for the rounding of the interpolated density.  `nz`
// is the prefix count (N + 1 values) of entries that are not exactly
// transparent (alpha != 0, or a non-finite colour).  `n_clear` counts the
// transparent bricks (one atomic per wave).
__device__ __forceinline__ bool brick_clear(float2 r, const int *__restrict__ nz, int n)
{
    return brick_clear_range(r.x, r.y, nz, n);
}

static __global__ void __launch_bounds__(256)
k_brick_flags(const float2 *__restrict__ range, uint32_t nbricks, const int *__restrict__ nz, int n,
              uint8_t *__restrict__ flags, uint32_t *__restrict__ n_clear)
{
    const uint32_t brick = blockIdx.x * 256 + threadIdx.x;
    bool clear = false;
    if (brick < nbricks) {
        clear = brick_clear(range[brick], nz, n);
        flags[brick] = clear ? 1 : 0;
    }
    const uint64_t m = __ballot(clear);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(n_clear, (uint32_t)__popcll(m));
}

// K3 -- upload: convert a slab of host float32 data (numpy C order = GL
// texel order, x fastest) into the volume's texel format and scatter each
// texel to every slot of the brick layout that holds it (its own slot, the
// apron copy in the previous brick, and the edge replicas at X = 0 and
// X = s+1).  part: 1 = scalar field, 2 = normal field (3 floats per voxel).
// The other part of an interleaved texel is preserved.
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

template <int F>
__device__ __forceinline__ void store_part(char *p, int part, const float *val)
{
    if (part == 1) {
        const float v = val[0];
        if constexpr (F == TF_S_F32) *(float *)p = v;
        else if constexpr (F == TF_S_F16) *(__half *)p = __float2half_rn(v);
        else if constexpr (F == TF_S_U8) *(uint8_t *)p = (uint8_t)enc_u8(v);
        else if constexpr (F == TF_C_F32_F32) ((float *)p)[3] = v;
        else if constexpr (F == TF_C_F16_F32) ((uint32_t *)p)[3] = (uint32_t)__half_as_ushort(__float2half_rn(v));
        else if constexpr (F == TF_C_U8_F32) ((uint32_t *)p)[3] = enc_u8(v);
        else if constexpr (F == TF_C_F32_SN8) ((float *)p)[0] = v;
        else if constexpr (F == TF_C_U8_SN8) ((uint8_t *)p)[3] = (uint8_t)enc_u8(v);
        else if constexpr (F == TF_C_F16_F16 || F == TF_C_F16_SN8) *(__half *)p = __float2half_rn(v);
    } else {
        const float nx = val[0], ny = val[1], nz = val[2];
        const uint32_t packed = enc_sn8(nx) | (enc_sn8(ny) << 8) | (enc_sn8(nz) << 16);
        if constexpr (F == TF_N_F32 || F == TF_C_F32_F32 || F == TF_C_F16_F32 || F == TF_C_U8_F32) {
            float *q = (float *)p;
            q[0] = nx; q[1] = ny; q[2] = nz;
            if constexpr (F == TF_N_F32) q[3] = 0.f;
        } else if constexpr (F == TF_N_SN8) {
            *(uint32_t *)p = packed;
        } else if constexpr (F == TF_C_F32_SN8) {
            ((uint32_t *)p)[1] = packed;
        } else if constexpr (F == TF_C_F16_SN8) {
            // the SNORM8 integers as halves (exact), which the march feeds
            // to v_fma_mix without a conversion per channel and tap
            __half *q = (__half *)p;   // halves 1..3; half 0 is the scalar
            q[1] = __int2half_rn((int)(int8_t)(packed & 0xffu));
            q[2] = __int2half_rn((int)(int8_t)((packed >> 8) & 0xffu));
            q[3] = __int2half_rn((int)(int8_t)((packed >> 16) & 0xffu));
        } else if constexpr (F == TF_C_U8_SN8) {
            uint8_t *q = (uint8_t *)p;
            q[0] = (uint8_t)(packed & 0xffu); q[1] = (uint8_t)((packed >> 8) & 0xffu);
            q[2] = (uint8_t)((packed >> 16) & 0xffu);
        } else if constexpr (F == TF_C_F16_F16) {
            __half *q = (__half *)p;   // halves 1..3; half 0 is the scalar
            q[1] = __float2half_rn(nx); q[2] = __float2half_rn(ny); q[3] = __float2half_rn(nz);
        } else if constexpr (F == TF_N_F16) {
            __half *q = (__half *)p;
            q[0] = __float2half_rn(nx); q[1] = __float2half_rn(ny); q[2] = __float2half_rn(nz);
            q[3] = __float2half_rn(0.f);
        }
    }
}

// Write one voxel's part (val: 1 or 3 floats) to every slot holding its
// texel; (x, y, z) are GL texel coordinates.
template <int F>
__device__ __forceinline__ void scatter_texel(char *__restrict__ dst, int x, int y, int z, int s0, int s1, int s2,
                                              int lb, uint32_t nbx, uint32_t nby, uint32_t nbz,
                                              uint32_t brick_bytes, int part, const float *val)
{
    uint32_t bx[4], lx[4], by[4], ly[4], bz[4], lz[4];
    const int nx = axis_slots(x, s0, lb, nbx, bx, lx);
    const int ny = axis_slots(y, s1, lb, nby, by, ly);
    const int nz = axis_slots(z, s2, lb, nbz, bz, lz);
    const uint32_t E = (1u << lb) + 1u;
    constexpr uint32_t T = (uint32_t)texel_bytes(F);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const uint32_t brick = (bz[k] * nby + by[j]) * nbx + bx[i];
                const uint32_t local = (lz[k] * E + ly[j]) * E + lx[i];
                store_part<F>(dst + (uint64_t)brick * brick_bytes + (uint64_t)local * T, part, val);
            }
}

template <int F>
__global__ void __launch_bounds__(256)
k_upload(char *__restrict__ dst, int s0, int s1, int s2, int lb, uint32_t nbx, uint32_t nby, uint32_t nbz,
         uint32_t brick_bytes, const float *__restrict__ src, int part, int z0, uint64_t nvox)
{
    const uint64_t l = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (l >= nvox) return;
    const uint64_t plane = (uint64_t)s0 * (uint64_t)s1;
    const int z = (int)(l / plane) + z0;
    const uint64_t rem = l % plane;
    const int y = (int)(rem / (uint64_t)s0), x = (int)(rem % (uint64_t)s0);
    scatter_texel<F>(dst, x, y, z, s0, s1, s2, lb, nbx, nby, nbz, brick_bytes, part,
                     src + (part == 1 ? l : 3 * l));
}

// Normals of a device float32 field (numpy C order, shape (s0, s1, s2)),
// exactly as compute_normal_volume (pyvr/datasets/synthetic.py:109-122)
// computes them from a float32 array: np.gradient along each numpy axis
// (central differences halved inside, one-sided at the two faces, unit
// spacing), the norm sqrt((gx^2 + gy^2) + gz^2) summed left to right as
// numpy does, plus float32(1e-8), and each component divided by it -- all
// correctly rounded float32 operations.  Every axis has at least 2 voxels.
__device__ __forceinline__ float np_gradient(const float *__restrict__ f, uint64_t l, int i, int n, uint64_t st)
{
    if (i == 0) return __fsub_rn(f[l + st], f[l]);
    if (i == n - 1) return __fsub_rn(f[l], f[l - st]);
    return __fmul_rn(__fsub_rn(f[l + st], f[l - st]), 0.5f);  // /2, exact
}

template <int F>
__global__ void __launch_bounds__(256)
k_normals_from_field(char *__restrict__ dst, int s0, int s1, int s2, int lb, uint32_t nbx, uint32_t nby,
                     uint32_t nbz, uint32_t brick_bytes, const float *__restrict__ field, uint64_t nvox)
{
    const uint64_t l = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (l >= nvox) return;
    // numpy coordinates of flat index l
    const uint64_t n12 = (uint64_t)s1 * (uint64_t)s2;
    const int i0 = (int)(l / n12);
    const uint64_t r12 = l % n12;
    const int i1 = (int)(r12 / (uint64_t)s2), i2 = (int)(r12 % (uint64_t)s2);
    const float gx = np_gradient(field, l, i0, s0, n12);
    const float gy = np_gradient(field, l, i1, s1, (uint64_t)s2);
    const float gz = np_gradient(field, l, i2, s2, 1);
    const float nn = __fadd_rn(__fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(gx, gx), __fmul_rn(gy, gy)), __fmul_rn(gz, gz))),
                               1e-8f);
    const float val[3] = {__fdiv_rn(gx, nn), __fdiv_rn(gy, nn), __fdiv_rn(gz, nn)};
    // GL texel coordinates of the same flat index
    const uint64_t plane = (uint64_t)s0 * (uint64_t)s1;
    const int z = (int)(l / plane);
    const uint64_t rem = l % plane;
    const int y = (int)(rem / (uint64_t)s0), x = (int)(rem % (uint64_t)s0);
    scatter_texel<F>(dst, x, y, z, s0, s1, s2, lb, nbx, nby, nbz, brick_bytes, 2, val);
}

// Seeded Gaussian-blob field generated on the device (numpy C order): the
// family of pyvr_amd.datasets.blob_field -- axes linspace(-1, 1, s),
// amplitude * exp(-(axis - centre)^2 / (2 sigma^2)) per axis multiplied as
// amp * (fx * (fy * fz)) in double, summed over the blobs -- plus uniform
// noise in [-noise, noise] from a counter-based hash of (seed, voxel),
// clipped to [0, 1] and rounded to float32.
static constexpr int kMaxBlobs = 32;
struct BlobSpec {
    double c[3 * kMaxBlobs], s[kMaxBlobs], a[kMaxBlobs];
    double noise;
    uint64_t seed;
    int n;
};

__device__ __forceinline__ uint64_t splitmix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double np_linspace(int i, int n)
{
    if (n <= 1) return -1.0;
    if (i == n - 1) return 1.0;
    return (double)i * (2.0 / (double)(n - 1)) + -1.0;
}

static __global__ void __launch_bounds__(256)
k_blob_field(float *__restrict__ out, int s0, int s1, int s2, const BlobSpec *__restrict__ spec, uint64_t nvox)
{
    const uint64_t l = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (l >= nvox) return;
    const uint64_t n12 = (uint64_t)s1 * (uint64_t)s2;
    const int i0 = (int)(l / n12);
    const uint64_t r12 = l % n12;
    const int i1 = (int)(r12 / (uint64_t)s2), i2 = (int)(r12 % (uint64_t)s2);
    const double x = np_linspace(i0, s0), y = np_linspace(i1, s1), z = np_linspace(i2, s2);
    double acc = 0.0;
    for (int b = 0; b < spec->n; ++b) {
        const double k = 2.0 * spec->s[b] * spec->s[b];
        const double dx = x - spec->c[3 * b], dy = y - spec->c[3 * b + 1], dz = z - spec->c[3 * b + 2];
        const double fx = exp(-(dx * dx) / k), fy = exp(-(dy * dy) / k), fz = exp(-(dz * dz) / k);
        acc += spec->a[b] * (fx * (fy * fz));
    }
    if (spec->noise > 0.0) {
        const uint64_t h = splitmix64(spec->seed ^ splitmix64(l));
        const double u = (double)(h >> 11) * 0x1p-53;  // [0, 1)
        acc += (2.0 * u - 1.0) * spec->noise;
    }
    out[l] = (float)fmin(fmax(acc, 0.0), 1.0);
}

// ---------------------------------------------------------------------------
// Batched rendering: many requests (own volume, image size, camera, LUT) in
// one persistent launch per kernel instance.  The host groups requests by
// instance (storage format, normal mode, brick size) into classes and
// stages, for each class, a descriptor array ordered by decreasing cost;
// tiles of 256 pixels are numbered through the class's requests in that
// order.  Blocks take the next tile from an atomic counter until none is
// left, load the request's LUT into LDS when the request changes, and set
// up and march each pixel's ray exactly as the single-request kernels do.

struct ReqDesc {
    FrameConst fc;
    TexDesc ts;                    // scalar volume (also the normals: NM_SAME / NM_ALIAS)
    const float4 *lut;             // lut_n RGBA entries
    const int *nz;                 // lut_n + 1 prefix counts of non-transparent entries
    const float2 *range;           // per-brick density range of the volume
    uint8_t *flags;                // per-brick transparent flags for this LUT
    uchar4 *out;                   // W*H RGBA8, bottom row first
    uint32_t *counts;              // W*H composited-sample counts (counting runs)
    uint32_t nbricks;
    int lut_n;
    int ntx;                       // tiles along x
    int tile_end;                  // one past this request's last tile in its class
};

// Per class: where its descriptors are in the staged block, and how many.
struct BatchClass {
    uint32_t desc_offset;          // bytes from the block start
    int nreq;
};

static constexpr int kBatchMaxClasses = 8;

// The staged block starts with this header, then the descriptors, LUTs and
// prefix counts.
struct BatchHeader {
    BatchClass cls[kBatchMaxClasses];
    int nclass;
    int nreq;                      // all requests (descriptors are contiguous)
    uint32_t all_offset;           // first descriptor of the first class
    int pad;
};

// Ray-tile shapes: a wave covers WX x WY pixels; a block is W waves (2, 4
// or 8; 4 by default), 2 across and W / 2 down, and takes one tile of
// (2 WX) x (W / 2 * WY) pixels at a time.  Fewer waves per block halve the
// tile and double the counter traffic; more lengthen the wait at the
// per-tile barrier.  The autotuner times them.
static constexpr int kBatchWaves = 4;

template <int TS>
struct TileShape;
template <>
struct TileShape<0> {
    static constexpr int WX = 8, WY = 8;
};
template <>
struct TileShape<1> {
    static constexpr int WX = 16, WY = 4;
};

// W: waves per block.  SKIP: step over transparent bricks (the flags of
// k_batch_flags).  GATE: interpolate and light normals only for samples of
// non-zero alpha.  Every variant gives the same image.  For texels of up
// to 8 bytes the launch bounds ask for 8 waves per SIMD (64 VGPRs): the
// compiler then keeps a few per-ray values in scratch outside the march
// loop rather than run 7 waves.  16-byte texels would spill inside the
// loop, so their kernels keep the registers they need.
template <int SF, int NM, int LB, bool COUNT, int TS, bool BUF, int W = kBatchWaves, bool SKIP = true,
          bool GATE = true>
__global__ void __launch_bounds__(64 * W, texel_bytes(SF) > 8 ? 1 : 8)
k_batch(const char *__restrict__ block, int cls, uint32_t *__restrict__ counter)
{
    extern __shared__ float4 s_lut[];
    __shared__ int s_item[2];
    constexpr int WX = TileShape<TS>::WX, WY = TileShape<TS>::WY, BX = 2 * WX, BY = W / 2 * WY;
    const BatchHeader *hdr = (const BatchHeader *)block;
    const int nreq = hdr->cls[cls].nreq;
    if (nreq <= 0) return;
    const ReqDesc *__restrict__ reqs = (const ReqDesc *)(block + hdr->cls[cls].desc_offset);
    const int total = reqs[nreq - 1].tile_end;
    // the wave's index in the block as a scalar, so that no per-thread value
    // but the lane index is needed after a ray's march (the pixel is formed
    // again there from scalars and the lane index)
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x) >> 6, lane = threadIdx.x & 63;
    int cur = -1;
    for (;;) {
        if (threadIdx.x == 0) {
            const int t = (int)atomicAdd(counter, 1u);
            int q = 0;
            if (t < total) {
                int lo = 0, hi = nreq - 1;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (reqs[mid].tile_end > t) hi = mid;
                    else lo = mid + 1;
                }
                q = lo;
            }
            s_item[0] = t;
            s_item[1] = q;
        }
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(s_item[0]);
        const int q = __builtin_amdgcn_readfirstlane(s_item[1]);
        if (t >= total) break;
        const ReqDesc &R = reqs[q];
        if (q != cur) {
            load_lut_lds(R.lut, R.lut_n, s_lut);
            __syncthreads();
            cur = q;
        }
        const int lt = t - (q ? reqs[q - 1].tile_end : 0);
        const int tx = lt % R.ntx, ty = lt / R.ntx;
        const int px = tx * BX + (wv & 1) * WX + lane % WX;
        const int py = ty * BY + (wv >> 1) * WY + lane / WX;
        if (px < R.fc.W && py < R.fc.H) {
            const RayRec r = ray_setup(R.fc, px, py);
            float C0 = 0.f, C1 = 0.f, C2 = 0.f, A = 0.f;
            uint32_t ns = 0;
            march_ray<SF, NM, LB, COUNT, SKIP, BUF, GATE>(R.fc, R.ts, R.ts, R.lut, R.lut_n, s_lut, r,
                                                           R.flags, C0, C1, C2, A, ns);
            const int ln = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));   // = lane
            const int pix = (ty * BY + (wv >> 1) * WY + ln / WX) * R.fc.W + tx * BX + (wv & 1) * WX + ln % WX;
            st_global(R.out + pix, blend_store(R.fc, C0, C1, C2, A));
            if constexpr (COUNT) st_global(R.counts + pix, ns);
        }
        // every thread is done with this tile's LDS before the next fetch
        __syncthreads();
    }
}

// Transparent-brick flags of every request of a batch for its own LUT.
static __global__ void __launch_bounds__(256) k_batch_flags(const char *__restrict__ block)
{
    const BatchHeader *hdr = (const BatchHeader *)block;
    const ReqDesc *__restrict__ reqs = (const ReqDesc *)(block + hdr->all_offset);
    const uint32_t stride = gridDim.x * 256u;
    for (int q = 0; q < hdr->nreq; ++q) {
        const ReqDesc &R = reqs[q];
        for (uint32_t b = blockIdx.x * 256u + threadIdx.x; b < R.nbricks; b += stride)
            st_global(R.flags + b, (uint8_t)(brick_clear(ld_global(R.range + b), R.nz, R.lut_n) ? 1 : 0));
    }
}

static __global__ void __launch_bounds__(64) k_fill_u32(uint32_t *__restrict__ p, int n, uint32_t v)
{
    for (int i = threadIdx.x; i < n; i += 64) p[i] = v;
}

// A one-request batch block written on the device from kernel arguments (so
// that a frame needs no host-to-device copy and no staging allocation): the
// header at the start of the block (where k_batch and k_batch_flags read
// it), the tile counter reset to 0, the prefix counts of the unbound LUT
// (0, 1), and the descriptor, at the offsets the host chose.
struct StageOffsets {
    uint32_t counter, unbound_nz, desc;
};

static __global__ void __launch_bounds__(64)
k_stage_request(char *__restrict__ blk, const StageOffsets off, const BatchHeader hdr, const ReqDesc desc)
{
    const int lane = threadIdx.x;
    if (lane == 0) {
        *(uint32_t *)(blk + off.counter) = 0u;
        ((int *)(blk + off.unbound_nz))[0] = 0;
        ((int *)(blk + off.unbound_nz))[1] = 1;
    }
    static_assert(sizeof(BatchHeader) % 4 == 0 && sizeof(ReqDesc) % 4 == 0, "dword copies");
    const uint32_t *h = (const uint32_t *)&hdr, *d = (const uint32_t *)&desc;
    uint32_t *hh = (uint32_t *)blk, *dd = (uint32_t *)(blk + off.desc);
    for (int i = lane; i < (int)(sizeof(BatchHeader) / 4); i += 64) hh[i] = h[i];
    for (int i = lane; i < (int)(sizeof(ReqDesc) / 4); i += 64) dd[i] = d[i];
}

// Decoded texels of a volume in numpy C order (texel (x, y, z) at flat
// x + s0*(y + s1*z)), z-planes z0 .. z0 + nz - 1: the scalar as the march
// reads it (UNORM8 times 1/255) and the normal as stored (SNORM8 as its
// integer, FP16 as float).  Either output may be null.
template <int F>
__global__ void __launch_bounds__(256)
k_download(const char *__restrict__ texels, int s0, int s1, int lb, uint32_t nbx, uint32_t nby, uint32_t brick_bytes,
           int z0, uint64_t nvox, float *__restrict__ scalar, float *__restrict__ normal)
{
    const uint64_t l = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (l >= nvox) return;
    const uint64_t plane = (uint64_t)s0 * (uint64_t)s1;
    const uint32_t z = (uint32_t)(l / plane) + (uint32_t)z0;
    const uint64_t rem = l % plane;
    const uint32_t y = (uint32_t)(rem / (uint64_t)s0), x = (uint32_t)(rem % (uint64_t)s0);
    // the texel's own slot in the extended (apron) coordinates X = x + 1
    const uint32_t X = x + 1u, Y = y + 1u, Z = z + 1u, M = (1u << lb) - 1u, E = (1u << lb) + 1u;
    const uint64_t brick = ((uint64_t)(Z >> lb) * nby + (Y >> lb)) * nbx + (X >> lb);
    const uint32_t local = ((Z & M) * E + (Y & M)) * E + (X & M);
    const float4 t = fetch4<F>(texels + brick * brick_bytes + (uint64_t)local * (uint32_t)texel_bytes(F));
    if (scalar) scalar[l] = fmt_u8_scalar(F) ? t.w * (1.0f / 255.0f) : t.w;
    if (normal) {
        normal[3 * l] = t.x;
        normal[3 * l + 1] = t.y;
        normal[3 * l + 2] = t.z;
    }
}

}  // namespace hvr
