// hipvr.hip -- libhipvr.so: C ABI (include/hipvr.h) over the gfx950 kernels
// of hvr_device.h.  Host side: contexts (framebuffer, rays, LUT, stream,
// events), volumes (device texels in brick layout), FP64 uniform
// preparation and kernel dispatch.
#include "hvr_copy.h"
#include "hvr_launch.h"
#include "../../include/hipvr.h"

#include <math.h>
#include <cmath>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

using namespace hvr;

// ---------------------------------------------------------------------------
// errors
static thread_local std::string g_err;

// A failed HIP call (an allocation above the device's memory, say) also
// sets the thread's last error, which the launch checks of later calls read:
// once a failure is reported (or handled), the last error is cleared so that
// a later, unrelated call does not report it again.
static int fail(int code, const char *msg)
{
    g_err = msg;
    (void)hipGetLastError();
    return code;
}
static int fail_hip(hipError_t e, const char *where)
{
    g_err = std::string(where) + ": " + hipGetErrorString(e);
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? HVR_ENOMEM : HVR_EDEVICE;
}
#define HIPCHK(expr)                                   \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return fail_hip(e_, #expr); \
    } while (0)

// ---------------------------------------------------------------------------
// fault injection (hvr_debug_inject_fault): one armed entry point fails with
// a chosen code on its n-th call, before any device work
static std::mutex g_fault_mu;
static std::atomic<bool> g_fault_armed{false};
static std::string g_fault_fn;
static int g_fault_code = 0, g_fault_after = -1;

static int fault_point(const char *fn)
{
    if (!g_fault_armed.load(std::memory_order_relaxed)) return HVR_OK;
    std::lock_guard<std::mutex> lock(g_fault_mu);
    if (g_fault_after < 0 || g_fault_fn != fn) return HVR_OK;
    if (g_fault_after-- > 0) return HVR_OK;
    g_fault_armed.store(false);
    g_fault_after = -1;
    return fail(g_fault_code, (std::string("injected fault in ") + fn).c_str());
}
#define HVR_FAULT_POINT(fn)                             \
    do {                                               \
        if (const int rc_ = fault_point(fn)) return rc_; \
    } while (0)

// ---------------------------------------------------------------------------
// objects
struct hvr_volume {
    hvr_ctx *ctx;
    int s0, s1, s2;
    int fmt;               // TexFmt
    int lb;                // log2 of the brick interior edge
    uint32_t nbx, nby, nbz;
    uint32_t brick_bytes;  // (2^lb + 1)^3 texels
    uint64_t bytes;
    void *d_texels;
    bool scalar_loaded, normals_loaded;
    float2 *d_range;       // per-brick scalar range (after a scalar upload)
    uint32_t gen;          // bumped on every upload
};

struct hvr_ctx {
    int device;
    int W, H;
    hipStream_t stream;
    uchar4 *d_out;
    uint32_t *d_counts;
    float4 *d_lut;
    int lut_n;       // entries in d_lut (>= 1)
    int lut_cap;
    void *h_pinned;  // W*H*4 readback staging
    void *h_pinned_dev;   // its device address (the march's target in HVR_READBACK_DIRECT)
    int readback_mode;    // HVR_READBACK_COPY / _DIRECT
    bool last_host;       // the last frame was written to h_pinned
    float *d_stage;  // upload staging
    size_t stage_bytes;
    hipEvent_t ev[4];
    hipEvent_t rev[4];   // readback chunks (hvr_read)
    hipEvent_t last_ev[3];  // the last frame's events (ev[] or a timing region's)
    bool events_recorded;
    hvr_params params;
    bool params_set;
    hvr_volume *scalar_vol, *normal_vol;
    int scalar_role, normal_role;
    int brick_log2;
    // per-launch timing region (hvr_timing_begin/end)
    std::vector<hipEvent_t> *tev;
    int timing_max, timing_n;
    // transparent-brick skipping
    int *d_nz;             // prefix counts of non-transparent LUT entries (lut_n + 1)
    int nz_cap;            // entries d_nz holds counts for
    uint32_t lut_gen;
    uint8_t *d_flags;      // per-brick transparent flags of flags_vol
    uint32_t *d_nclear;    // transparent-brick count written by k_brick_flags
    uint32_t *h_nclear;    // its copy in pinned host memory, valid once flags_ev completed
    hipEvent_t flags_ev;   // recorded after that copy
    bool flags_count_known;   // flags_clear holds the count of the current flags
    uint64_t flags_cap;
    const hvr_volume *flags_vol;
    uint32_t flags_vol_gen, flags_lut_gen;
    uint64_t flags_clear, flags_total;
    int skip_mode;         // HVR_SKIP_OFF / _ON / _AUTO
    int skip_used;         // whether the last frame's march stepped over bricks
    bool force_global;     // HVR_GATHER=global at creation: 64-bit loads for every volume
    int ncu;               // compute units of the device
    size_t pix_cap;        // pixels d_out and h_pinned hold
    size_t counts_cap;     // pixels d_counts holds
    // cost-ordered dispatch of single frames (k_order_blocks)
    uint32_t *d_order;     // two sets of 8 x kOrderBuckets counters, then the slots
    size_t order_words;    // words d_order holds
    int order_mode;        // HVR_ORDER_OFF / _ON / _AUTO
    int order_used;        // whether the last frame was cost-ordered
    int order_parity;      // counter set of the next ordered frame
    // LUT uploads (hvr_set_lut): pinned staging slots holding the entries and
    // their prefix counts, each reused once the copy that last read it is done
    void *h_lut_stage[4];
    hipEvent_t lut_ev[4];
    bool lut_ev_live[4];
    size_t lut_stage_bytes;   // bytes each slot holds
    int lut_slot;             // the next slot
};

// HVR_SKIP_AUTO steps over transparent bricks when at least this fraction of
// the bound volume's bricks are transparent for the current LUT.
static constexpr double kSkipAutoFraction = 0.25;

// HVR_ORDER_AUTO orders a frame's blocks by cost when they are more than the
// march keeps resident at once (8 blocks of 256 threads per compute unit):
// only then do blocks start late.
static constexpr int kOrderAutoBlocksPerCU = 8;


// ---------------------------------------------------------------------------
// FP64 uniform preparation

// Uniforms -> kernel constants.  The matrices arrive as numpy row-major
// float32 bytes that GLSL reads column-major, i.e. transposed
// (manager.py:192); they are inverted in double and rounded once to float.
static FrameConst frame_const_of(const hvr_params &p, bool params_set, int W, int H)
{
    FrameConst f;
    memset(&f, 0, sizeof(f));
    f.W = W;
    f.H = H;
    f.fW = (float)W;
    f.fH = (float)H;
    double pg[16], vg[16], pinv[16], vinv[16];
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) {
            pg[4 * r + k] = (double)p.proj[4 * k + r];
            vg[4 * r + k] = (double)p.view[4 * k + r];
        }
    f.ok = params_set && p.camera_set && inverse4(pg, pinv) && inverse4(vg, vinv);
    if (f.ok) {
        for (int k = 0; k < 8; ++k) f.prow[k] = (float)pinv[k];
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) f.rot[3 * r + k] = (float)vinv[4 * r + k];
    }
    for (int k = 0; k < 3; ++k) {
        f.o[k] = p.cam_pos[k];
        f.bmin[k] = p.bmin[k];
        f.bmax[k] = p.bmax[k];
        f.ext[k] = p.bmax[k] - p.bmin[k];  // float subtraction, as in the shader
    }
    f.step = p.step;
    // exp(-a*step/ref) = exp2(-a * log2(e) * step/ref)
    f.alpha_k = (float)(1.4426950408889634 * (double)p.step / (double)p.ref_step);
    f.ka = p.ambient;
    f.kd = p.diffuse;
    f.tau = p.et_threshold;
    f.max_steps = p.max_steps;
    double l[3], ll = 0.0;
    for (int k = 0; k < 3; ++k) {
        l[k] = (double)p.light_target[k] - (double)p.light_pos[k];
        ll += l[k] * l[k];
    }
    f.L_ok = ll > 0.0;
    for (int k = 0; k < 3; ++k) f.L[k] = f.L_ok ? (float)(l[k] / sqrt(ll)) : 0.f;
    for (int k = 0; k < 4; ++k) f.clear[k] = p.clear_rgba[k];
    f.blend = p.blend != 0;
    return f;
}

static FrameConst frame_const(const hvr_ctx *c) { return frame_const_of(c->params, c->params_set, c->W, c->H); }

// Texture descriptor for a frame: texture x samples world z, y samples y,
// z samples x (the shader's swizzle), so k0 = s0/ext_z, k1 = s1/ext_y,
// k2 = s2/ext_x with ext the float box extent of the in-box test.
static TexDesc tex_desc(const hvr_volume *v, int role, const FrameConst &fc)
{
    TexDesc t;
    memset(&t, 0, sizeof(t));
    if (!v) return t;
    t.base = v->d_texels;
    t.s0 = v->s0;
    t.s1 = v->s1;
    t.s2 = v->s2;
    t.k0 = (float)((double)v->s0 / (double)fc.ext[2]);
    t.k1 = (float)((double)v->s1 / (double)fc.ext[1]);
    t.k2 = (float)((double)v->s2 / (double)fc.ext[0]);
    t.lb = v->lb;
    t.nbx = v->nbx;
    t.nby = v->nby;
    t.brick_bytes = v->brick_bytes;
    const uint64_t nbricks = (uint64_t)v->nbx * v->nby * v->nbz;
    t.nbytes32 = v->bytes <= 0xffffffffull && nbricks < (1ull << 24) ? (uint32_t)v->bytes : 0u;
    t.fmt = v->fmt;
    t.role = role;
    return t;
}

// ---------------------------------------------------------------------------
// kernel dispatch
// The generic march instances (runtime format; no scalar field bound).
template <int SF, int NM, bool COUNT>
static void launch_march_generic(const MarchArgs &a)
{
    const size_t lds = NM == NM_ANY ? 0 : 2 * (size_t)a.lut_n * sizeof(float4);
    hipLaunchKernelGGL((k_march<SF, NM, 3, COUNT, false, false>), dim3(a.nblocks), dim3(256), lds, a.stream, a.fc,
                       a.ts, a.tn, a.lut, a.lut_n, a.out, a.counts, a.skip, a.ntx, a.nblocks,
                       a.order_cnt, a.order_slots, a.order_cap);
}

// The transparent-brick count of the current flags, once its copy to the
// host has completed (waiting for it when `wait`).  Returns whether
// flags_clear holds it.
static bool flags_count(hvr_ctx *c, bool wait)
{
    if (c->flags_count_known) return true;
    const hipError_t q = wait ? hipEventSynchronize(c->flags_ev) : hipEventQuery(c->flags_ev);
    if (q == hipSuccess) {
        c->flags_clear = *c->h_nclear;
        c->flags_count_known = true;
    } else if (q == hipErrorNotReady) {
        (void)hipGetLastError();   // still in flight, not an error
    }
    return c->flags_count_known;
}

// Transparent-brick flags of volume v for the current LUT, recomputed when
// either changed; their count comes back to the host without a wait.
// Returns the flags when the march should step over transparent bricks:
// always in HVR_SKIP_ON; in HVR_SKIP_AUTO when enough bricks are
// transparent, or while the count is not back yet (skipping gives the same
// image, so the choice only affects speed); nullptr otherwise.
static const uint8_t *brick_flags(hvr_ctx *c, const hvr_volume *v)
{
    if (c->skip_mode == HVR_SKIP_OFF || !v || !v->d_range || !c->d_nz) return nullptr;
    const uint64_t nb = (uint64_t)v->nbx * v->nby * v->nbz;
    // without flags the frame takes the non-skipping kernel: a failure here
    // is handled, not reported
    const auto none = [] {
        (void)hipGetLastError();
        return (const uint8_t *)nullptr;
    };
    if (!(c->flags_vol == v && c->flags_vol_gen == v->gen && c->flags_lut_gen == c->lut_gen)) {
        c->flags_vol = nullptr;
        if (!c->d_nclear && hipMalloc(&c->d_nclear, sizeof(uint32_t)) != hipSuccess) return none();
        if (nb > c->flags_cap) {
            if (hipStreamSynchronize(c->stream) != hipSuccess) return none();
            (void)hipFree(c->d_flags);
            c->d_flags = nullptr;
            c->flags_cap = 0;
            if (hipMalloc(&c->d_flags, nb) != hipSuccess) return none();
            c->flags_cap = nb;
        }
        // the count comes back asynchronously (no wait in the frame's path)
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, c->d_nclear, 1, 0u);
        hipLaunchKernelGGL(k_brick_flags, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, c->stream,
                           (const float2 *)v->d_range, (uint32_t)nb, (const int *)c->d_nz, c->lut_n, c->d_flags,
                           c->d_nclear);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(c->h_nclear, c->d_nclear, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream) !=
                hipSuccess ||
            hipEventRecord(c->flags_ev, c->stream) != hipSuccess)
            return none();
        c->flags_vol = v;
        c->flags_vol_gen = v->gen;
        c->flags_lut_gen = c->lut_gen;
        c->flags_count_known = false;
        c->flags_total = nb;
    }
    if (c->skip_mode == HVR_SKIP_ON) return c->d_flags;
    if (!flags_count(c, false)) return c->d_flags;   // not back yet: skipping is exact either way
    return (double)c->flags_clear >= kSkipAutoFraction * (double)c->flags_total ? c->d_flags : nullptr;
}

// The march of one frame into `out` (W*H RGBA8 on the device): the fast
// instances of the scalar volume's format (normals from the same texels, as
// (rho, 0, 0) of the same volume, or unbound; brick edge 8 or 16), with
// transparent-brick skipping when the flags say so and buffer loads below
// 4 GiB, else the generic instance.
static void launch_march(hvr_ctx *c, const FrameConst &fc, uchar4 *out, bool count, const uint32_t *order_cnt,
                         const uint32_t *order_slots, int order_cap)
{
    MarchArgs a;
    a.fc = fc;
    a.order_cnt = order_cnt;
    a.order_slots = order_slots;
    a.order_cap = order_cap;
    a.ntx = (c->W + 15) / 16;
    a.nblocks = a.ntx * ((c->H + 15) / 16);
    const hvr_volume *S = c->scalar_vol, *N = c->normal_vol;
    a.ts = tex_desc(S, c->scalar_role, fc);
    a.tn = tex_desc(N, c->normal_role, fc);
    a.lut = (const float4 *)c->d_lut;
    a.lut_n = c->lut_n;
    a.out = out;
    a.counts = c->d_counts;
    a.skip = nullptr;
    a.stream = c->stream;
    const bool small_lut = c->lut_n <= kLutLdsMax;
    c->skip_used = 0;
    if (small_lut && !S && !N) {
        a.tn = a.ts = tex_desc(nullptr, 0, fc);
        if (count) launch_march_generic<TF_NONE, NM_ZERO, true>(a);
        else launch_march_generic<TF_NONE, NM_ZERO, false>(a);
        return;
    }
    if (small_lut && S && c->scalar_role == HVR_ROLE_SCALAR && (S->lb == 3 || S->lb == 4)) {
        int nm = -1;
        if (!N) nm = NM_ZERO;
        else if (N == S && c->normal_role == HVR_ROLE_SCALAR) nm = NM_ALIAS;
        else if (N == S && c->normal_role == HVR_ROLE_NORMAL && fmt_has_scalar(S->fmt)) nm = NM_SAME;
        if (nm >= 0) {
            const uint8_t *skip = brick_flags(c, S);
            const bool buf = a.ts.nbytes32 != 0 && !c->force_global;
            a.skip = skip;
            a.tn = tex_desc(nullptr, 0, fc);
            bool done = false;
            switch (S->fmt) {
#define HVR_CASE(F)                                                            \
    case F:                                                                    \
        done = launch_march_##F(a, nm, S->lb, count, skip != nullptr, buf);    \
        break;
                HVR_FAST_FORMATS(HVR_CASE)
#undef HVR_CASE
            default: break;
            }
            if (done) {
                c->skip_used = skip != nullptr;
                return;
            }
            a.skip = nullptr;
            a.tn = tex_desc(N, c->normal_role, fc);
        }
    }
    if (count) launch_march_generic<TF_NONE, NM_ANY, true>(a);
    else launch_march_generic<TF_NONE, NM_ANY, false>(a);
}

static int enqueue_frame(hvr_ctx *c, bool count, uchar4 *dst, bool count_shaded = false,
                         uint8_t *touched = nullptr)
{
    FrameConst fc = frame_const(c);
    fc.count_shaded = count_shaded;
    fc.touched = touched;
    // events of this frame: the context's own, or the next slots of an
    // open timing region; hvr_last_kernel_ms reads the frame's events
    hipEvent_t e0 = c->ev[0], e1 = c->ev[1], e2 = c->ev[2];
    if (c->tev && c->timing_n < c->timing_max) {
        const size_t k = 3 * (size_t)c->timing_n++;
        e0 = (*c->tev)[k];
        e1 = (*c->tev)[k + 1];
        e2 = (*c->tev)[k + 2];
    }
    // cost-ordered dispatch (hvr_set_dispatch_order): the blocks of the
    // first multiple-of-8 block rows are filed by k_order_blocks and marched
    // in decreasing order of their rays' iteration bounds
    const int ntx = (c->W + 15) / 16, nty = (c->H + 15) / 16, nblocks = ntx * nty;
    const int head_rows = nty >> 3 << 3, cap = (head_rows >> 3) * ntx;
    const bool ordered = head_rows > 0 && (c->order_mode == HVR_ORDER_ON ||
                                           (c->order_mode == HVR_ORDER_AUTO &&
                                            nblocks > kOrderAutoBlocksPerCU * c->ncu));
    constexpr int kSet = 8 * kOrderBuckets;   // counters per set
    if (ordered) {
        const size_t words = 2 * (size_t)kSet + (size_t)kSet * (size_t)cap;
        if (words > c->order_words) {
            HIPCHK(hipStreamSynchronize(c->stream));
            (void)hipFree(c->d_order);
            c->d_order = nullptr;
            c->order_words = 0;
            HIPCHK(hipMalloc(&c->d_order, words * sizeof(uint32_t)));
            c->order_words = words;
            // both counter sets start at zero (a kernel: no memset node in
            // a graph a caller captures)
            hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, c->d_order, 2 * kSet, 0u);
            HIPCHK(hipGetLastError());
            c->order_parity = 0;
        }
    }
    c->order_used = ordered;
    uint32_t *cnt = nullptr, *slots = nullptr;
    HIPCHK(hipEventRecord(e0, c->stream));
    if (ordered) {
        // this frame files into one counter set and clears the other
        const int p = c->order_parity;
        c->order_parity = p ^ 1;
        cnt = c->d_order + p * kSet;
        slots = c->d_order + 2 * kSet;
        // the blocks of the head rows only (the others keep row order)
        hipLaunchKernelGGL(k_order_blocks, dim3(head_rows * ntx), dim3(256), 0, c->stream, fc, cnt,
                           c->d_order + (p ^ 1) * kSet, slots, ntx, head_rows, cap);
        HIPCHK(hipGetLastError());
    }
    // the march sets up its rays itself
    HIPCHK(hipEventRecord(e1, c->stream));
    const bool host = !dst && c->readback_mode == HVR_READBACK_DIRECT;
    launch_march(c, fc, dst ? dst : host ? (uchar4 *)c->h_pinned_dev : c->d_out, count, cnt, slots, cap);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e2, c->stream));
    c->last_ev[0] = e0;
    c->last_ev[1] = e1;
    c->last_ev[2] = e2;
    c->events_recorded = true;
    if (!dst) c->last_host = host;
    return HVR_OK;
}

// Per-device address of the no-allocation entry point's scratch array,
// looked up once per device at context creation (defined with the batches).
static void noalloc_setup(int device);

// ---------------------------------------------------------------------------
// C ABI
template <int F, int LB>
static void launch_range(hvr_volume *v, uint32_t nb)
{
    hipLaunchKernelGGL((k_brick_range<F, LB>), dim3((nb + 3) / 4), dim3(256), 0, v->ctx->stream,
                       (const char *)v->d_texels, v->brick_bytes, nb, v->d_range);
}

// Per-brick scalar range of a volume (for transparent-brick skipping).
static int compute_range(hvr_volume *v)
{
    hvr_ctx *c = v->ctx;
    const uint64_t nb = (uint64_t)v->nbx * v->nby * v->nbz;
    if (!v->d_range) HIPCHK(hipMalloc(&v->d_range, nb * sizeof(float2)));
#define HVR_RG(F)                                                 \
    case F:                                                       \
        if (v->lb == 2) launch_range<F, 2>(v, (uint32_t)nb);      \
        else if (v->lb == 3) launch_range<F, 3>(v, (uint32_t)nb); \
        else launch_range<F, 4>(v, (uint32_t)nb);                 \
        break;
    switch (v->fmt) {
        HVR_RG(TF_S_F32)
        HVR_RG(TF_S_F16)
        HVR_RG(TF_S_U8)
        HVR_RG(TF_C_F32_F32)
        HVR_RG(TF_C_F16_F32)
        HVR_RG(TF_C_U8_F32)
        HVR_RG(TF_C_F32_SN8)
        HVR_RG(TF_C_F16_SN8)
        HVR_RG(TF_C_U8_SN8)
        HVR_RG(TF_C_F16_F16)
    default: return HVR_OK;
    }
#undef HVR_RG
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return HVR_OK;
}

extern "C" {

int hvr_abi_version(void) { return HVR_ABI_VERSION; }

int hvr_debug_inject_fault(const char *function, int code, int after)
{
    std::lock_guard<std::mutex> lock(g_fault_mu);
    if (!function) {
        g_fault_armed.store(false);
        g_fault_after = -1;
        g_fault_fn.clear();
        return HVR_OK;
    }
    if (code >= 0 || after < 0) return fail(HVR_EINVAL, "injected faults need a negative code and after >= 0");
    g_fault_fn = function;
    g_fault_code = code;
    g_fault_after = after;
    g_fault_armed.store(true);
    return HVR_OK;
}

const char *hvr_last_error(void) { return g_err.c_str(); }

int hvr_device_count(int *count)
{
    if (!count) return fail(HVR_EINVAL, "count is null");
    HIPCHK(hipGetDeviceCount(count));
    return HVR_OK;
}

int hvr_device_memory(int device, unsigned long long *free_bytes, unsigned long long *total_bytes)
{
    HIPCHK(hipSetDevice(device));
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return HVR_OK;
}

int hvr_create(int device, int width, int height, hvr_ctx **out)
{
    HVR_FAULT_POINT("hvr_create");
    if (!out) return fail(HVR_EINVAL, "out is null");
    *out = nullptr;
    if (width <= 0 || height <= 0 || (long long)width * height > (1ll << 28))
        return fail(HVR_EINVAL, "width and height must be positive (at most 2^28 pixels)");
    HIPCHK(hipSetDevice(device));
    hvr_ctx *c = new hvr_ctx();
    memset(c, 0, sizeof(*c));
    c->device = device;
    c->W = width;
    c->H = height;
    c->brick_log2 = 3;
    c->skip_mode = HVR_SKIP_AUTO;
    c->order_mode = HVR_ORDER_AUTO;
    c->readback_mode = HVR_READBACK_DIRECT;
    const char *gather = getenv("HVR_GATHER");
    c->force_global = gather && strcmp(gather, "global") == 0;
    const size_t npix = (size_t)width * height;
    noalloc_setup(device);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipMalloc(&c->d_out, npix * sizeof(uchar4));
    if (e == hipSuccess) e = hipHostMalloc(&c->h_pinned, npix * 4, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&c->h_pinned_dev, c->h_pinned, 0);
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_nclear, sizeof(uint32_t), hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->flags_ev, hipEventDisableTiming);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&c->lut_ev[i], hipEventDisableTiming);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipEventCreate(&c->ev[i]);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&c->rev[i], hipEventDisableTiming);
    if (e == hipSuccess) {
        // unbound LUT: a 1-texel (0, 0, 0, 1) texture (GL incomplete texture)
        c->lut_cap = 1;
        c->lut_n = 1;
        e = hipMalloc(&c->d_lut, sizeof(float4));
        const float4 unbound = make_float4(0.f, 0.f, 0.f, 1.f);
        if (e == hipSuccess) e = hipMemcpy(c->d_lut, &unbound, sizeof(float4), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemset(c->d_out, 0, npix * sizeof(uchar4));
    if (e != hipSuccess) {
        hvr_destroy(c);
        return fail_hip(e, "hvr_create");
    }
    c->pix_cap = npix;
    *out = c;
    return HVR_OK;
}

int hvr_resize(hvr_ctx *c, int width, int height)
{
    HVR_FAULT_POINT("hvr_resize");
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    if (width <= 0 || height <= 0 || (long long)width * height > (1ll << 28))
        return fail(HVR_EINVAL, "width and height must be positive (at most 2^28 pixels)");
    HIPCHK(hipSetDevice(c->device));
    const size_t npix = (size_t)width * height;
    // Frames in flight keep their own sizes (kernel arguments); only
    // buffers that must grow wait for them.
    if (npix > c->pix_cap) {
        HIPCHK(hipStreamSynchronize(c->stream));
        (void)hipFree(c->d_out);
        if (c->h_pinned) (void)hipHostFree(c->h_pinned);
        c->d_out = nullptr;
        c->h_pinned = nullptr;
        c->h_pinned_dev = nullptr;
        c->last_host = false;
        c->pix_cap = 0;
        HIPCHK(hipMalloc(&c->d_out, npix * sizeof(uchar4)));
        HIPCHK(hipHostMalloc(&c->h_pinned, npix * 4, hipHostMallocDefault));
        HIPCHK(hipHostGetDevicePointer(&c->h_pinned_dev, c->h_pinned, 0));
        HIPCHK(hipMemset(c->d_out, 0, npix * sizeof(uchar4)));
        c->pix_cap = npix;
    }
    c->W = width;
    c->H = height;
    return HVR_OK;
}

int hvr_destroy(hvr_ctx *c)
{
    if (!c) return HVR_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->d_out);
    (void)hipFree(c->d_counts);
    (void)hipFree(c->d_lut);
    (void)hipFree(c->d_stage);
    (void)hipFree(c->d_nz);
    (void)hipFree(c->d_flags);
    (void)hipFree(c->d_nclear);
    (void)hipFree(c->d_order);
    if (c->h_pinned) (void)hipHostFree(c->h_pinned);
    if (c->h_nclear) (void)hipHostFree(c->h_nclear);
    if (c->flags_ev) (void)hipEventDestroy(c->flags_ev);
    for (int i = 0; i < 4; ++i) {
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
        if (c->rev[i]) (void)hipEventDestroy(c->rev[i]);
        if (c->lut_ev[i]) (void)hipEventDestroy(c->lut_ev[i]);
        if (c->h_lut_stage[i]) (void)hipHostFree(c->h_lut_stage[i]);
    }
    if (c->tev) {
        for (hipEvent_t e : *c->tev) (void)hipEventDestroy(e);
        delete c->tev;
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return HVR_OK;
}

int hvr_set_empty_skip(hvr_ctx *c, int mode)
{
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    if (mode != HVR_SKIP_OFF && mode != HVR_SKIP_ON && mode != HVR_SKIP_AUTO)
        return fail(HVR_EINVAL, "skip mode must be HVR_SKIP_OFF, HVR_SKIP_ON or HVR_SKIP_AUTO");
    c->skip_mode = mode;
    return HVR_OK;
}

int hvr_set_readback(hvr_ctx *c, int mode)
{
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    if (mode != HVR_READBACK_COPY && mode != HVR_READBACK_DIRECT)
        return fail(HVR_EINVAL, "readback mode must be HVR_READBACK_COPY or HVR_READBACK_DIRECT");
    c->readback_mode = mode;
    return HVR_OK;
}

int hvr_set_dispatch_order(hvr_ctx *c, int mode)
{
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    if (mode != HVR_ORDER_OFF && mode != HVR_ORDER_ON && mode != HVR_ORDER_AUTO)
        return fail(HVR_EINVAL, "dispatch order mode must be HVR_ORDER_OFF, HVR_ORDER_ON or HVR_ORDER_AUTO");
    c->order_mode = mode;
    return HVR_OK;
}

int hvr_dispatch_order_info(hvr_ctx *c, int *used)
{
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    if (used) *used = c->order_used;
    return HVR_OK;
}

int hvr_skip_info(hvr_ctx *c, unsigned long long *clear_bricks, unsigned long long *bricks, int *used)
{
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    const bool valid = c->flags_vol != nullptr;
    if (valid && !flags_count(c, true)) return fail(HVR_EDEVICE, "hvr_skip_info: reading the brick count failed");
    if (clear_bricks) *clear_bricks = valid ? c->flags_clear : 0;
    if (bricks) *bricks = valid ? c->flags_total : 0;
    if (used) *used = c->skip_used;
    return HVR_OK;
}

int hvr_set_brick_log2(hvr_ctx *c, int lb)
{
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    if (lb < 2 || lb > 4) return fail(HVR_EINVAL, "brick log2 must be in [2, 4]");
    c->brick_log2 = lb;
    return HVR_OK;
}

static int texel_format(int sf, int nf)
{
    if (nf == HVR_NORMAL_NONE) {
        if (sf == HVR_SCALAR_F32) return TF_S_F32;
        if (sf == HVR_SCALAR_F16) return TF_S_F16;
        if (sf == HVR_SCALAR_U8) return TF_S_U8;
        return -1;
    }
    if (nf == HVR_NORMAL_F32X3) {
        if (sf == HVR_SCALAR_NONE) return TF_N_F32;
        if (sf == HVR_SCALAR_F32) return TF_C_F32_F32;
        if (sf == HVR_SCALAR_F16) return TF_C_F16_F32;
        if (sf == HVR_SCALAR_U8) return TF_C_U8_F32;
        return -1;
    }
    if (nf == HVR_NORMAL_SN8X3) {
        if (sf == HVR_SCALAR_NONE) return TF_N_SN8;
        if (sf == HVR_SCALAR_F32) return TF_C_F32_SN8;
        if (sf == HVR_SCALAR_F16) return TF_C_F16_SN8;
        if (sf == HVR_SCALAR_U8) return TF_C_U8_SN8;
        return -1;
    }
    if (nf == HVR_NORMAL_F16X3) {
        // 8-byte texels; FP32 and UNORM8 scalars are not paired with it
        if (sf == HVR_SCALAR_NONE) return TF_N_F16;
        if (sf == HVR_SCALAR_F16) return TF_C_F16_F16;
    }
    return -1;
}

int hvr_volume_create(hvr_ctx *c, int s0, int s1, int s2, int sf, int nf, hvr_volume **out)
{
    HVR_FAULT_POINT("hvr_volume_create");
    if (!c || !out) return fail(HVR_EINVAL, "null argument");
    *out = nullptr;
    if (s0 <= 0 || s1 <= 0 || s2 <= 0) return fail(HVR_EINVAL, "volume dimensions must be positive");
    const int fmt = texel_format(sf, nf);
    if (fmt < 0) return fail(HVR_EINVAL, "unsupported scalar/normal format combination");
    if (s0 > (1 << 20) || s1 > (1 << 20) || s2 > (1 << 20))
        return fail(HVR_EINVAL, "volume dimension above 2^20");
    const int lb = c->brick_log2;
    const uint64_t E = (1u << lb) + 1u;
    hvr_volume *v = new hvr_volume();
    memset(v, 0, sizeof(*v));
    v->ctx = c;
    v->s0 = s0;
    v->s1 = s1;
    v->s2 = s2;
    v->fmt = fmt;
    v->lb = lb;
    v->nbx = bricks_along(s0, lb);
    v->nby = bricks_along(s1, lb);
    v->nbz = bricks_along(s2, lb);
    v->brick_bytes = (uint32_t)(E * E * E * (uint64_t)texel_bytes(fmt));
    const uint64_t nbricks = (uint64_t)v->nbx * v->nby * v->nbz;
    if (nbricks >= (1ull << 32)) {
        delete v;
        return fail(HVR_EINVAL, "volume too large (brick index exceeds 32 bits)");
    }
    v->bytes = nbricks * v->brick_bytes;
    const size_t bytes = (size_t)v->bytes;
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) e = hipMalloc(&v->d_texels, bytes);
    if (e == hipSuccess) e = hipMemsetAsync(v->d_texels, 0, bytes, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        (void)hipFree(v->d_texels);
        delete v;
        return fail_hip(e, "hvr_volume_create");
    }
    *out = v;
    return HVR_OK;
}

static int upload_part(hvr_volume *v, const float *data, int part)
{
    hvr_ctx *c = v->ctx;
    HIPCHK(hipSetDevice(c->device));
    const int comps = part == 1 ? 1 : 3;
    const uint64_t plane = (uint64_t)v->s0 * v->s1;
    // stage whole z-planes, up to 256 MiB per chunk
    const size_t max_stage = (size_t)256 << 20;
    uint64_t zper = max_stage / (plane * comps * sizeof(float));
    if (zper < 1) zper = 1;
    if (zper > (uint64_t)v->s2) zper = v->s2;
    const size_t need = (size_t)(zper * plane * comps * sizeof(float));
    if (c->stage_bytes < need) {
        HIPCHK(hipStreamSynchronize(c->stream));
        (void)hipFree(c->d_stage);
        c->d_stage = nullptr;
        c->stage_bytes = 0;
        HIPCHK(hipMalloc(&c->d_stage, need));
        c->stage_bytes = need;
    }
    for (uint64_t z0 = 0; z0 < (uint64_t)v->s2; z0 += zper) {
        const uint64_t nz = (z0 + zper <= (uint64_t)v->s2) ? zper : v->s2 - z0;
        const uint64_t nvox = nz * plane;
        HIPCHK(hipMemcpyAsync(c->d_stage, data + z0 * plane * comps, nvox * comps * sizeof(float),
                              hipMemcpyHostToDevice, c->stream));
        const unsigned nblk = (unsigned)((nvox + 255) / 256);
#define HVR_UP(F)                                                                                      \
    case F:                                                                                            \
        hipLaunchKernelGGL(k_upload<F>, dim3(nblk), dim3(256), 0, c->stream, (char *)v->d_texels, v->s0, \
                           v->s1, v->s2, v->lb, v->nbx, v->nby, v->nbz, v->brick_bytes,                     \
                           (const float *)c->d_stage,                                                   \
                           part, (int)z0, nvox);                                                        \
        break;
        switch (v->fmt) {
            HVR_UP(TF_S_F32)
            HVR_UP(TF_S_F16)
            HVR_UP(TF_S_U8)
            HVR_UP(TF_N_F32)
            HVR_UP(TF_N_SN8)
            HVR_UP(TF_C_F32_F32)
            HVR_UP(TF_C_F16_F32)
            HVR_UP(TF_C_U8_F32)
            HVR_UP(TF_C_F32_SN8)
            HVR_UP(TF_C_F16_SN8)
            HVR_UP(TF_C_U8_SN8)
            HVR_UP(TF_C_F16_F16)
            HVR_UP(TF_N_F16)
        default: return fail(HVR_EINVAL, "bad texel format");
        }
#undef HVR_UP
        HIPCHK(hipGetLastError());
        // the staging buffer is reused by the next chunk
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return HVR_OK;
}

// Scalar part (and, with `normals`, the normal part computed on the device)
// of volume v from a device float32 field in numpy C order.
static int volume_from_field(hvr_volume *v, const float *d_field, bool normals)
{
    hvr_ctx *c = v->ctx;
    const uint64_t nvox = (uint64_t)v->s0 * v->s1 * v->s2;
    const unsigned nblk = (unsigned)((nvox + 255) / 256);
#define HVR_UPF(F)                                                                                           \
    case F:                                                                                                  \
        if (fmt_has_scalar(F))                                                                               \
            hipLaunchKernelGGL(k_upload<F>, dim3(nblk), dim3(256), 0, c->stream, (char *)v->d_texels, v->s0,   \
                               v->s1, v->s2, v->lb, v->nbx, v->nby, v->nbz, v->brick_bytes, d_field, 1, 0, nvox); \
        if (normals && fmt_has_normal(F))                                                                    \
            hipLaunchKernelGGL(k_normals_from_field<F>, dim3(nblk), dim3(256), 0, c->stream,                 \
                               (char *)v->d_texels, v->s0, v->s1, v->s2, v->lb, v->nbx, v->nby, v->nbz,      \
                               v->brick_bytes, d_field, nvox);                                               \
        break;
    switch (v->fmt) {
        HVR_UPF(TF_S_F32)
        HVR_UPF(TF_S_F16)
        HVR_UPF(TF_S_U8)
        HVR_UPF(TF_C_F32_F32)
        HVR_UPF(TF_C_F16_F32)
        HVR_UPF(TF_C_U8_F32)
        HVR_UPF(TF_C_F32_SN8)
        HVR_UPF(TF_C_F16_SN8)
        HVR_UPF(TF_C_U8_SN8)
        HVR_UPF(TF_C_F16_F16)
    default: return fail(HVR_EINVAL, "the volume has no scalar field");
    }
#undef HVR_UPF
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    v->scalar_loaded = true;
    if (normals && fmt_has_normal(v->fmt)) v->normals_loaded = true;
    ++v->gen;
    return compute_range(v);
}

int hvr_volume_upload_scalar_normals(hvr_volume *v, const float *data)
{
    HVR_FAULT_POINT("hvr_volume_upload_scalar_normals");
    if (!v || !data) return fail(HVR_EINVAL, "null argument");
    if (!fmt_has_scalar(v->fmt) || !fmt_has_normal(v->fmt))
        return fail(HVR_ESTATE, "the volume needs scalar and normal fields");
    if (v->s0 < 2 || v->s1 < 2 || v->s2 < 2)
        return fail(HVR_EINVAL, "normals need at least 2 voxels along every axis (np.gradient)");
    hvr_ctx *c = v->ctx;
    HIPCHK(hipSetDevice(c->device));
    const size_t bytes = (size_t)v->s0 * v->s1 * v->s2 * sizeof(float);
    float *d_field = nullptr;
    HIPCHK(hipMalloc(&d_field, bytes));
    int rc = HVR_OK;
    if (hipMemcpy(d_field, data, bytes, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(HVR_EDEVICE, "copying the scalar field to the device failed");
    if (rc == HVR_OK) rc = volume_from_field(v, d_field, true);
    (void)hipStreamSynchronize(c->stream);   // also after a failure: nothing queued may still read the field
    (void)hipFree(d_field);
    return rc;
}

int hvr_volume_generate_blobs(hvr_volume *v, int n_blobs, const double *centers, const double *sigmas,
                              const double *amps, double noise, unsigned long long seed)
{
    HVR_FAULT_POINT("hvr_volume_generate_blobs");
    if (!v || (n_blobs > 0 && (!centers || !sigmas || !amps))) return fail(HVR_EINVAL, "null argument");
    if (n_blobs < 0 || n_blobs > kMaxBlobs) return fail(HVR_EINVAL, "0 to 32 blobs");
    if (!fmt_has_scalar(v->fmt)) return fail(HVR_ESTATE, "the volume has no scalar field");
    const bool normals = fmt_has_normal(v->fmt);
    if (normals && (v->s0 < 2 || v->s1 < 2 || v->s2 < 2))
        return fail(HVR_EINVAL, "normals need at least 2 voxels along every axis (np.gradient)");
    hvr_ctx *c = v->ctx;
    HIPCHK(hipSetDevice(c->device));
    BlobSpec spec;
    memset(&spec, 0, sizeof(spec));
    spec.n = n_blobs;
    for (int b = 0; b < n_blobs; ++b) {
        for (int k = 0; k < 3; ++k) spec.c[3 * b + k] = centers[3 * b + k];
        spec.s[b] = sigmas[b];
        spec.a[b] = amps[b];
    }
    spec.noise = noise;
    spec.seed = seed;
    const uint64_t nvox = (uint64_t)v->s0 * v->s1 * v->s2;
    float *d_field = nullptr;
    BlobSpec *d_spec = nullptr;
    HIPCHK(hipMalloc(&d_field, nvox * sizeof(float)));
    int rc = HVR_OK;
    if (hipMalloc(&d_spec, sizeof(spec)) != hipSuccess ||
        hipMemcpy(d_spec, &spec, sizeof(spec), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(HVR_EDEVICE, "staging the blob parameters failed");
    if (rc == HVR_OK) {
        hipLaunchKernelGGL(k_blob_field, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, c->stream, d_field,
                           v->s0, v->s1, v->s2, (const BlobSpec *)d_spec, nvox);
        if (hipGetLastError() != hipSuccess) rc = fail(HVR_EDEVICE, "blob field kernel launch failed");
    }
    if (rc == HVR_OK) rc = volume_from_field(v, d_field, normals);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(d_spec);
    (void)hipFree(d_field);
    return rc;
}

int hvr_volume_upload_scalar(hvr_volume *v, const float *data)
{
    HVR_FAULT_POINT("hvr_volume_upload_scalar");
    if (!v || !data) return fail(HVR_EINVAL, "null argument");
    if (!fmt_has_scalar(v->fmt)) return fail(HVR_ESTATE, "volume has no scalar field");
    int rc = upload_part(v, data, 1);
    if (rc != HVR_OK) return rc;
    v->scalar_loaded = true;
    ++v->gen;
    return compute_range(v);
}

int hvr_volume_upload_normals(hvr_volume *v, const float *data)
{
    HVR_FAULT_POINT("hvr_volume_upload_normals");
    if (!v || !data) return fail(HVR_EINVAL, "null argument");
    if (!fmt_has_normal(v->fmt)) return fail(HVR_ESTATE, "volume has no normal field");
    int rc = upload_part(v, data, 2);
    if (rc == HVR_OK) {
        v->normals_loaded = true;
        ++v->gen;
    }
    return rc;
}

int hvr_volume_bytes(const hvr_volume *v, uint64_t *bytes)
{
    if (!v || !bytes) return fail(HVR_EINVAL, "null argument");
    *bytes = v->bytes;
    return HVR_OK;
}

int hvr_volume_destroy(hvr_volume *v)
{
    if (!v) return HVR_OK;
    hvr_ctx *c = v->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->scalar_vol == v) c->scalar_vol = nullptr;
    if (c->normal_vol == v) c->normal_vol = nullptr;
    if (c->flags_vol == v) c->flags_vol = nullptr;
    (void)hipFree(v->d_texels);
    (void)hipFree(v->d_range);
    delete v;
    return HVR_OK;
}

int hvr_set_lut(hvr_ctx *c, const float *rgba, int n)
{
    HVR_FAULT_POINT("hvr_set_lut");
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    if (n < 0 || (n > 0 && !rgba)) return fail(HVR_EINVAL, "bad LUT");
    HIPCHK(hipSetDevice(c->device));
    const float unbound[4] = {0.f, 0.f, 0.f, 1.f};
    const float *src = n > 0 ? rgba : unbound;
    const int cnt = n > 0 ? n : 1;
    // The copies are ordered on the stream after the frames already queued
    // (which read the previous LUT), through pinned staging, so a LUT change
    // does not wait for them; growing a buffer does.
    const size_t lut_bytes = (size_t)cnt * sizeof(float4), nz_bytes = ((size_t)cnt + 1) * sizeof(int);
    if (cnt > c->lut_cap || cnt > c->nz_cap || lut_bytes + nz_bytes > c->lut_stage_bytes)
        HIPCHK(hipStreamSynchronize(c->stream));
    if (cnt > c->lut_cap) {
        (void)hipFree(c->d_lut);
        c->d_lut = nullptr;
        c->lut_cap = 0;
        HIPCHK(hipMalloc(&c->d_lut, lut_bytes));
        c->lut_cap = cnt;
    }
    if (cnt > c->nz_cap) {
        (void)hipFree(c->d_nz);
        c->d_nz = nullptr;
        c->nz_cap = 0;
        HIPCHK(hipMalloc(&c->d_nz, nz_bytes));
        c->nz_cap = cnt;
    }
    if (lut_bytes + nz_bytes > c->lut_stage_bytes) {
        size_t want = 4096;
        while (want < lut_bytes + nz_bytes) want *= 2;
        for (int i = 0; i < 4; ++i) {
            if (c->h_lut_stage[i]) (void)hipHostFree(c->h_lut_stage[i]);
            c->h_lut_stage[i] = nullptr;
            c->lut_ev_live[i] = false;
        }
        c->lut_stage_bytes = 0;
        for (int i = 0; i < 4; ++i) HIPCHK(hipHostMalloc(&c->h_lut_stage[i], want, hipHostMallocDefault));
        c->lut_stage_bytes = want;
    }
    const int k = c->lut_slot;
    c->lut_slot = (k + 1) & 3;
    if (c->lut_ev_live[k]) HIPCHK(hipEventSynchronize(c->lut_ev[k]));
    char *stage = (char *)c->h_lut_stage[k];
    memcpy(stage, src, lut_bytes);
    lut_prefix_counts(src, cnt, (int *)(stage + lut_bytes));
    HIPCHK(hipMemcpyAsync(c->d_lut, stage, lut_bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_nz, stage + lut_bytes, nz_bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->lut_ev[k], c->stream));
    c->lut_ev_live[k] = true;
    c->lut_n = cnt;
    ++c->lut_gen;
    return HVR_OK;
}

static int check_params(const hvr_params *p)
{
    if (!(p->step > 0.f) || !(p->ref_step > 0.f)) return fail(HVR_EINVAL, "step sizes must be positive");
    if (p->max_steps < 0) return fail(HVR_EINVAL, "max_steps must be >= 0");
    for (int k = 0; k < 3; ++k) {
        // the float extent of the in-box test must be positive, with the
        // texel scale (size / extent) finite
        const float ext = p->bmax[k] - p->bmin[k];
        if (!std::isfinite(p->bmin[k]) || !std::isfinite(p->bmax[k]) || !(ext >= 1e-20f) || !std::isfinite(ext))
            return fail(HVR_EINVAL, "box bounds must be finite with bmax - bmin >= 1e-20 on every axis");
    }
    return HVR_OK;
}

int hvr_set_params(hvr_ctx *c, const hvr_params *p)
{
    HVR_FAULT_POINT("hvr_set_params");
    if (!c || !p) return fail(HVR_EINVAL, "null argument");
    if (const int rc = check_params(p)) return rc;
    c->params = *p;
    c->params_set = true;
    return HVR_OK;
}

int hvr_bind(hvr_ctx *c, hvr_volume *s, int srole, hvr_volume *n, int nrole)
{
    HVR_FAULT_POINT("hvr_bind");
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    if ((s && s->ctx != c) || (n && n->ctx != c)) return fail(HVR_EINVAL, "volume belongs to another context");
    if (srole != HVR_ROLE_SCALAR && srole != HVR_ROLE_NORMAL) return fail(HVR_EINVAL, "bad scalar role");
    if (nrole != HVR_ROLE_SCALAR && nrole != HVR_ROLE_NORMAL) return fail(HVR_EINVAL, "bad normal role");
    for (int k = 0; k < 2; ++k) {
        const hvr_volume *v = k ? n : s;
        const int role = k ? nrole : srole;
        if (!v) continue;
        if (role == HVR_ROLE_SCALAR && !fmt_has_scalar(v->fmt)) return fail(HVR_EINVAL, "volume has no scalar field");
        if (role == HVR_ROLE_NORMAL && !fmt_has_normal(v->fmt)) return fail(HVR_EINVAL, "volume has no normal field");
    }
    c->scalar_vol = s;
    c->normal_vol = n;
    c->scalar_role = srole;
    c->normal_role = nrole;
    return HVR_OK;
}

int hvr_render_async(hvr_ctx *c)
{
    HVR_FAULT_POINT("hvr_render_async");
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    HIPCHK(hipSetDevice(c->device));
    return enqueue_frame(c, false, nullptr);
}

int hvr_render_async_to(hvr_ctx *c, void *dst)
{
    HVR_FAULT_POINT("hvr_render_async_to");
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    HIPCHK(hipSetDevice(c->device));
    return enqueue_frame(c, false, (uchar4 *)dst);
}

int hvr_device_alloc(hvr_ctx *c, uint64_t bytes, void **ptr)
{
    if (!c || !ptr) return fail(HVR_EINVAL, "null argument");
    *ptr = nullptr;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMalloc(ptr, bytes ? bytes : 1));
    return HVR_OK;
}

int hvr_device_free(hvr_ctx *c, void *ptr)
{
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(ptr));
    return HVR_OK;
}

int hvr_copy_to_host(hvr_ctx *c, void *host, const void *dev, uint64_t bytes)
{
    if (!c || !host || !dev) return fail(HVR_EINVAL, "null argument");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return HVR_OK;
}

int hvr_device_copy(hvr_ctx *c, void *dst, const void *src, uint64_t bytes)
{
    if (!c || !dst || !src) return fail(HVR_EINVAL, "null argument");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return HVR_OK;
}

int hvr_timing_begin(hvr_ctx *c, int max_frames)
{
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    if (max_frames <= 0 || max_frames > (1 << 16)) return fail(HVR_EINVAL, "max_frames out of range");
    HIPCHK(hipSetDevice(c->device));
    if (!c->tev) c->tev = new std::vector<hipEvent_t>();
    while (c->tev->size() < 3 * (size_t)max_frames) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        c->tev->push_back(e);
    }
    c->timing_max = max_frames;
    c->timing_n = 0;
    return HVR_OK;
}

int hvr_timing_end(hvr_ctx *c, double *march_ms, double *setup_ms, int *frames)
{
    if (!c || !c->tev) return fail(HVR_ESTATE, "no timing region");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    double m = 0.0, s = 0.0;
    for (int i = 0; i < c->timing_n; ++i) {
        float a = 0.f, b = 0.f;
        HIPCHK(hipEventElapsedTime(&a, (*c->tev)[3 * i], (*c->tev)[3 * i + 1]));
        HIPCHK(hipEventElapsedTime(&b, (*c->tev)[3 * i + 1], (*c->tev)[3 * i + 2]));
        s += a;
        m += b;
    }
    if (march_ms) *march_ms = m;
    if (setup_ms) *setup_ms = s;
    if (frames) *frames = c->timing_n;
    c->timing_max = 0;
    c->timing_n = 0;
    return HVR_OK;
}

// The copy runs through the pinned buffer in chunks of at least 1 MiB (at
// most four): each chunk's host copy into `out` starts as soon as its
// device-to-host copy has landed, overlapping the transfer of the next.
// Images up to 1 MiB take one copy, since the per-chunk synchronisation
// costs more than the overlap saves there.

int hvr_read(hvr_ctx *c, uint8_t *out)
{
    HVR_FAULT_POINT("hvr_read");
    if (!c || !out) return fail(HVR_EINVAL, "null argument");
    HIPCHK(hipSetDevice(c->device));
    const size_t bytes = (size_t)c->W * c->H * 4;
    if (c->last_host) {
        // the march wrote the frame to the pinned buffer: wait, copy once
        HIPCHK(hipStreamSynchronize(c->stream));
        hvr::CopyPool::get().copy(out, (const uint8_t *)c->h_pinned, bytes);
        return HVR_OK;
    }
    constexpr int KMAX = 4;
    const int K = (int)std::min<size_t>(KMAX, std::max<size_t>(1, bytes >> 20));
    const size_t chunk = ((bytes + K - 1) / K + 4095) & ~(size_t)4095;
    size_t off[KMAX + 1];
    for (int k = 0; k <= K; ++k) off[k] = std::min(bytes, (size_t)k * chunk);
    uint8_t *h = (uint8_t *)c->h_pinned;
    const uint8_t *d = (const uint8_t *)c->d_out;
    for (int k = 0; k < K; ++k) {
        if (off[k + 1] > off[k])
            HIPCHK(hipMemcpyAsync(h + off[k], d + off[k], off[k + 1] - off[k], hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipEventRecord(c->rev[k], c->stream));
    }
    for (int k = 0; k < K; ++k) {
        HIPCHK(hipEventSynchronize(c->rev[k]));
        memcpy(out + off[k], h + off[k], off[k + 1] - off[k]);
    }
    return HVR_OK;
}

int hvr_read_pinned(hvr_ctx *c, const uint8_t **rgba)
{
    if (!c || !rgba) return fail(HVR_EINVAL, "null argument");
    *rgba = nullptr;
    HIPCHK(hipSetDevice(c->device));
    const size_t bytes = (size_t)c->W * c->H * 4;
    if (!c->last_host) HIPCHK(hipMemcpyAsync(c->h_pinned, c->d_out, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *rgba = (const uint8_t *)c->h_pinned;
    return HVR_OK;
}

int hvr_render(hvr_ctx *c, uint8_t *out)
{
    int rc = hvr_render_async(c);
    if (rc != HVR_OK) return rc;
    return hvr_read(c, out);
}

int hvr_synchronize(hvr_ctx *c)
{
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return HVR_OK;
}

int hvr_framebuffer(hvr_ctx *c, void **ptr)
{
    if (!c || !ptr) return fail(HVR_EINVAL, "null argument");
    *ptr = c->d_out;
    return HVR_OK;
}

static int sample_counts(hvr_ctx *c, uint32_t *counts, bool shaded)
{
    if (!c || !counts) return fail(HVR_EINVAL, "null argument");
    HIPCHK(hipSetDevice(c->device));
    const size_t npix = (size_t)c->W * c->H;
    if (c->counts_cap < npix) {
        (void)hipFree(c->d_counts);
        c->d_counts = nullptr;
        c->counts_cap = 0;
        HIPCHK(hipMalloc(&c->d_counts, npix * sizeof(uint32_t)));
        c->counts_cap = npix;
    }
    int rc = enqueue_frame(c, true, nullptr, shaded);
    if (rc != HVR_OK) return rc;
    HIPCHK(hipMemcpyAsync(counts, c->d_counts, npix * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return HVR_OK;
}

int hvr_sample_counts(hvr_ctx *c, uint32_t *counts)
{
    HVR_FAULT_POINT("hvr_sample_counts");
    return sample_counts(c, counts, false);
}

int hvr_shaded_sample_counts(hvr_ctx *c, uint32_t *counts)
{
    return sample_counts(c, counts, true);
}

int hvr_touched_bricks(hvr_ctx *c, unsigned long long *touched, unsigned long long *bricks,
                       unsigned long long *brick_bytes)
{
    if (!c || !touched) return fail(HVR_EINVAL, "null argument");
    const hvr_volume *S = c->scalar_vol;
    if (!S || !(S->lb == 3 || S->lb == 4) || c->lut_n > kLutLdsMax)
        return fail(HVR_ESTATE, "hvr_touched_bricks: needs a bound scalar volume with 8^3 or 16^3 bricks and a LUT "
                                "of at most 2048 entries");
    HIPCHK(hipSetDevice(c->device));
    const size_t npix = (size_t)c->W * c->H;
    if (c->counts_cap < npix) {
        (void)hipFree(c->d_counts);
        c->d_counts = nullptr;
        c->counts_cap = 0;
        HIPCHK(hipMalloc(&c->d_counts, npix * sizeof(uint32_t)));
        c->counts_cap = npix;
    }
    const uint64_t nb = (uint64_t)S->nbx * S->nby * S->nbz;
    uint8_t *d_marks = nullptr;
    HIPCHK(hipMalloc(&d_marks, nb));
    std::vector<uint8_t> marks(nb);
    int rc = hipMemsetAsync(d_marks, 0, nb, c->stream) == hipSuccess ? HVR_OK : fail(HVR_EDEVICE, "memset");
    if (rc == HVR_OK) rc = enqueue_frame(c, true, nullptr, false, d_marks);
    if (rc == HVR_OK && (hipMemcpyAsync(marks.data(), d_marks, nb, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                         hipStreamSynchronize(c->stream) != hipSuccess))
        rc = fail(HVR_EDEVICE, "hvr_touched_bricks: reading the marks failed");
    (void)hipStreamSynchronize(c->stream);   // nothing queued may still write the marks
    (void)hipFree(d_marks);
    if (rc != HVR_OK) return rc;
    unsigned long long n = 0;
    for (uint8_t m : marks) n += m != 0;
    *touched = n;
    if (bricks) *bricks = nb;
    if (brick_bytes) *brick_bytes = S->brick_bytes;
    return HVR_OK;
}

int hvr_last_kernel_ms(hvr_ctx *c, float *march_ms, float *setup_ms)
{
    if (!c) return fail(HVR_EINVAL, "ctx is null");
    if (!c->events_recorded) return fail(HVR_ESTATE, "no frame rendered yet");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipEventSynchronize(c->last_ev[2]));
    float a = 0.f, b = 0.f;
    HIPCHK(hipEventElapsedTime(&a, c->last_ev[0], c->last_ev[1]));
    HIPCHK(hipEventElapsedTime(&b, c->last_ev[1], c->last_ev[2]));
    if (setup_ms) *setup_ms = a;
    if (march_ms) *march_ms = b;
    return HVR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: requests staged on the host, rendered by one persistent launch
// per kernel class, captured once in a hipGraph per batch shape (classes,
// LUT capacity, staging block) and replayed with new staged contents.

struct BatchReq {
    hvr_volume *vol;
    int nm;                  // NM_ZERO / NM_ALIAS / NM_SAME
    hvr_params params;
    int W, H;
    std::vector<float> lut;  // n x RGBA
};

struct hvr_batch {
    hvr_ctx *ctx;
    hipStream_t stream;
    std::vector<BatchReq> reqs;
    int tile_shape;                       // HVR_TILE_8X8 / HVR_TILE_16X4
    int variant;                          // kernel variant of non-counting runs (BV_*)
    char *h_stage, *d_stage;              // staged block (pinned host, device)
    size_t stage_cap;
    uchar4 *d_out;                        // output pool, one RGBA8 image per request
    uint8_t *h_out;                       // its pinned host copy
    uchar4 *h_out_dev;                    // h_out's device address (HVR_READBACK_DIRECT)
    const uint8_t *h_out_dev_of;          // the h_out that h_out_dev maps
    size_t out_cap;
    int readback;                         // HVR_READBACK_DIRECT (default) / HVR_READBACK_COPY
    uint32_t *d_counts, *h_counts;        // sample-count pool (counting runs)
    size_t counts_cap;                    // in pixels
    uint8_t *d_flags;                     // transparent-brick flags, per request
    size_t flags_cap;
    uint32_t *d_counters;                 // tile counter per class
    std::vector<size_t> out_off;          // per request (add order): pixel offset in the pools
    hipGraph_t graph;
    hipGraphExec_t exec;
    std::string key;                      // shape the graph was captured for
    hipEvent_t ev0, ev1, ev2;             // graph start / end, copies back done
    bool ran, counted;
    size_t out_pixels;
    int ncu;
};

// The serving layer's memory budget bounds a staged request by 512 bytes of
// descriptor (pyvr_amd/hip_renderer/serve.py, batch_pool_bytes).
static_assert(sizeof(ReqDesc) <= 512, "request descriptor above the serving budget's bound");

// The batched launcher for a class: format, normal mode, brick log2, counting,
// ray tile, buffer loads and kernel variant (BV_*); nullptr when the library
// has no such instance.  buf: the volume is read through buffer loads
// (below 4 GiB and 2^24 bricks).
static BatchLaunchFn batch_fn(int fmt, int nm, int lb, bool count, int ts, bool buf, int variant = BV_DEFAULT)
{
    switch (fmt) {
#define HVR_CASE(F) \
    case F: return batch_fn_##F(nm, lb, count, ts, buf, variant);
        HVR_FAST_FORMATS(HVR_CASE)
#undef HVR_CASE
    default: return nullptr;
    }
}

static bool volume_buf(const hvr_volume *v)
{
    return v->bytes <= 0xffffffffull && (uint64_t)v->nbx * v->nby * v->nbz < (1ull << 24);
}

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Grow a device (and optionally pinned host) buffer to at least `need`
// bytes.  Returns true when it was reallocated.
static int grow(void **d, void **h, size_t *cap, size_t need, bool *moved)
{
    if (need <= *cap) return HVR_OK;
    size_t n = *cap ? *cap : 4096;
    while (n < need) n *= 2;
    (void)hipFree(*d);
    *d = nullptr;
    if (h) {
        (void)hipHostFree(*h);
        *h = nullptr;
    }
    *cap = 0;
    HIPCHK(hipMalloc(d, n));
    if (h) HIPCHK(hipHostMalloc(h, n, hipHostMallocDefault));
    *cap = n;
    *moved = true;
    return HVR_OK;
}

static void batch_drop_graph(hvr_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->key.clear();
}

extern "C" {

int hvr_batch_create(hvr_ctx *c, hvr_batch **out)
{
    HVR_FAULT_POINT("hvr_batch_create");
    if (!c || !out) return fail(HVR_EINVAL, "null argument");
    *out = nullptr;
    HIPCHK(hipSetDevice(c->device));
    hvr_batch *b = new hvr_batch();
    b->ctx = c;
    b->tile_shape = HVR_TILE_8X8;
    b->variant = BV_DEFAULT;
    b->readback = HVR_READBACK_DIRECT;
    b->exec = nullptr;
    b->graph = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev2, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&b->d_counters, kBatchMaxClasses * sizeof(uint32_t));
    if (e == hipSuccess) e = hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, c->device);
    if (e != hipSuccess) {
        hvr_batch_destroy(b);
        return fail_hip(e, "hvr_batch_create");
    }
    *out = b;
    return HVR_OK;
}

int hvr_batch_destroy(hvr_batch *b)
{
    if (!b) return HVR_OK;
    (void)hipSetDevice(b->ctx->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graph(b);
    (void)hipFree(b->d_stage);
    (void)hipFree(b->d_out);
    (void)hipFree(b->d_counts);
    (void)hipFree(b->d_flags);
    (void)hipFree(b->d_counters);
    if (b->h_stage) (void)hipHostFree(b->h_stage);
    if (b->h_out) (void)hipHostFree(b->h_out);
    if (b->h_counts) (void)hipHostFree(b->h_counts);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->ev2) (void)hipEventDestroy(b->ev2);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return HVR_OK;
}

int hvr_batch_clear(hvr_batch *b)
{
    if (!b) return fail(HVR_EINVAL, "batch is null");
    b->reqs.clear();
    b->ran = b->counted = false;
    return HVR_OK;
}

int hvr_batch_set_tile_shape(hvr_batch *b, int shape)
{
    if (!b) return fail(HVR_EINVAL, "batch is null");
    if (shape != HVR_TILE_8X8 && shape != HVR_TILE_16X4) return fail(HVR_EINVAL, "unknown tile shape");
    b->tile_shape = shape;
    return HVR_OK;
}

int hvr_batch_set_readback(hvr_batch *b, int mode)
{
    if (!b) return fail(HVR_EINVAL, "batch is null");
    if (mode != HVR_READBACK_COPY && mode != HVR_READBACK_DIRECT)
        return fail(HVR_EINVAL, "readback mode must be HVR_READBACK_COPY or HVR_READBACK_DIRECT");
    // the next run's descriptors point the march at the chosen pool; a run
    // in flight keeps the one it was enqueued with, and its images end in
    // the pinned pool either way
    b->readback = mode;
    return HVR_OK;
}

int hvr_batch_set_kernel(hvr_batch *b, int waves_per_block, int skip, int gate)
{
    if (!b) return fail(HVR_EINVAL, "batch is null");
    int v = -1;
    if (waves_per_block == 4 && skip && gate) v = BV_DEFAULT;
    else if (waves_per_block == 2 && skip && gate) v = BV_WAVES2;
    else if (waves_per_block == 8 && skip && gate) v = BV_WAVES8;
    else if (waves_per_block == 4 && skip && !gate) v = BV_NOGATE;
    else if (waves_per_block == 4 && !skip && gate) v = BV_NOSKIP;
    if (v < 0)
        return fail(HVR_EINVAL, "kernel variants: 4 waves with skipping and gating, or one of 2 / 8 waves, "
                                "no gating, no skipping");
    b->variant = v;
    return HVR_OK;
}

int hvr_batch_add(hvr_batch *b, hvr_volume *s, int srole, hvr_volume *n, int nrole, const float *lut_rgba,
                  int lut_n, const hvr_params *p, int width, int height, int *index)
{
    HVR_FAULT_POINT("hvr_batch_add");
    if (!b || !p) return fail(HVR_EINVAL, "null argument");
    if (!s || s->ctx != b->ctx || (n && n->ctx != b->ctx))
        return fail(HVR_EINVAL, "a batch request needs a scalar volume of the batch's context");
    if (srole != HVR_ROLE_SCALAR || !fmt_has_scalar(s->fmt) || !s->scalar_loaded || !s->d_range)
        return fail(HVR_EINVAL, "the scalar sampler must read an uploaded scalar field");
    int nm;
    if (!n) nm = NM_ZERO;
    else if (n == s && nrole == HVR_ROLE_SCALAR) nm = NM_ALIAS;
    else if (n == s && nrole == HVR_ROLE_NORMAL && fmt_has_normal(s->fmt)) nm = NM_SAME;
    else return fail(HVR_EINVAL, "a batch request reads its normals from its scalar volume, or none");
    if (!batch_fn(s->fmt, nm, s->lb, false, HVR_TILE_8X8, true))
        return fail(HVR_EINVAL, "this volume layout and normal binding need the single-request path");
    if (lut_n < 0 || lut_n > kLutLdsMax || (lut_n > 0 && !lut_rgba))
        return fail(HVR_EINVAL, "batched LUTs hold 0 to 2048 entries");
    if (width <= 0 || height <= 0 || (long long)width * height > (1ll << 28))
        return fail(HVR_EINVAL, "width and height must be positive (at most 2^28 pixels)");
    if (const int rc = check_params(p)) return rc;
    BatchReq r;
    r.vol = s;
    r.nm = nm;
    r.params = *p;
    r.W = width;
    r.H = height;
    if (lut_n > 0) r.lut.assign(lut_rgba, lut_rgba + 4 * (size_t)lut_n);
    else r.lut = {0.f, 0.f, 0.f, 1.f};  // unbound LUT texture
    b->reqs.push_back(std::move(r));
    b->ran = b->counted = false;
    if (index) *index = (int)b->reqs.size() - 1;
    return HVR_OK;
}

int hvr_batch_reserve(hvr_batch *b, int max_requests, unsigned long long max_pixels, unsigned long long max_bricks)
{
    if (!b || max_requests < 0) return fail(HVR_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(b->ctx->device));
    // growing frees buffers the previous run may still read
    HIPCHK(hipStreamSynchronize(b->stream));
    const size_t nreq = (size_t)max_requests;
    // the largest staged block: header, descriptors, LUTs of 2048 entries
    // and their prefix counts (the layout of hvr_batch_run_async)
    const size_t desc_off = align_up(sizeof(BatchHeader), 256);
    const size_t lut_off = align_up(desc_off + nreq * sizeof(ReqDesc), 256);
    const size_t nz_off = align_up(lut_off + nreq * kLutLdsMax * sizeof(float4), 256);
    const size_t stage = align_up(nz_off + nreq * (kLutLdsMax + 1) * sizeof(int), 256);
    const size_t pixels = (size_t)max_pixels + 64 * nreq;   // each request padded to 64 pixels
    const size_t flags = nreq * align_up((size_t)max_bricks, 256);
    bool moved = false, moved_out = false;
    if (const int rc = grow((void **)&b->d_stage, (void **)&b->h_stage, &b->stage_cap, stage, &moved)) return rc;
    if (const int rc = grow((void **)&b->d_out, (void **)&b->h_out, &b->out_cap, pixels * 4, &moved_out)) return rc;
    if (const int rc = grow((void **)&b->d_flags, nullptr, &b->flags_cap, flags, &moved_out)) return rc;
    if (moved_out) b->h_out_dev_of = nullptr;   // a new pinned pool may reuse the old address
    if (moved) batch_drop_graph(b);
    return HVR_OK;
}

int hvr_batch_run_async(hvr_batch *b, int count)
{
    HVR_FAULT_POINT("hvr_batch_run_async");
    if (!b) return fail(HVR_EINVAL, "batch is null");
    hvr_ctx *c = b->ctx;
    HIPCHK(hipSetDevice(c->device));
    // the previous run's copies read and write the pinned buffers
    HIPCHK(hipStreamSynchronize(b->stream));
    const int nreq = (int)b->reqs.size();
    b->ran = b->counted = false;
    if (nreq == 0) return HVR_OK;
    // classes: one kernel instance each, in order of first appearance
    struct Cls {
        int fmt, nm, lb;
        bool buf;
        std::vector<int> members;
    };
    std::vector<Cls> cls;
    int lut_cap = 1;
    for (int i = 0; i < nreq; ++i) {
        const BatchReq &r = b->reqs[i];
        const bool buf = volume_buf(r.vol) && !c->force_global;
        int k = 0;
        while (k < (int)cls.size() && !(cls[k].fmt == r.vol->fmt && cls[k].nm == r.nm && cls[k].lb == r.vol->lb &&
                                        cls[k].buf == buf))
            ++k;
        if (k == (int)cls.size()) {
            if (k == kBatchMaxClasses) return fail(HVR_EINVAL, "a batch mixes more than 8 volume layouts");
            cls.push_back({r.vol->fmt, r.nm, r.vol->lb, buf, {}});
        }
        cls[k].members.push_back(i);
        lut_cap = std::max(lut_cap, (int)(r.lut.size() / 4));
    }
    // LDS capacity in steps of 256 entries, so that graphs are reused
    lut_cap = (int)align_up((size_t)lut_cap, 256);
    // costliest requests first inside each class (tail balance)
    for (Cls &k : cls)
        std::stable_sort(k.members.begin(), k.members.end(), [&](int a, int z) {
            return (long long)b->reqs[a].W * b->reqs[a].H > (long long)b->reqs[z].W * b->reqs[z].H;
        });
    // sizes
    // the block tile of k_batch: 2 waves across, W / 2 down (counting runs
    // use the default variant)
    const int variant = count ? (int)BV_DEFAULT : b->variant;
    const int bx = b->tile_shape == HVR_TILE_16X4 ? 32 : 16;
    const int by = (b->tile_shape == HVR_TILE_16X4 ? 4 : 8) * (batch_variant_waves(variant) / 2);
    size_t desc_off = align_up(sizeof(BatchHeader), 256);
    size_t lut_off = align_up(desc_off + (size_t)nreq * sizeof(ReqDesc), 256);
    size_t lut_bytes = 0, nz_bytes = 0, pixels = 0, flag_bytes = 0;
    b->out_off.assign(nreq, 0);
    for (int i = 0; i < nreq; ++i) {
        const BatchReq &r = b->reqs[i];
        lut_bytes += r.lut.size() * sizeof(float);
        nz_bytes += (r.lut.size() / 4 + 1) * sizeof(int);
        b->out_off[i] = pixels;
        pixels += align_up((size_t)r.W * r.H, 64);
        flag_bytes += align_up((size_t)r.vol->nbx * r.vol->nby * r.vol->nbz, 256);
    }
    const size_t nz_off = align_up(lut_off + lut_bytes, 256);
    const size_t stage_need = align_up(nz_off + nz_bytes, 256);
    bool moved = false;
    if (const int rc = grow((void **)&b->d_stage, (void **)&b->h_stage, &b->stage_cap, stage_need, &moved)) return rc;
    bool moved_out = false;
    if (const int rc = grow((void **)&b->d_out, (void **)&b->h_out, &b->out_cap, pixels * 4, &moved_out)) return rc;
    if (const int rc = grow((void **)&b->d_flags, nullptr, &b->flags_cap, flag_bytes, &moved_out)) return rc;
    if (count) {
        size_t cb = b->counts_cap * 4;
        if (const int rc = grow((void **)&b->d_counts, (void **)&b->h_counts, &cb, pixels * 4, &moved_out)) return rc;
        b->counts_cap = cb / 4;
    }
    if (moved_out) b->h_out_dev_of = nullptr;   // a new pinned pool may reuse the old address
    if (moved) batch_drop_graph(b);
    // direct: the march stores each pixel straight into the pinned pool
    // through its device address (posted writes over the host link while
    // the batch runs, no copy after it)
    const bool direct = b->readback == HVR_READBACK_DIRECT;
    if (direct && b->h_out_dev_of != b->h_out) {
        HIPCHK(hipHostGetDevicePointer((void **)&b->h_out_dev, b->h_out, 0));
        b->h_out_dev_of = b->h_out;
    }
    uchar4 *const out_pool = direct ? b->h_out_dev : b->d_out;
    // stage: header, descriptors (class by class), LUTs, prefix counts
    memset(b->h_stage, 0, desc_off);
    BatchHeader *hdr = (BatchHeader *)b->h_stage;
    hdr->nclass = (int)cls.size();
    hdr->nreq = nreq;
    hdr->all_offset = (uint32_t)desc_off;
    ReqDesc *descs = (ReqDesc *)(b->h_stage + desc_off);
    size_t lut_at = lut_off, nz_at = nz_off, flag_at = 0;
    int slot = 0;
    for (size_t k = 0; k < cls.size(); ++k) {
        hdr->cls[k].desc_offset = (uint32_t)(desc_off + (size_t)slot * sizeof(ReqDesc));
        hdr->cls[k].nreq = (int)cls[k].members.size();
        int tiles = 0;
        for (int i : cls[k].members) {
            const BatchReq &r = b->reqs[i];
            ReqDesc &d = descs[slot++];
            memset(&d, 0, sizeof(d));
            d.fc = frame_const_of(r.params, true, r.W, r.H);
            d.ts = tex_desc(r.vol, HVR_ROLE_SCALAR, d.fc);
            const int n = (int)(r.lut.size() / 4);
            memcpy(b->h_stage + lut_at, r.lut.data(), r.lut.size() * sizeof(float));
            lut_prefix_counts(r.lut.data(), n, (int *)(b->h_stage + nz_at));
            d.lut = (const float4 *)(b->d_stage + lut_at);
            d.nz = (const int *)(b->d_stage + nz_at);
            lut_at += r.lut.size() * sizeof(float);
            nz_at += (size_t)(n + 1) * sizeof(int);
            d.lut_n = n;
            d.range = r.vol->d_range;
            d.nbricks = r.vol->nbx * r.vol->nby * r.vol->nbz;
            d.flags = b->d_flags + flag_at;
            flag_at += align_up(d.nbricks, 256);
            d.out = out_pool + b->out_off[i];
            d.counts = count ? b->d_counts + b->out_off[i] : nullptr;
            d.ntx = (r.W + bx - 1) / bx;
            tiles += d.ntx * ((r.H + by - 1) / by);
            d.tile_end = tiles;
        }
    }
    // the graph for this shape
    std::string key = std::to_string(b->stage_cap) + "/" + std::to_string(lut_cap) + "/" +
                      std::to_string(count) + "/" + std::to_string(b->tile_shape) + "/" + std::to_string(variant);
    for (const Cls &k : cls)
        key += "/" + std::to_string(k.fmt) + "." + std::to_string(k.nm) + "." + std::to_string(k.lb) + "." +
               std::to_string(k.buf);
    if (key != b->key || !b->exec) {
        batch_drop_graph(b);
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = HVR_OK;
        if (hipMemcpyAsync(b->d_stage, b->h_stage, b->stage_cap, hipMemcpyHostToDevice, b->stream) != hipSuccess)
            rc = HVR_EDEVICE;
        if (rc == HVR_OK) {
            // counters reset by a kernel (a graph replays kernels reliably)
            hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, b->stream, b->d_counters, kBatchMaxClasses, 0u);
            hipLaunchKernelGGL(k_batch_flags, dim3(1024), dim3(256), 0, b->stream, (const char *)b->d_stage);
            if (hipGetLastError() != hipSuccess) rc = HVR_EDEVICE;
        }
        for (size_t k = 0; k < cls.size() && rc == HVR_OK; ++k) {
            const BatchLaunchFn fn =
                batch_fn(cls[k].fmt, cls[k].nm, cls[k].lb, count != 0, b->tile_shape, cls[k].buf, variant);
            rc = fn ? fn(b->d_stage, (int)k, b->d_counters + k, lut_cap, b->ncu, b->stream) : HVR_EINVAL;
        }
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc != HVR_OK || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(HVR_EDEVICE, "capturing the batch graph failed");
        }
        b->graph = g;
        HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->key = key;
    }
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    HIPCHK(hipGraphLaunch(b->exec, b->stream));
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    if (!direct) HIPCHK(hipMemcpyAsync(b->h_out, b->d_out, pixels * 4, hipMemcpyDeviceToHost, b->stream));
    if (count) HIPCHK(hipMemcpyAsync(b->h_counts, b->d_counts, pixels * 4, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipEventRecord(b->ev2, b->stream));
    b->out_pixels = pixels;
    b->ran = true;
    b->counted = count != 0;
    return HVR_OK;
}

int hvr_batch_wait(hvr_batch *b)
{
    HVR_FAULT_POINT("hvr_batch_wait");
    if (!b) return fail(HVR_EINVAL, "batch is null");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    return HVR_OK;
}

int hvr_batch_query(hvr_batch *b, int *done)
{
    HVR_FAULT_POINT("hvr_batch_query");
    if (!b || !done) return fail(HVR_EINVAL, "null argument");
    if (!b->ran) {
        *done = 1;
        return HVR_OK;
    }
    HIPCHK(hipSetDevice(b->ctx->device));
    const hipError_t e = hipEventQuery(b->ev2);
    if (e == hipSuccess) *done = 1;
    else if (e == hipErrorNotReady) {
        *done = 0;
        (void)hipGetLastError();   // still in flight, not an error
    }
    else return fail_hip(e, "hvr_batch_query");
    return HVR_OK;
}

int hvr_batch_run(hvr_batch *b, int count)
{
    if (const int rc = hvr_batch_run_async(b, count)) return rc;
    return hvr_batch_wait(b);
}

int hvr_batch_output(hvr_batch *b, int index, const uint8_t **rgba)
{
    if (!b || !rgba) return fail(HVR_EINVAL, "null argument");
    if (!b->ran) return fail(HVR_ESTATE, "the batch has not been run since its last change");
    if (index < 0 || index >= (int)b->reqs.size()) return fail(HVR_EINVAL, "request index out of range");
    *rgba = b->h_out + 4 * b->out_off[index];
    return HVR_OK;
}

int hvr_batch_counts(hvr_batch *b, int index, const uint32_t **counts)
{
    if (!b || !counts) return fail(HVR_EINVAL, "null argument");
    if (!b->ran || !b->counted) return fail(HVR_ESTATE, "the batch has not been run with counting");
    if (index < 0 || index >= (int)b->reqs.size()) return fail(HVR_EINVAL, "request index out of range");
    *counts = b->h_counts + b->out_off[index];
    return HVR_OK;
}

int hvr_batch_last_ms(hvr_batch *b, float *ms)
{
    if (!b || !ms) return fail(HVR_EINVAL, "null argument");
    if (!b->ran) return fail(HVR_ESTATE, "the batch has not been run");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipEventSynchronize(b->ev1));
    HIPCHK(hipEventElapsedTime(ms, b->ev0, b->ev1));
    return HVR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// A frame rendered without allocating (hvr_render_async_to_noalloc): the
// one-request form of the batch launch, whose block (descriptor, tile
// counter) and per-brick transparency flags live in a static device array
// instead of buffers the library allocates.  The block is written on the
// device by k_stage_request from kernel arguments, so the frame is three
// kernels on the context's stream: stage, brick flags, persistent march.

static constexpr size_t kNoallocFlagsOffset = 4096;
static_assert(HVR_NOALLOC_SCRATCH_BYTES == kNoallocFlagsOffset + HVR_NOALLOC_MAX_BRICKS, "scratch layout");

// Layout of the scratch array: the batch header at offset 0 (the batch
// kernels read it at the block's start), the tile counter, the prefix counts
// of the unbound LUT, the request descriptor, then one flag byte per brick.
static constexpr StageOffsets kNoallocOffsets = {128u, 192u, 256u};
static_assert(sizeof(BatchHeader) <= 128, "header fits before the counter");
static_assert(256 + sizeof(ReqDesc) <= kNoallocFlagsOffset, "descriptor fits before the flags");

static __device__ __align__(256) unsigned char scratch[HVR_NOALLOC_SCRATCH_BYTES];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];         // device address of `scratch` per device
static hipError_t g_scratch_err[kMaxDevices];    // the lookup's result per device

// With `device` current: the address of this device's copy of `scratch`
// (each device holds its own), looked up once.  The array's name is a
// host-side handle, not a device address; only this pointer is passed to
// kernels.
static void noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    if (e == hipSuccess) g_scratch_ptr[device] = p;
}

static int noalloc_scratch(int device, char **out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return fail(HVR_EINVAL, "device index out of range for scratch");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        const hipError_t e = g_scratch_err[device];
        g_err = std::string("no scratch address for this device (hipGetSymbolAddress: ") +
                (e == hipSuccess ? "not looked up" : hipGetErrorString(e)) + ")";
        return HVR_EDEVICE;
    }
    *out = (char *)g_scratch_ptr[device];
    return HVR_OK;
}

extern "C" {

int hvr_render_async_to_noalloc(hvr_ctx *c, void *dst)
{
    if (!c || !dst) return fail(HVR_EINVAL, "null argument");
    HIPCHK(hipSetDevice(c->device));
    const hvr_volume *S = c->scalar_vol, *N = c->normal_vol;
    if (!S || c->scalar_role != HVR_ROLE_SCALAR || !fmt_has_scalar(S->fmt) || !S->scalar_loaded || !S->d_range)
        return fail(HVR_ESTATE, "hvr_render_async_to_noalloc: the scalar sampler must read an uploaded scalar field");
    int nm;
    if (!N) nm = NM_ZERO;
    else if (N == S && c->normal_role == HVR_ROLE_SCALAR) nm = NM_ALIAS;
    else if (N == S && c->normal_role == HVR_ROLE_NORMAL && fmt_has_normal(S->fmt)) nm = NM_SAME;
    else return fail(HVR_ESTATE, "hvr_render_async_to_noalloc: normals must come from the scalar volume, or none");
    const bool buf = volume_buf(S) && !c->force_global;
    const BatchLaunchFn fn = batch_fn(S->fmt, nm, S->lb, false, HVR_TILE_8X8, buf);
    if (!fn) return fail(HVR_EINVAL, "hvr_render_async_to_noalloc: brick edge must be 8 or 16 (log2 3 or 4)");
    if (c->lut_n > kLutLdsMax) return fail(HVR_EINVAL, "hvr_render_async_to_noalloc: LUT above 2048 entries");
    const uint64_t nb = (uint64_t)S->nbx * S->nby * S->nbz;
    if (nb > HVR_NOALLOC_MAX_BRICKS)
        return fail(HVR_EINVAL, "hvr_render_async_to_noalloc: volume has more bricks than the scratch array's "
                                "flags hold (HVR_NOALLOC_MAX_BRICKS)");
    char *blk = nullptr;
    if (const int rc = noalloc_scratch(c->device, &blk)) return rc;

    const StageOffsets off = kNoallocOffsets;
    BatchHeader hdr;
    memset(&hdr, 0, sizeof(hdr));
    hdr.nclass = 1;
    hdr.nreq = 1;
    hdr.all_offset = off.desc;
    hdr.cls[0].desc_offset = off.desc;
    hdr.cls[0].nreq = 1;
    // the block tile of k_batch with 8x8 ray tiles and the default variant
    const int bx = 16, by = 8 * (kBatchWaves / 2);
    ReqDesc d;
    memset(&d, 0, sizeof(d));
    d.fc = frame_const(c);
    d.ts = tex_desc(S, HVR_ROLE_SCALAR, d.fc);
    d.lut = c->d_lut;
    d.nz = c->d_nz ? (const int *)c->d_nz : (const int *)(blk + off.unbound_nz);
    d.range = S->d_range;
    d.flags = (uint8_t *)(blk + kNoallocFlagsOffset);
    d.out = (uchar4 *)dst;
    d.counts = nullptr;
    d.nbricks = (uint32_t)nb;
    d.lut_n = c->lut_n;
    d.ntx = (c->W + bx - 1) / bx;
    d.tile_end = d.ntx * ((c->H + by - 1) / by);

    hipEvent_t e0 = c->ev[0], e1 = c->ev[1], e2 = c->ev[2];
    HIPCHK(hipEventRecord(e0, c->stream));
    hipLaunchKernelGGL(k_stage_request, dim3(1), dim3(64), 0, c->stream, blk, off, hdr, d);
    hipLaunchKernelGGL(k_batch_flags, dim3(1024), dim3(256), 0, c->stream, (const char *)blk);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e1, c->stream));
    if (const int rc = fn(blk, 0, (uint32_t *)(blk + off.counter), (int)align_up((size_t)c->lut_n, 256), c->ncu,
                          c->stream))
        return fail(rc, "hvr_render_async_to_noalloc: march launch failed");
    HIPCHK(hipEventRecord(e2, c->stream));
    c->last_ev[0] = e0;
    c->last_ev[1] = e1;
    c->last_ev[2] = e2;
    c->events_recorded = true;
    return HVR_OK;
}

int hvr_volume_download(hvr_volume *v, float *scalar, float *normals)
{
    if (!v || (!scalar && !normals)) return fail(HVR_EINVAL, "null argument");
    if (scalar && (!fmt_has_scalar(v->fmt) || !v->scalar_loaded))
        return fail(HVR_ESTATE, "the volume holds no uploaded scalar field");
    if (normals && (!fmt_has_normal(v->fmt) || !v->normals_loaded))
        return fail(HVR_ESTATE, "the volume holds no uploaded normal field");
    hvr_ctx *c = v->ctx;
    HIPCHK(hipSetDevice(c->device));
    const uint64_t plane = (uint64_t)v->s0 * v->s1;
    // whole z-planes per chunk, up to 256 MiB of floats (scalar + normals)
    const uint64_t per_voxel = (scalar ? 4 : 0) + (normals ? 12 : 0);
    uint64_t zper = ((uint64_t)256 << 20) / (plane * per_voxel);
    zper = std::max<uint64_t>(1, std::min<uint64_t>(zper, (uint64_t)v->s2));
    float *d_s = nullptr, *d_n = nullptr;
    int rc = HVR_OK;
    if (scalar && hipMalloc(&d_s, zper * plane * 4) != hipSuccess) rc = fail(HVR_ENOMEM, "download staging");
    if (rc == HVR_OK && normals && hipMalloc(&d_n, zper * plane * 12) != hipSuccess)
        rc = fail(HVR_ENOMEM, "download staging");
    for (uint64_t z0 = 0; rc == HVR_OK && z0 < (uint64_t)v->s2; z0 += zper) {
        const uint64_t nz = std::min<uint64_t>(zper, (uint64_t)v->s2 - z0);
        const uint64_t nvox = nz * plane;
        const unsigned nblk = (unsigned)((nvox + 255) / 256);
#define HVR_DL(F)                                                                                          \
    case F:                                                                                                \
        hipLaunchKernelGGL(k_download<F>, dim3(nblk), dim3(256), 0, c->stream, (const char *)v->d_texels,  \
                           v->s0, v->s1, v->lb, v->nbx, v->nby, v->brick_bytes, (int)z0, nvox, d_s, d_n); \
        break;
        switch (v->fmt) {
            HVR_DL(TF_S_F32)
            HVR_DL(TF_S_F16)
            HVR_DL(TF_S_U8)
            HVR_DL(TF_N_F32)
            HVR_DL(TF_N_SN8)
            HVR_DL(TF_C_F32_F32)
            HVR_DL(TF_C_F16_F32)
            HVR_DL(TF_C_U8_F32)
            HVR_DL(TF_C_F32_SN8)
            HVR_DL(TF_C_F16_SN8)
            HVR_DL(TF_C_U8_SN8)
            HVR_DL(TF_C_F16_F16)
            HVR_DL(TF_N_F16)
        default: rc = fail(HVR_EINVAL, "bad texel format");
        }
#undef HVR_DL
        if (rc != HVR_OK) break;
        if (hipGetLastError() != hipSuccess) rc = fail(HVR_EDEVICE, "download kernel launch failed");
        if (rc == HVR_OK && scalar &&
            hipMemcpyAsync(scalar + z0 * plane, d_s, nvox * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = fail(HVR_EDEVICE, "download copy failed");
        if (rc == HVR_OK && normals &&
            hipMemcpyAsync(normals + 3 * z0 * plane, d_n, nvox * 12, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = fail(HVR_EDEVICE, "download copy failed");
        if (rc == HVR_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(HVR_EDEVICE, "download wait");
    }
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(d_s);
    (void)hipFree(d_n);
    return rc;
}

}  // extern "C"
