// hvr_launch.h -- host-side launchers of the march kernels, one set per
// texel format.  Each format's kernel instances live in their own
// translation unit (inst.hip, compiled once per format with HVR_INST_FMT set
// to the format's TexFmt name), so that the library builds in parallel;
// hipvr.hip dispatches to them by format.
#pragma once

#include "hvr_device.h"

namespace hvr {

// Everything a k_march launch takes besides its template parameters.
struct MarchArgs {
    FrameConst fc;
    TexDesc ts, tn;
    const float4 *lut;
    int lut_n;
    uchar4 *out;
    uint32_t *counts;
    const uint8_t *skip;
    int ntx, nblocks;
    const uint32_t *order_cnt;     // cost-ordered dispatch (k_order_blocks), or nullptr
    const uint32_t *order_slots;
    int order_cap;
    hipStream_t stream;
};

// Kernel variants of the batched march (the autotuner's space beyond brick
// size and ray tile): waves per block, transparent-brick skipping, normals
// gated on non-zero alpha.  Every variant gives the same image.
enum BatchVariant : int {
    BV_DEFAULT = 0,   // 4 waves per block, skipping, gating
    BV_WAVES2 = 1,    // 2 waves per block
    BV_WAVES8 = 2,    // 8 waves per block
    BV_NOGATE = 3,    // normals for every shaded sample
    BV_NOSKIP = 4,    // no transparent-brick skipping
    BV_COUNT = 5
};
constexpr int batch_variant_waves(int v) { return v == BV_WAVES2 ? 2 : v == BV_WAVES8 ? 8 : kBatchWaves; }

// A persistent batched launch of class `cls` of the staged block.
typedef int (*BatchLaunchFn)(const char *blk, int cls, uint32_t *ctr, int lut_cap, int ncu, hipStream_t st);

// The formats with fast-path (templated) kernels.
#define HVR_FAST_FORMATS(X) \
    X(TF_S_F32)             \
    X(TF_S_F16)             \
    X(TF_S_U8)              \
    X(TF_C_F32_F32)         \
    X(TF_C_F16_F32)         \
    X(TF_C_U8_F32)          \
    X(TF_C_F32_SN8)         \
    X(TF_C_F16_SN8)         \
    X(TF_C_U8_SN8)          \
    X(TF_C_F16_F16)

// Per format F:
//   launch_march_F: k_march for normal mode nm (NM_ZERO / NM_ALIAS /
//     NM_SAME), brick log2 lb (3 or 4), counting, skipping, buffer loads;
//     returns false when this format has no such instance (NM_SAME without
//     normals);
//   batch_fn_F: the batched launcher for those choices, tile shape ts and
//     kernel variant (BV_*), or nullptr when there is none (counting runs
//     use BV_DEFAULT only).
#define HVR_DECLARE_FORMAT(F)                                                                   \
    bool launch_march_##F(const MarchArgs &a, int nm, int lb, bool count, bool skip, bool buf); \
    BatchLaunchFn batch_fn_##F(int nm, int lb, bool count, int ts, bool buf, int variant);
HVR_FAST_FORMATS(HVR_DECLARE_FORMAT)
#undef HVR_DECLARE_FORMAT

}  // namespace hvr
