// inst.hip -- the k_march and k_batch instances of one texel format and their
// launchers (hvr_launch.h).  Compiled once per format with
// -DHVR_INST_FMT=<TexFmt name> (see pyvr_amd/_build.py), so the library's
// kernel instances build in parallel.
#include "hvr_launch.h"
#include "../../include/hipvr.h"

#ifndef HVR_INST_FMT
#error "compile with -DHVR_INST_FMT=<TexFmt name>"
#endif

#define HVR_CAT2(a, b) a##b
#define HVR_CAT(a, b) HVR_CAT2(a, b)

namespace hvr {
namespace {

constexpr int F = HVR_INST_FMT;

// ---- single-frame march (K2) ----
template <int NM, int LB, bool COUNT, bool SKIP, bool BUF>
void march(const MarchArgs &a)
{
    const size_t lds = 2 * (size_t)a.lut_n * sizeof(float4);
    hipLaunchKernelGGL((k_march<F, NM, LB, COUNT, SKIP, BUF>), dim3(a.nblocks), dim3(256), lds, a.stream, a.fc, a.ts,
                       a.tn, a.lut, a.lut_n, a.out, a.counts, a.skip, a.ntx, a.nblocks,
                       a.order_cnt, a.order_slots, a.order_cap);
}

template <int NM, int LB, bool COUNT>
void march_sb(const MarchArgs &a, bool skip, bool buf)
{
    if (skip && buf) march<NM, LB, COUNT, true, true>(a);
    else if (skip) march<NM, LB, COUNT, true, false>(a);
    else if (buf) march<NM, LB, COUNT, false, true>(a);
    else march<NM, LB, COUNT, false, false>(a);
}

template <int NM>
void march_nm(const MarchArgs &a, int lb, bool count, bool skip, bool buf)
{
    if (lb == 4) {
        if (count) march_sb<NM, 4, true>(a, skip, buf);
        else march_sb<NM, 4, false>(a, skip, buf);
    } else {
        if (count) march_sb<NM, 3, true>(a, skip, buf);
        else march_sb<NM, 3, false>(a, skip, buf);
    }
}

// ---- batched march ----
template <int NM, int LB, bool COUNT, int TS, bool BUF, int W, bool SKIP, bool GATE>
int batch_launch(const char *blk, int cls, uint32_t *ctr, int lut_cap, int ncu, hipStream_t st)
{
    const size_t lds = 2 * (size_t)lut_cap * sizeof(float4);
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch<F, NM, LB, COUNT, TS, BUF, W, SKIP, GATE>, 64 * W,
                                                     lds) != hipSuccess ||
        per < 1)
        per = 1;
    hipLaunchKernelGGL((k_batch<F, NM, LB, COUNT, TS, BUF, W, SKIP, GATE>), dim3(per * ncu), dim3(64 * W), lds, st,
                       blk, cls, ctr);
    return hipGetLastError() == hipSuccess ? HVR_OK : HVR_EDEVICE;
}

template <int NM, int LB, bool COUNT, int TS, bool BUF>
BatchLaunchFn batch_variant(int v)
{
    if constexpr (COUNT) {
        return v == BV_DEFAULT ? batch_launch<NM, LB, true, TS, BUF, kBatchWaves, true, true> : nullptr;
    } else {
        switch (v) {
        case BV_DEFAULT: return batch_launch<NM, LB, false, TS, BUF, kBatchWaves, true, true>;
        case BV_WAVES2: return batch_launch<NM, LB, false, TS, BUF, 2, true, true>;
        case BV_WAVES8: return batch_launch<NM, LB, false, TS, BUF, 8, true, true>;
        case BV_NOGATE: return batch_launch<NM, LB, false, TS, BUF, kBatchWaves, true, false>;
        case BV_NOSKIP: return batch_launch<NM, LB, false, TS, BUF, kBatchWaves, false, true>;
        default: return nullptr;
        }
    }
}

template <int NM, bool COUNT, bool BUF>
BatchLaunchFn batch_lb_ts(int lb, int ts, int v)
{
    if (lb == 4) return ts == HVR_TILE_16X4 ? batch_variant<NM, 4, COUNT, 1, BUF>(v) : batch_variant<NM, 4, COUNT, 0, BUF>(v);
    return ts == HVR_TILE_16X4 ? batch_variant<NM, 3, COUNT, 1, BUF>(v) : batch_variant<NM, 3, COUNT, 0, BUF>(v);
}

template <int NM>
BatchLaunchFn batch_nm(int lb, bool count, int ts, bool buf, int v)
{
    if (count) return buf ? batch_lb_ts<NM, true, true>(lb, ts, v) : batch_lb_ts<NM, true, false>(lb, ts, v);
    return buf ? batch_lb_ts<NM, false, true>(lb, ts, v) : batch_lb_ts<NM, false, false>(lb, ts, v);
}

}  // namespace

bool HVR_CAT(launch_march_, HVR_INST_FMT)(const MarchArgs &a, int nm, int lb, bool count, bool skip, bool buf)
{
    if (nm == NM_ZERO) {
        march_nm<NM_ZERO>(a, lb, count, skip, buf);
        return true;
    }
    if (nm == NM_ALIAS) {
        march_nm<NM_ALIAS>(a, lb, count, skip, buf);
        return true;
    }
    if constexpr (fmt_has_normal(F)) {
        if (nm == NM_SAME) {
            march_nm<NM_SAME>(a, lb, count, skip, buf);
            return true;
        }
    }
    return false;
}

// Batched launchers: scalar-only formats read normals as (rho, 0, 0) or none,
// combined formats from their own texels.
BatchLaunchFn HVR_CAT(batch_fn_, HVR_INST_FMT)(int nm, int lb, bool count, int ts, bool buf, int variant)
{
    if (lb != 3 && lb != 4) return nullptr;
    if constexpr (fmt_has_normal(F)) {
        return nm == NM_SAME ? batch_nm<NM_SAME>(lb, count, ts, buf, variant) : nullptr;
    } else {
        if (nm == NM_ZERO) return batch_nm<NM_ZERO>(lb, count, ts, buf, variant);
        if (nm == NM_ALIAS) return batch_nm<NM_ALIAS>(lb, count, ts, buf, variant);
        return nullptr;
    }
}

}  // namespace hvr
