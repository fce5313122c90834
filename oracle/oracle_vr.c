/*
 * oracle_vr.c -- CPU restatement of the reference's render path (the
 * fragment-shader ray marcher of pyvr/shaders/volume.frag.glsl:70-122 plus
 * the GL state of pyvr/moderngl_renderer/manager.py that surrounds it).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP renderer: only
 * the tests, the smoke check and the CPU-baseline leg of the benchmark load
 * it.  The product never links or calls it.
 *
 * Two builds of the same per-pixel code (oracle_core.h):
 *   ovr_render_f64  canonical oracle: every operation in double, sample
 *                   positions p_i = o + d*(t_near + i*step) evaluated
 *                   directly (no accumulated rounding);
 *   ovr_render_f32  the shader's own arithmetic: float everywhere,
 *                   positions stepped incrementally (p += d*step), as
 *                   volume.frag.glsl:118 does.  This is also the CPU
 *                   baseline of the benchmark (OpenMP over pixels).
 *
 * Exactness of the loop shortcuts (both builds; full_loop=1 disables them):
 *   - Each coordinate of p_i is monotone in i (a fixed step added with
 *     round-to-nearest, or an affine function of i), and so is the
 *     normalised coordinate; the set of in-box samples is therefore one
 *     contiguous range of i.  Once a ray has been inside and a sample is
 *     outside, no later sample is inside: the loop stops there.
 *   - Rays that are never inside stop after 1.0625*(t_far - t_near)/step
 *     + 16 iterations instead of max_steps.  Beyond t_far the exact position
 *     is outside the box; the margin covers the drift of incrementally
 *     stepped FP32 positions (well under 6% of the steps taken unless the
 *     ray runs within ~1e-5 of parallel to the face it leaves through).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp; no fast-math).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OVR_FLT_MIN 1.17549435082228750797e-38  /* 2^-126: smallest normal float */

/* Uniform block of one frame.  view/proj are the float32 matrices exactly as
 * the reference uploads them (numpy row-major bytes of get_view_matrix /
 * get_projection_matrix). */
typedef struct {
    int width, height;
    float view[16];
    float proj[16];
    float cam_pos[3];
    float bmin[3], bmax[3];
    float step, ref_step;
    int max_steps;
    float ambient, diffuse;
    float light_pos[3], light_target[3];
    float et_threshold;        /* 0.99 in the reference shader */
    int full_loop;             /* 1: iterate every i < max_steps like the shader */
    int direct_positions;      /* set by the f64 entry point */
} ovr_params;

/* A bound 3-D texture: float32 data with GL size (s0, s1, s2) = numpy shape
 * and `channels` components per texel (1 = scalar, 3 = normals). */
typedef struct {
    const float *data;
    int s0, s1, s2;
    int channels;
} ovr_tex3;

/* ---- double instantiation ---- */
#define REAL double
#define SUF(n) n##_f64
#define RSQRT_FN(x) (1.0 / sqrt(x))
#define EXP_FN(x) exp(x)
#define FLOOR_FN(x) floor(x)
#include "oracle_core.h"
#undef REAL
#undef SUF
#undef RSQRT_FN
#undef EXP_FN
#undef FLOOR_FN

/* ---- float instantiation ---- */
#define REAL float
#define SUF(n) n##_f32
#define RSQRT_FN(x) (1.0f / sqrtf(x))
#define EXP_FN(x) expf(x)
#define FLOOR_FN(x) floorf(x)
#include "oracle_core.h"
#undef REAL
#undef SUF
#undef RSQRT_FN
#undef EXP_FN
#undef FLOOR_FN

/* Render pixels [pix_begin, pix_end) (linear index py*width + px, py = 0 is
 * the bottom row).  scalar / normal may be NULL (an unbound sampler reads 0);
 * lut may be NULL (reads (0,0,0,1)).  Outputs (each may be NULL) are indexed
 * by the linear pixel index over the full frame:
 *   rgba8  W*H*4 bytes as stored by the blend (bottom row first),
 *   rgba_f W*H*4 doubles: the shader's output (C.rgb, A) before the blend,
 *   counts W*H samples composited per pixel. */
int ovr_render_f64(const ovr_params *p, const ovr_tex3 *scalar, const ovr_tex3 *normal,
                   const float *lut, int lut_n, long pix_begin, long pix_end,
                   uint8_t *rgba8, double *rgba_f, uint32_t *counts, int nthreads)
{
    if (!p || p->width <= 0 || p->height <= 0 || (lut && lut_n <= 0)) return -1;
    ovr_params q = *p;
    q.direct_positions = 1;
    render_rows_f64(&q, scalar, normal, lut, lut_n, pix_begin, pix_end, rgba8, rgba_f,
                    counts, nthreads);
    return 0;
}

int ovr_render_f32(const ovr_params *p, const ovr_tex3 *scalar, const ovr_tex3 *normal,
                   const float *lut, int lut_n, long pix_begin, long pix_end,
                   uint8_t *rgba8, double *rgba_f, uint32_t *counts, int nthreads)
{
    if (!p || p->width <= 0 || p->height <= 0 || (lut && lut_n <= 0)) return -1;
    ovr_params q = *p;
    q.direct_positions = 0;
    render_rows_f32(&q, scalar, normal, lut, lut_n, pix_begin, pix_end, rgba8, rgba_f,
                    counts, nthreads);
    return 0;
}

/* Single-sample helpers exported for the tests (closed-form pins of the GL
 * sampling rules). */
void ovr_tex3_f64(const ovr_tex3 *t, double cx, double cy, double cz, double out[3])
{
    tex3_f64(t, cx, cy, cz, out);
}

void ovr_lut_f64(const float *lut, int n, double rho, double out[4])
{
    lut1_f64(lut, n, rho, out);
}

int ovr_inv4(const double *m, double *inv)
{
    return inv4_f64(m, inv);
}

unsigned ovr_blend8(double c, double a)
{
    unsigned aq = unorm8_f64(a), cq = unorm8_f64(c);
    return ((cq * aq + 127) / 255) | (((aq * aq + 127) / 255) << 8);
}

int ovr_abi_version(void) { return 1; }
