/*
 * Per-pixel ray-march core of the oracle, written once and compiled twice by
 * oracle_vr.c: with REAL=double (canonical FP64 oracle) and REAL=float (the
 * shader's FP32 arithmetic).  TEST INFRASTRUCTURE ONLY: the checker for the
 * HIP renderer, never part of the product.
 *
 * Every step restates the reference fragment shader and the GL state around
 * it (paths relative to the reference repository):
 *   pixel centre            pyvr/shaders/volume.vert.glsl:8, manager.py:63-75
 *   eye/world direction     pyvr/shaders/volume.frag.glsl:47-54
 *   slab test               volume.frag.glsl:56-68, 75-78
 *   march + composite       volume.frag.glsl:80-121
 *   3-D texture sampling    manager.py:95-101, 123-129 (GL_LINEAR,
 *                           CLAMP_TO_EDGE, size (w,h,d) = numpy shape)
 *   transfer-function LUT   manager.py:163-181, volume.frag.glsl:96
 *   blend + RGBA8 store     manager.py:29-31, 212-220, 228-230
 *
 * Required macros: REAL, SUF(name) (suffixes symbol names), RSQRT_FN,
 * EXP_FN, FLOOR_FN.
 */

/* GL_LINEAR sample of a 3-D texture at normalised coordinate (cx, cy, cz).
 * Texel (x, y, z) of channel ch is data[(x + s0*(y + s1*z))*channels + ch]:
 * the texture is the numpy array's bytes with width s0, height s1, depth s2.
 * CLAMP_TO_EDGE: both taps of each axis are clamped to [0, size-1].
 * Channels beyond the texture's own read as 0 (GL fills G, B with 0). */
static void SUF(tex3)(const ovr_tex3 *t, REAL cx, REAL cy, REAL cz, REAL out[3])
{
    const REAL half = (REAL)0.5;
    REAL u = cx * (REAL)t->s0 - half;
    REAL v = cy * (REAL)t->s1 - half;
    REAL w = cz * (REAL)t->s2 - half;
    REAL fu = FLOOR_FN(u), fv = FLOOR_FN(v), fw = FLOOR_FN(w);
    REAL a = u - fu, b = v - fv, g = w - fw;
    long x0 = (long)fu, y0 = (long)fv, z0 = (long)fw;
    long x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
#define OVR_CLAMP(i, n) ((i) < 0 ? 0 : ((i) > (long)(n) - 1 ? (long)(n) - 1 : (i)))
    x0 = OVR_CLAMP(x0, t->s0); x1 = OVR_CLAMP(x1, t->s0);
    y0 = OVR_CLAMP(y0, t->s1); y1 = OVR_CLAMP(y1, t->s1);
    z0 = OVR_CLAMP(z0, t->s2); z1 = OVR_CLAMP(z1, t->s2);
#undef OVR_CLAMP
    const long sx = t->channels, sy = (long)t->s0 * sx, sz = (long)t->s1 * sy;
    const float *p000 = t->data + x0 * sx + y0 * sy + z0 * sz;
    const long dx = (x1 - x0) * sx, dy = (y1 - y0) * sy, dz = (z1 - z0) * sz;
    for (int c = 0; c < 3; ++c) {
        if (c >= t->channels) { out[c] = 0; continue; }
        const float *p = p000 + c;
        /* nested linear interpolation: x, then y, then z */
        REAL c00 = (REAL)p[0] + a * ((REAL)p[dx] - (REAL)p[0]);
        REAL c10 = (REAL)p[dy] + a * ((REAL)p[dy + dx] - (REAL)p[dy]);
        REAL c01 = (REAL)p[dz] + a * ((REAL)p[dz + dx] - (REAL)p[dz]);
        REAL c11 = (REAL)p[dz + dy] + a * ((REAL)p[dz + dy + dx] - (REAL)p[dz + dy]);
        REAL c0 = c00 + b * (c10 - c00);
        REAL c1 = c01 + b * (c11 - c01);
        out[c] = c0 + g * (c1 - c0);
    }
}

/* Linear sample of the N x 1 RGBA LUT at (rho, 0.5): u = rho*N - 0.5,
 * clamp-to-edge; the single row is selected exactly. */
static void SUF(lut1)(const float *lut, int n, REAL rho, REAL out[4])
{
    REAL u = rho * (REAL)n - (REAL)0.5;
    /* keep the int conversion in range (both clamps pick the same edge
     * texel); a NaN density reads the first entry, as max(NaN, 0) = 0 in
     * the IEEE maxNum clamp */
    if (!(u >= (REAL)-1)) u = (REAL)-1;
    if (u > (REAL)n) u = (REAL)n;
    REAL fu = FLOOR_FN(u);
    REAL a = u - fu;
    long i0 = (long)fu, i1 = i0 + 1;
    if (i0 < 0) i0 = 0;
    if (i0 > n - 1) i0 = n - 1;
    if (i1 < 0) i1 = 0;
    if (i1 > n - 1) i1 = n - 1;
    for (int c = 0; c < 4; ++c) {
        REAL q0 = (REAL)lut[4 * i0 + c], q1 = (REAL)lut[4 * i1 + c];
        out[c] = q0 + a * (q1 - q0);
    }
}

/* 4x4 inverse by Gauss-Jordan elimination with partial pivoting, in double.
 * m and inv are row-major in the mathematical sense (m[4*r + c]).
 * Returns 0 when singular. */
static int SUF(inv4)(const double *m, double *inv)
{
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 8; ++c)
            a[r][c] = c < 4 ? m[4 * r + c] : (c - 4 == r ? 1.0 : 0.0);
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int r = col + 1; r < 4; ++r)
            if (fabs(a[r][col]) > fabs(a[piv][col])) piv = r;
        if (a[piv][col] == 0.0) return 0;
        if (piv != col)
            for (int c = 0; c < 8; ++c) { double t = a[col][c]; a[col][c] = a[piv][c]; a[piv][c] = t; }
        double d = a[col][col];
        for (int c = 0; c < 8; ++c) a[col][c] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == col) continue;
            double f = a[r][col];
            if (f != 0.0)
                for (int c = 0; c < 8; ++c) a[r][c] -= f * a[col][c];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[4 * r + c] = a[r][c + 4];
    return 1;
}

/* Shared per-frame constants, derived once from the uniforms. */
typedef struct {
    REAL prow[8];              /* rows 0-1 of inverse(P_gl): eye.xy = prow*(ndc,-1,1) */
    REAL rot[9];               /* 3x3 block of inverse(V_gl): world = rot*(ex,ey,-1) */
    int ok;                    /* 0: a matrix was singular (shader output: NaN rays) */
    REAL L[3];                 /* normalize(light_target - light_position)           */
    int L_ok;                  /* 0: zero-length light vector -> diffuse term is 0   */
} SUF(frame);

static void SUF(frame_setup)(const ovr_params *p, SUF(frame) *f)
{
    /* GLSL receives the numpy row-major bytes as a column-major mat4, i.e.
     * the transpose: M_gl[r][c] = M_np[c][r]. */
    double pg[16], vg[16], pinv[16], vinv[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            pg[4 * r + c] = (double)p->proj[4 * c + r];
            vg[4 * r + c] = (double)p->view[4 * c + r];
        }
    f->ok = SUF(inv4)(pg, pinv) && SUF(inv4)(vg, vinv);
    /* eye = inverse(P_gl) * (ndc.x, ndc.y, -1, 1), then eye.zw := (-1, 0) */
    for (int k = 0; k < 8; ++k) f->prow[k] = (REAL)pinv[k];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) f->rot[3 * r + c] = (REAL)vinv[4 * r + c];
    REAL lx = (REAL)p->light_target[0] - (REAL)p->light_pos[0];
    REAL ly = (REAL)p->light_target[1] - (REAL)p->light_pos[1];
    REAL lz = (REAL)p->light_target[2] - (REAL)p->light_pos[2];
    REAL ll = lx * lx + ly * ly + lz * lz;
    f->L_ok = ll > (REAL)0;
    REAL inv = f->L_ok ? RSQRT_FN(ll) : (REAL)0;
    f->L[0] = lx * inv; f->L[1] = ly * inv; f->L[2] = lz * inv;
}

/* Quantise-then-blend into a cleared RGBA8 target: source (C, A) are clamped
 * and converted to 8-bit, then blended with (SRC_ALPHA, ONE_MINUS_SRC_ALPHA)
 * over 0 in fixed point: rgb = round(Cq*Aq/255), a = round(Aq*Aq/255). */
static inline unsigned SUF(unorm8)(REAL x)
{
    if (!(x > (REAL)0)) return 0;          /* also maps NaN to 0 */
    if (x >= (REAL)1) return 255;
    return (unsigned)FLOOR_FN(x * (REAL)255 + (REAL)0.5);
}

/* March one pixel.  Returns the number of samples composited. */
static unsigned SUF(march_pixel)(const ovr_params *p, const SUF(frame) *f,
                                 const ovr_tex3 *scalar, const ovr_tex3 *normal,
                                 const float *lut, int lut_n,
                                 int px, int py, REAL out[4])
{
    out[0] = out[1] = out[2] = out[3] = 0;
    if (!f->ok) return 0;
    /* A1: pixel centre, bottom row first */
    REAL ux = ((REAL)px + (REAL)0.5) / (REAL)p->width;
    REAL uy = ((REAL)py + (REAL)0.5) / (REAL)p->height;
    REAL nx = ux * (REAL)2 - (REAL)1, ny = uy * (REAL)2 - (REAL)1;
    /* A2: eye-space direction */
    REAL ex = f->prow[0] * nx + f->prow[1] * ny - f->prow[2] + f->prow[3];
    REAL ey = f->prow[4] * nx + f->prow[5] * ny - f->prow[6] + f->prow[7];
    /* A3: world direction, normalised */
    REAL w0 = f->rot[0] * ex + f->rot[1] * ey - f->rot[2];
    REAL w1 = f->rot[3] * ex + f->rot[4] * ey - f->rot[5];
    REAL w2 = f->rot[6] * ex + f->rot[7] * ey - f->rot[8];
    REAL wn = RSQRT_FN(w0 * w0 + w1 * w1 + w2 * w2);
    REAL d[3] = {w0 * wn, w1 * wn, w2 * wn};
    REAL o[3] = {(REAL)p->cam_pos[0], (REAL)p->cam_pos[1], (REAL)p->cam_pos[2]};
    REAL bmin[3] = {(REAL)p->bmin[0], (REAL)p->bmin[1], (REAL)p->bmin[2]};
    REAL bmax[3] = {(REAL)p->bmax[0], (REAL)p->bmax[1], (REAL)p->bmax[2]};
    REAL ext[3] = {bmax[0] - bmin[0], bmax[1] - bmin[1], bmax[2] - bmin[2]};
    /* A4: slab test (IEEE infinities for zero direction components) */
    REAL tn = -(REAL)INFINITY, tf = (REAL)INFINITY;
    for (int k = 0; k < 3; ++k) {
        REAL inv = (REAL)1 / d[k];
        REAL t0 = (bmin[k] - o[k]) * inv, t1 = (bmax[k] - o[k]) * inv;
        REAL lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
        if (lo != lo) lo = -(REAL)INFINITY;   /* 0*inf: the slab does not constrain */
        if (hi != hi) hi = (REAL)INFINITY;
        if (lo > tn) tn = lo;
        if (hi < tf) tf = hi;
    }
    if (!(tn <= tf && tf > (REAL)0)) return 0;
    /* A5: start at the entry point (or the camera when it is inside) */
    if (tn < (REAL)0) tn = 0;
    REAL step = (REAL)p->step;
    REAL ref = (REAL)p->ref_step;
    REAL ka = (REAL)p->ambient, kd = (REAL)p->diffuse;
    REAL tau = (REAL)p->et_threshold;
    REAL pos[3], ds[3];
    for (int k = 0; k < 3; ++k) {
        REAL dt = d[k] * tn;
        pos[k] = o[k] + dt;
        ds[k] = d[k] * step;
    }
    /* Iteration bound for rays that never enter (see oracle_vr.c) */
    long ilim = p->max_steps;
    if (!p->full_loop) {
        double span = ((double)tf - (double)tn) / (double)p->step * 1.0625 + 16.0;
        if (span < (double)ilim) ilim = (long)span;
    }
    REAL C0 = 0, C1 = 0, C2 = 0, A = 0;
    unsigned nsamp = 0;
    int entered = 0;
    for (long i = 0; i < ilim && A < tau; ++i) {
        REAL s[3];
        if (p->direct_positions) {
            /* FP64 ideal: p_i = o + d*(tn + i*step), no accumulated rounding */
            REAL t = tn + (REAL)i * step;
            for (int k = 0; k < 3; ++k) s[k] = o[k] + d[k] * t;
        } else {
            s[0] = pos[0]; s[1] = pos[1]; s[2] = pos[2];
        }
        /* A6: normalised coordinate, swizzled (z, y, x), inclusive [0, 1] test */
        REAL qx = (s[0] - bmin[0]) / ext[0];
        REAL qy = (s[1] - bmin[1]) / ext[1];
        REAL qz = (s[2] - bmin[2]) / ext[2];
        int valid = qz >= 0 && qz <= 1 && qy >= 0 && qy <= 1 && qx >= 0 && qx <= 1;
        if (valid) {
            entered = 1;
            ++nsamp;
            REAL rho = 0;
            if (scalar) { REAL t3[3]; SUF(tex3)(scalar, qz, qy, qx, t3); rho = t3[0]; }
            REAL rgba[4];
            if (lut) SUF(lut1)(lut, lut_n, rho, rgba);
            else { rgba[0] = rgba[1] = rgba[2] = 0; rgba[3] = 1; }   /* incomplete texture */
            REAL a = (REAL)1 - EXP_FN(-(rgba[3] * step) / ref);
            REAL diff = 0;
            if (normal && f->L_ok) {
                REAL nv[3];
                SUF(tex3)(normal, qz, qy, qx, nv);
                REAL nn = nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2];
                /* A10: a zero (or FP32-subnormal) vector has no direction; the
                 * diffuse term is 0 (IEEE maxNum(NaN, 0)). */
                if (nn >= (REAL)OVR_FLT_MIN) {
                    REAL r = RSQRT_FN(nn);
                    REAL dt = (nv[0] * r) * f->L[0] + (nv[1] * r) * f->L[1] + (nv[2] * r) * f->L[2];
                    diff = dt > (REAL)0 ? dt : (REAL)0;
                }
            }
            REAL lam = ka + kd * diff;
            REAL tr = (REAL)1 - A;
            C0 += tr * ((rgba[0] * lam) * a);
            C1 += tr * ((rgba[1] * lam) * a);
            C2 += tr * ((rgba[2] * lam) * a);
            A += tr * a;
        } else if (entered && !p->full_loop) {
            break;   /* left the box: no later sample can be inside */
        }
        pos[0] += ds[0]; pos[1] += ds[1]; pos[2] += ds[2];
    }
    out[0] = C0; out[1] = C1; out[2] = C2; out[3] = A;
    return nsamp;
}

static void SUF(render_rows)(const ovr_params *p, const ovr_tex3 *scalar,
                             const ovr_tex3 *normal, const float *lut, int lut_n,
                             long pix_begin, long pix_end,
                             unsigned char *rgba8, double *rgba_f, unsigned *counts,
                             int nthreads)
{
    SUF(frame) f;
    SUF(frame_setup)(p, &f);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
#endif
    for (long idx = pix_begin; idx < pix_end; ++idx) {
        int px = (int)(idx % p->width), py = (int)(idx / p->width);
        REAL c[4];
        unsigned n = SUF(march_pixel)(p, &f, scalar, normal, lut, lut_n, px, py, c);
        if (rgba_f)
            for (int k = 0; k < 4; ++k) rgba_f[4 * idx + k] = (double)c[k];
        if (counts) counts[idx] = n;
        if (rgba8) {
            unsigned aq = SUF(unorm8)(c[3]);
            for (int k = 0; k < 3; ++k) {
                unsigned cq = SUF(unorm8)(c[k]);
                rgba8[4 * idx + k] = (unsigned char)((cq * aq + 127) / 255);
            }
            rgba8[4 * idx + 3] = (unsigned char)((aq * aq + 127) / 255);
        }
    }
    (void)nthreads;
}
