/* Test infrastructure: a driver that runs the C oracle (oracle_vr.c) over
 * random small scenes and edge cases, built with AddressSanitizer and
 * UndefinedBehaviorSanitizer (`make -C oracle sanitize`; run by
 * tests/test_sanitizers.py).  It checks memory safety and undefined behaviour
 * of the oracle's code paths, not pixel values (the pins in
 * tests/test_oracle_pins.py do that): every output must be written inside
 * its buffer, counts stay within max_steps, and outputs are finite where the
 * inputs are.  Exit status 0 on success. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

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
    float et_threshold;
    int full_loop;
    int direct_positions;
} ovr_params;

typedef struct {
    const float *data;
    int s0, s1, s2;
    int channels;
} ovr_tex3;

int ovr_render_f64(const ovr_params *p, const ovr_tex3 *scalar, const ovr_tex3 *normal, const float *lut,
                   int lut_n, long pix_begin, long pix_end, uint8_t *rgba8, double *rgba_f, uint32_t *counts,
                   int nthreads);
int ovr_render_f32(const ovr_params *p, const ovr_tex3 *scalar, const ovr_tex3 *normal, const float *lut,
                   int lut_n, long pix_begin, long pix_end, uint8_t *rgba8, double *rgba_f, uint32_t *counts,
                   int nthreads);
void ovr_tex3_f64(const ovr_tex3 *t, double cx, double cy, double cz, double out[3]);
void ovr_lut_f64(const float *lut, int n, double rho, double out[4]);
int ovr_inv4(const double *m, double *inv);
unsigned ovr_blend8(double c, double a);

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static double urand(void)
{
    rng_state = rng_state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(rng_state >> 11) * 0x1p-53;
}
static int irand(int n) { return (int)(urand() * n); }

/* lookAt (pre-transposed as the reference authors it) and perspective, in the
 * float32 row-major layout the renderer uploads */
static void camera(ovr_params *p, const double pos[3], double fov, double aspect)
{
    double f[3] = {-pos[0], -pos[1], -pos[2]}, up[3] = {0, 0, 1}, r[3], u[3];
    double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (int k = 0; k < 3; ++k) f[k] /= fl;
    if (fabs(f[2]) > 0.999) { up[0] = 0; up[1] = 1; up[2] = 0; }
    r[0] = f[1] * up[2] - f[2] * up[1];
    r[1] = f[2] * up[0] - f[0] * up[2];
    r[2] = f[0] * up[1] - f[1] * up[0];
    double rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (int k = 0; k < 3; ++k) r[k] /= rl;
    u[0] = r[1] * f[2] - r[2] * f[1];
    u[1] = r[2] * f[0] - r[0] * f[2];
    u[2] = r[0] * f[1] - r[1] * f[0];
    double v[16] = {0};
    for (int k = 0; k < 3; ++k) {
        v[4 * k + 0] = r[k];
        v[4 * k + 1] = u[k];
        v[4 * k + 2] = -f[k];
    }
    v[12] = -(r[0] * pos[0] + r[1] * pos[1] + r[2] * pos[2]);
    v[13] = -(u[0] * pos[0] + u[1] * pos[1] + u[2] * pos[2]);
    v[14] = f[0] * pos[0] + f[1] * pos[1] + f[2] * pos[2];
    v[15] = 1;
    double t = tan(fov / 2), n = 0.1, fa = 100.0, pr[16] = {0};
    pr[0] = 1 / (aspect * t);
    pr[5] = 1 / t;
    pr[10] = -(fa + n) / (fa - n);
    pr[11] = -1;
    pr[14] = -2 * fa * n / (fa - n);
    for (int k = 0; k < 16; ++k) {
        p->view[k] = (float)v[k];
        p->proj[k] = (float)pr[k];
    }
    for (int k = 0; k < 3; ++k) p->cam_pos[k] = (float)pos[k];
}

static int check_frame(const ovr_params *p, long b, long e, const uint8_t *rgba8, const double *rgba_f,
                       const uint32_t *counts, int finite_inputs)
{
    for (long i = b; i < e; ++i) {
        if ((long)counts[i] > (long)p->max_steps) {
            fprintf(stderr, "pixel %ld: %u samples > max_steps %d\n", i, counts[i], p->max_steps);
            return 1;
        }
        if (finite_inputs)
            for (int c = 0; c < 4; ++c)
                if (!isfinite(rgba_f[4 * i + c])) {
                    fprintf(stderr, "pixel %ld: non-finite output\n", i);
                    return 1;
                }
    }
    (void)rgba8;
    return 0;
}

static int run_scene(int iter)
{
    ovr_params p;
    memset(&p, 0, sizeof(p));
    p.width = 1 + irand(12);
    p.height = 1 + irand(12);
    const int s0 = 1 + irand(7), s1 = 1 + irand(7), s2 = 1 + irand(7);
    const long nv = (long)s0 * s1 * s2;
    float *sc = malloc(sizeof(float) * nv), *nm = malloc(sizeof(float) * nv * 3);
    int plateau = irand(3) == 0, finite = 1;
    for (long i = 0; i < nv; ++i) {
        sc[i] = plateau ? (float)(irand(2)) : (float)urand();
        for (int c = 0; c < 3; ++c) nm[3 * i + c] = plateau ? 0.f : (float)(2 * urand() - 1);
    }
    const int lut_n = 1 + irand(64);
    float *lut = malloc(sizeof(float) * 4 * lut_n);
    for (int i = 0; i < 4 * lut_n; ++i) lut[i] = (float)urand();
    if (irand(4) == 0) {             /* non-finite LUT entries */
        lut[4 * irand(lut_n) + irand(4)] = irand(2) ? NAN : INFINITY;
        finite = 0;
    }
    for (int k = 0; k < 3; ++k) {
        p.bmin[k] = (float)(-0.5 - urand());
        p.bmax[k] = (float)(0.5 + urand());
    }
    double pos[3];
    const double d = irand(4) == 0 ? 0.3 : 1.5 + 3 * urand();   /* sometimes inside the box */
    const double az = 2 * M_PI * urand(), el = M_PI * (urand() - 0.5);
    pos[0] = d * cos(el) * cos(az);
    pos[1] = d * cos(el) * sin(az);
    pos[2] = d * sin(el);
    camera(&p, pos, M_PI / 4, (double)p.width / p.height);
    p.step = (float)(0.005 + 0.05 * urand());
    p.ref_step = 0.01f;
    p.max_steps = irand(300);
    p.ambient = 0.2f;
    p.diffuse = 0.8f;
    p.light_pos[0] = 10.f;
    p.light_pos[1] = -10.f;
    p.light_target[0] = p.light_target[1] = p.light_target[2] = 0.f;
    if (irand(8) == 0) p.light_target[0] = p.light_pos[0], p.light_target[1] = p.light_pos[1]; /* zero L */
    p.et_threshold = (float)(0.5 + urand());
    p.full_loop = irand(2);

    ovr_tex3 ts = {sc, s0, s1, s2, 1}, tn = {nm, s0, s1, s2, 3};
    const long npix = (long)p.width * p.height;
    uint8_t *rgba8 = malloc(4 * npix);
    double *rgba_f = malloc(sizeof(double) * 4 * npix);
    uint32_t *counts = malloc(sizeof(uint32_t) * npix);
    int rc = 0;
    for (int mode = 0; mode < 2 && !rc; ++mode) {
        /* a random pixel range, then the whole frame */
        long b = irand((int)npix), e = b + irand((int)(npix - b) + 1);
        for (int pass = 0; pass < 2 && !rc; ++pass) {
            if (pass == 1) { b = 0; e = npix; }
            const ovr_tex3 *S = irand(6) == 0 ? NULL : &ts;
            const ovr_tex3 *N = irand(6) == 0 ? NULL : (irand(3) == 0 ? &ts : &tn);
            const float *L = irand(8) == 0 ? NULL : lut;
            memset(counts, 0, sizeof(uint32_t) * npix);
            int r = mode ? ovr_render_f32(&p, S, N, L, L ? lut_n : 0, b, e, rgba8, rgba_f, counts, 1 + irand(3))
                         : ovr_render_f64(&p, S, N, L, L ? lut_n : 0, b, e, rgba8, rgba_f, counts, 1 + irand(3));
            if (r != 0) {
                fprintf(stderr, "scene %d: render failed\n", iter);
                rc = 1;
            } else {
                rc = check_frame(&p, b, e, rgba8, rgba_f, counts, finite && L != NULL);
            }
        }
    }
    /* single-sample helpers, including coordinates outside [0, 1] */
    double out[4];
    for (int k = 0; k < 4; ++k) {
        ovr_tex3_f64(&ts, 2 * urand() - 0.5, 2 * urand() - 0.5, 2 * urand() - 0.5, out);
        ovr_tex3_f64(&tn, 2 * urand() - 0.5, 2 * urand() - 0.5, 2 * urand() - 0.5, out);
        ovr_lut_f64(lut, lut_n, 3 * urand() - 1, out);
    }
    ovr_lut_f64(lut, lut_n, NAN, out);
    ovr_lut_f64(lut, lut_n, INFINITY, out);
    (void)ovr_blend8(3 * urand() - 1, 3 * urand() - 1);
    (void)ovr_blend8(NAN, NAN);
    free(sc);
    free(nm);
    free(lut);
    free(rgba8);
    free(rgba_f);
    free(counts);
    return rc;
}

int main(int argc, char **argv)
{
    const int n = argc > 1 ? atoi(argv[1]) : 400;
    double m[16], inv[16];
    for (int k = 0; k < 16; ++k) m[k] = urand();
    (void)ovr_inv4(m, inv);
    memset(m, 0, sizeof(m));
    if (ovr_inv4(m, inv)) {
        fprintf(stderr, "singular matrix inverted\n");
        return 1;
    }
    for (int i = 0; i < n; ++i)
        if (run_scene(i)) return 1;
    printf("sanitized oracle: %d scenes ok\n", n);
    return 0;
}
