/*
 * The oracle's path arithmetic: lib/ref of cyanreg/libcraw restated, one
 * source for three modes (see oracle.h). TEST INFRASTRUCTURE.
 *
 * Build with ORACLE_MODE 1 (float, glibc), 2 (float, detmath.h) or 3
 * (double), always with -ffp-contract=off: every expression below is
 * evaluated as written, left to right, the way lib/ref's C is, so that the
 * float modes round exactly where the reference rounds. Constants are
 * given as a float literal and a double literal (K(f, d)) so that the
 * float modes use the reference's own float constants.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "detmath.h"

#if ORACLE_MODE == 1
typedef float real;
#define PFX(n) oracle_f32l_##n
#define EXPF(x) expf(x)
#define LOGF(x) logf(x)
#define POWF(x, y) powf(x, y)
#define SQRTF(x) sqrtf(x)
#define RINTF(x) rintf(x)
#define FLOORF(x) floorf(x)
#define ROUNDF(x) roundf(x)
#define FABSF(x) fabsf(x)
#define K(f, d) (f)
#elif ORACLE_MODE == 2
typedef float real;
#define PFX(n) oracle_f32d_##n
#define EXPF(x) craw_dm_expf(x)
#define LOGF(x) craw_dm_logf(x)
#define POWF(x, y) craw_dm_powf(x, y)
#define SQRTF(x) sqrtf(x)
#define RINTF(x) rintf(x)
#define FLOORF(x) floorf(x)
#define ROUNDF(x) roundf(x)
#define FABSF(x) fabsf(x)
#define K(f, d) (f)
#elif ORACLE_MODE == 3
typedef double real;
#define PFX(n) oracle_f64_##n
#define EXPF(x) exp(x)
#define LOGF(x) log(x)
#define POWF(x, y) pow(x, y)
#define SQRTF(x) sqrt(x)
#define RINTF(x) rint(x)
#define FLOORF(x) floor(x)
#define ROUNDF(x) round(x)
#define FABSF(x) fabs(x)
#define K(f, d) (d)
#else
#error "ORACLE_MODE must be 1, 2 or 3"
#endif

/* The reference's ternary min/max/clamp (utils.h:41-43), NaN behaviour
 * included. */
#define OMIN(a, b) ((a) < (b) ? (a) : (b))
#define OMAX(a, b) ((a) > (b) ? (a) : (b))
#define OCLAMP(v, lo, hi) OMIN(OMAX((v), (lo)), (hi))

static int o_threads = 1;

void PFX(set_threads)(int n)
{
    o_threads = n > 0 ? n : 1;
}

#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static) num_threads(o_threads)")
#else
#define OMP_FOR
#endif

/* RGGB site to channel (ref_debayer.c:14). */
static const int o_site_chan[4] = { 0, 1, 1, 2 };

/* ref_debayer.c:20-33: mirror without repeating the edge sample. */
static inline int o_mirror(int v, int n)
{
    if (n < 2)
        return 0;
    while (v < 0 || v >= n) {
        if (v < 0)
            v = -v;
        if (v >= n)
            v = 2 * n - 2 - v;
    }
    return v;
}

static inline real o_fetch(const real *p, uint32_t w, uint32_t h, int x, int y)
{
    return p[(size_t)o_mirror(y, (int)h) * w + o_mirror(x, (int)w)];
}

/* lib/ref/ref_debayer.c:42-63: black, range, gain, clip at 1.0;
 * below black is kept. */
void PFX(prep)(const uint16_t *bayer, uint32_t width, uint32_t height, size_t stride,
               const CRAWProfile *p, const float wb[3], real *v)
{
    OMP_FOR
    for (uint32_t y = 0; y < height; y++) {
        for (uint32_t x = 0; x < width; x++) {
            int s = (int)((y & 1) * 2 + (x & 1));
            real raw = (real)bayer[(size_t)y * stride + x] * K(1.0f / 65535.0f, 1.0 / 65535.0);
            real t = (raw - (real)p->black_level[s]) /
                     ((real)p->white_level[s] - (real)p->black_level[s]);

            t *= (real)wb[o_site_chan[s]];
            v[(size_t)y * width + x] = OMIN(t, K(1.0f, 1.0));
        }
    }
}

#define V(dx, dy) o_fetch(v, width, height, x + (dx), y + (dy))

/* lib/ref/ref_debayer.c:65-96. */
static void o_bilinear(const real *v, uint32_t width, uint32_t height, real *rgb)
{
    OMP_FOR
    for (int y = 0; y < (int)height; y++) {
        for (int x = 0; x < (int)width; x++) {
            real *o = &rgb[((size_t)y * width + x) * 4];
            int s = (y & 1) * 2 + (x & 1);
            real edge = K(0.25f, 0.25) * (V(0, -1) + V(0, 1) + V(-1, 0) + V(1, 0));
            real corner = K(0.25f, 0.25) * (V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1));
            real horiz = K(0.5f, 0.5) * (V(-1, 0) + V(1, 0));
            real vert = K(0.5f, 0.5) * (V(0, -1) + V(0, 1));

            switch (s) {
            case 0:  o[0] = V(0, 0); o[1] = edge;    o[2] = corner;  break;
            case 1:  o[0] = horiz;   o[1] = V(0, 0); o[2] = vert;    break;
            case 2:  o[0] = vert;    o[1] = V(0, 0); o[2] = horiz;   break;
            default: o[0] = corner;  o[1] = edge;    o[2] = V(0, 0); break;
            }
            o[3] = K(1.0f, 1.0);
        }
    }
}

/* lib/ref/ref_debayer.c:101-145: the four MHC kernels over 8. */
static inline real o_mhc_green(const real *v, uint32_t width, uint32_t height, int x, int y)
{
    return (K(4.0f, 4.0) * V(0, 0)
            + K(2.0f, 2.0) * (V(0, -1) + V(0, 1) + V(-1, 0) + V(1, 0))
            - (V(0, -2) + V(0, 2) + V(-2, 0) + V(2, 0))) * K(0.125f, 0.125);
}

static inline real o_mhc_cross_h(const real *v, uint32_t width, uint32_t height, int x, int y)
{
    return (K(5.0f, 5.0) * V(0, 0)
            + K(4.0f, 4.0) * (V(-1, 0) + V(1, 0))
            - (V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1) + V(-2, 0) + V(2, 0))
            + K(0.5f, 0.5) * (V(0, -2) + V(0, 2))) * K(0.125f, 0.125);
}

static inline real o_mhc_cross_v(const real *v, uint32_t width, uint32_t height, int x, int y)
{
    return (K(5.0f, 5.0) * V(0, 0)
            + K(4.0f, 4.0) * (V(0, -1) + V(0, 1))
            - (V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1) + V(0, -2) + V(0, 2))
            + K(0.5f, 0.5) * (V(-2, 0) + V(2, 0))) * K(0.125f, 0.125);
}

static inline real o_mhc_diag(const real *v, uint32_t width, uint32_t height, int x, int y)
{
    return (K(6.0f, 6.0) * V(0, 0)
            + K(2.0f, 2.0) * (V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1))
            - K(1.5f, 1.5) * (V(0, -2) + V(0, 2) + V(-2, 0) + V(2, 0))) * K(0.125f, 0.125);
}

/* ref_debayer.c:147-176 */
static void o_mhc(const real *v, uint32_t width, uint32_t height, real *rgb)
{
    OMP_FOR
    for (int y = 0; y < (int)height; y++) {
        for (int x = 0; x < (int)width; x++) {
            real *o = &rgb[((size_t)y * width + x) * 4];
            int s = (y & 1) * 2 + (x & 1);
            real c = V(0, 0);

            switch (s) {
            case 0:
                o[0] = c;
                o[1] = o_mhc_green(v, width, height, x, y);
                o[2] = o_mhc_diag(v, width, height, x, y);
                break;
            case 1:
                o[0] = o_mhc_cross_h(v, width, height, x, y);
                o[1] = c;
                o[2] = o_mhc_cross_v(v, width, height, x, y);
                break;
            case 2:
                o[0] = o_mhc_cross_v(v, width, height, x, y);
                o[1] = c;
                o[2] = o_mhc_cross_h(v, width, height, x, y);
                break;
            default:
                o[0] = o_mhc_diag(v, width, height, x, y);
                o[1] = o_mhc_green(v, width, height, x, y);
                o[2] = c;
                break;
            }
            o[3] = K(1.0f, 1.0);
        }
    }
}

#undef V

/* ref_debayer.c:179-190: 3x3 mean, mirrored, summed row by row. */
static inline real o_box3(const real *s, uint32_t width, uint32_t height, int x, int y)
{
    real acc = K(0.0f, 0.0);

    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++)
            acc += o_fetch(s, width, height, x + dx, y + dy);
    return acc * K(1.0f / 9.0f, 1.0 / 9.0);
}

#define O_RCD_EPS K(1e-5f, 1e-5)

/* ref_debayer.c:195-202: signed low-pass ratio, below black taken as 0. */
static inline real o_rcd_ratio(real a, real b)
{
    a = OMAX(a, K(0.0f, 0.0));
    b = OMAX(b, K(0.0f, 0.0));
    return (a - b) / (a + b + O_RCD_EPS);
}

/* lib/ref/ref_debayer.c:204-395: RCD in full-frame sweeps. */
static int o_rcd(const real *v, uint32_t width, uint32_t height, real *rgb)
{
    size_t n = (size_t)width * height;
    real *lp = malloc(n * sizeof(real)), *wv = malloc(n * sizeof(real));
    real *wp = malloc(n * sizeof(real)), *gp = malloc(n * sizeof(real));
    real *cp = malloc(n * sizeof(real)), *ga = malloc(n * sizeof(real));
    real *gb = malloc(n * sizeof(real));
    int ret = 0;

    if (!lp || !wv || !wp || !gp || !cp || !ga || !gb) {
        ret = -ENOMEM;
        goto fail;
    }

#define V(dx, dy) o_fetch(v, width, height, x + (dx), y + (dy))
#define L(dx, dy) o_fetch(lp, width, height, x + (dx), y + (dy))
#define G(dx, dy) o_fetch(gp, width, height, x + (dx), y + (dy))
#define C(dx, dy) o_fetch(cp, width, height, x + (dx), y + (dy))

    /* Pass A: binomial low pass. */
    OMP_FOR
    for (int y = 0; y < (int)height; y++)
        for (int x = 0; x < (int)width; x++)
            lp[(size_t)y * width + x] =
                (K(4.0f, 4.0) * V(0, 0)
                 + K(2.0f, 2.0) * (V(0, -1) + V(0, 1) + V(-1, 0) + V(1, 0))
                 + V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1)) * K(1.0f / 16.0f, 1.0 / 16.0);

    /* Pass A: vertical / horizontal gradients, then their 3x3 means. */
    OMP_FOR
    for (int y = 0; y < (int)height; y++) {
        for (int x = 0; x < (int)width; x++) {
            size_t i = (size_t)y * width + x;
            ga[i] = FABSF(V(0, -1) - V(0, 1))
                  + FABSF(K(2.0f, 2.0) * V(0, 0) - V(0, -2) - V(0, 2));
            gb[i] = FABSF(V(-1, 0) - V(1, 0))
                  + FABSF(K(2.0f, 2.0) * V(0, 0) - V(-2, 0) - V(2, 0));
        }
    }
    OMP_FOR
    for (int y = 0; y < (int)height; y++) {
        for (int x = 0; x < (int)width; x++) {
            real vs = o_box3(ga, width, height, x, y);
            real hs = o_box3(gb, width, height, x, y);
            wv[(size_t)y * width + x] = hs / (hs + vs + O_RCD_EPS);
        }
    }

    /* Pass A: diagonal gradients (P: NE-SW, Q: NW-SE). */
    OMP_FOR
    for (int y = 0; y < (int)height; y++) {
        for (int x = 0; x < (int)width; x++) {
            size_t i = (size_t)y * width + x;
            ga[i] = FABSF(V(-1, 1) - V(1, -1))
                  + FABSF(K(2.0f, 2.0) * V(0, 0) - V(-2, 2) - V(2, -2));
            gb[i] = FABSF(V(-1, -1) - V(1, 1))
                  + FABSF(K(2.0f, 2.0) * V(0, 0) - V(-2, -2) - V(2, 2));
        }
    }
    OMP_FOR
    for (int y = 0; y < (int)height; y++) {
        for (int x = 0; x < (int)width; x++) {
            real ps = o_box3(ga, width, height, x, y);
            real qs = o_box3(gb, width, height, x, y);
            wp[(size_t)y * width + x] = qs / (ps + qs + O_RCD_EPS);
        }
    }

    /* Pass B: green at the red and blue sites. */
    OMP_FOR
    for (int y = 0; y < (int)height; y++) {
        for (int x = 0; x < (int)width; x++) {
            size_t i = (size_t)y * width + x;
            int s = (y & 1) * 2 + (x & 1);
            real l00, lu, ld, ll, lr, gv, gh, w;

            if (s == 1 || s == 2) {
                gp[i] = v[i];
                continue;
            }
            l00 = L(0, 0);
            lu = L(0, -2); ld = L(0, 2);
            ll = L(-2, 0); lr = L(2, 0);
            gv = K(0.5f, 0.5) * (V(0, -1) * (K(1.0f, 1.0) + o_rcd_ratio(l00, lu))
                               + V(0, 1) * (K(1.0f, 1.0) + o_rcd_ratio(l00, ld)));
            gh = K(0.5f, 0.5) * (V(-1, 0) * (K(1.0f, 1.0) + o_rcd_ratio(l00, ll))
                               + V(1, 0) * (K(1.0f, 1.0) + o_rcd_ratio(l00, lr)));
            w = wv[i];
            gp[i] = w * gv + (K(1.0f, 1.0) - w) * gh;
        }
    }

    /* Pass C, first half: red/blue at the blue/red sites from the diagonal
     * colour differences. */
    OMP_FOR
    for (int y = 0; y < (int)height; y++) {
        for (int x = 0; x < (int)width; x++) {
            size_t i = (size_t)y * width + x;
            int s = (y & 1) * 2 + (x & 1);
            real g00, dsw, dne, dnw, dse, c, w;
            real *o = &rgb[i * 4];

            if (s == 1 || s == 2)
                continue;
            g00 = gp[i];
            dsw = V(-1, 1) - G(-1, 1);
            dne = V(1, -1) - G(1, -1);
            dnw = V(-1, -1) - G(-1, -1);
            dse = V(1, 1) - G(1, 1);
            w = wp[i];
            c = w * (g00 + K(0.5f, 0.5) * (dsw + dne))
              + (K(1.0f, 1.0) - w) * (g00 + K(0.5f, 0.5) * (dnw + dse));
            cp[i] = c;
            o[0] = s == 0 ? v[i] : c;
            o[1] = g00;
            o[2] = s == 0 ? c : v[i];
            o[3] = K(1.0f, 1.0);
        }
    }

    /* Pass C, second half: both colours at the green sites. */
    OMP_FOR
    for (int y = 0; y < (int)height; y++) {
        for (int x = 0; x < (int)width; x++) {
            size_t i = (size_t)y * width + x;
            int s = (y & 1) * 2 + (x & 1);
            real g00, dh, dv, eh, ev, ch, cv, w;
            real *o = &rgb[i * 4];

            if (s == 0 || s == 3)
                continue;
            g00 = gp[i];
            w = wv[i];
            dh = K(0.5f, 0.5) * ((V(-1, 0) - G(-1, 0)) + (V(1, 0) - G(1, 0)));
            ev = K(0.5f, 0.5) * ((C(0, -1) - G(0, -1)) + (C(0, 1) - G(0, 1)));
            ch = w * (g00 + ev) + (K(1.0f, 1.0) - w) * (g00 + dh);
            dv = K(0.5f, 0.5) * ((V(0, -1) - G(0, -1)) + (V(0, 1) - G(0, 1)));
            eh = K(0.5f, 0.5) * ((C(-1, 0) - G(-1, 0)) + (C(1, 0) - G(1, 0)));
            cv = w * (g00 + dv) + (K(1.0f, 1.0) - w) * (g00 + eh);
            o[0] = s == 1 ? ch : cv;
            o[1] = v[i];
            o[2] = s == 1 ? cv : ch;
            o[3] = K(1.0f, 1.0);
        }
    }

#undef V
#undef L
#undef G
#undef C

fail:
    free(lp); free(wv); free(wp); free(gp); free(cp); free(ga); free(gb);
    return ret;
}

/* ref_debayer.c:456-472 */
int PFX(debayer_mosaic)(const real *v, uint32_t width, uint32_t height, int method, real *rgb)
{
    if (!v || !rgb)
        return -EINVAL;
    if (width < 2 || height < 2 || (width & 1) || (height & 1))
        return -EINVAL;
    switch (method) {
    case CRAW_DEBAYER_BILINEAR: o_bilinear(v, width, height, rgb); return 0;
    case CRAW_DEBAYER_MHC:      o_mhc(v, width, height, rgb); return 0;
    case CRAW_DEBAYER_RCD:      return o_rcd(v, width, height, rgb);
    default:                    return -EINVAL;
    }
}

/* ref_debayer.c:397-454 */
int PFX(debayer_wb)(const uint16_t *bayer, uint32_t width, uint32_t height, size_t stride,
                    const CRAWProfile *p, const float wb[3], int method, real *rgb)
{
    real *v;
    int ret;

    if (!bayer || !p || !wb || !rgb)
        return -EINVAL;
    if ((unsigned)method >= CRAW_DEBAYER_NB)
        return -EINVAL;
    if (p->cfa != CRAW_CFA_RGGB)
        return -ENOTSUP;
    if (width < 2 || height < 2 || (width & 1) || (height & 1))
        return -EINVAL;
    if (stride < width)
        return -EINVAL;
    for (int s = 0; s < 4; s++)
        if (!(p->white_level[s] > p->black_level[s]))
            return -EINVAL;

    v = malloc((size_t)width * height * sizeof(*v));
    if (!v)
        return -ENOMEM;
    PFX(prep)(bayer, width, height, stride, p, wb, v);
    ret = PFX(debayer_mosaic)(v, width, height, method, rgb);
    free(v);
    return ret;
}

/* ------------------------------------------------------------ denoise */

#define O_DN_SEARCH_R 3
#define O_DN_PATCH_R  1
#define O_DN_GUARD_LO K(0.90f, 0.90)
#define O_DN_GUARD_HI K(1.00f, 1.00)
#define O_DN_VAR_MIN  K(1.0e-4f, 1.0e-4)
/* 2 sigma^2 with sigma = 1 in the stabilised domain: the part of a patch
 * distance that two noisy samples of one true value carry
 * (reference DESIGN.md:1679-1700, 2149-2160). */
#define O_DN_ALLOW    K(2.0f, 2.0)
/* 1 / 9: the 3x3 patch mean of squared differences, and the same 3x3 for
 * the temporal weight (ref_denoise.c:82-97, 114-154). */
#define O_DN_INV9     K(1.0f / 9.0f, 1.0 / 9.0)

/* ORACLE_MUTANT plants one mistake, for the test that shows each closed-form
 * pin catches it (tests/test_oracle_pins.py builds these variants in a
 * temporary directory; the oracle proper is built without it). */
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif
#if ORACLE_MUTANT == 1          /* NLM allowance 2 -> 1 */
#define O_MUT_NLM_ALLOW K(1.0f, 1.0)
#else
#define O_MUT_NLM_ALLOW O_DN_ALLOW
#endif
#if ORACLE_MUTANT == 2          /* patch mean 1/9 -> 1/8 */
#define O_MUT_NLM_INV_NP K(1.0f / 8.0f, 1.0 / 8.0)
#else
#define O_MUT_NLM_INV_NP O_DN_INV9
#endif
#if ORACLE_MUTANT == 4          /* temporal 3x3 mean 1/9 -> 1/8 */
#define O_MUT_ALPHA_INV9 K(1.0f / 8.0f, 1.0 / 8.0)
#else
#define O_MUT_ALPHA_INV9 O_DN_INV9
#endif
#if ORACLE_MUTANT == 5          /* temporal allowance 2 -> 1 */
#define O_MUT_ALPHA_ALLOW K(1.0f, 1.0)
#else
#define O_MUT_ALPHA_ALLOW O_DN_ALLOW
#endif

/* lib/ref/ref_denoise.c:26-58: the generalised Anscombe transform
 * with its tangent below black, and the algebraic inverse. */
struct o_gat {
    real a, c, sc, fwd, inv, off, t0;
};

static void o_gat_init(struct o_gat *g, float a, float b)
{
    real ra = (real)a, rb = (real)b;

    g->a = ra;
#if ORACLE_MUTANT == 3          /* c without b */
    g->c = K(3.0f, 3.0) * ra * ra / K(8.0f, 8.0);
    (void)rb;
#else
    g->c = K(3.0f, 3.0) * ra * ra / K(8.0f, 8.0) + OMAX(rb, K(0.0f, 0.0));
#endif
    g->sc = SQRTF(g->c);
#if ORACLE_MUTANT == 7          /* forward scale 2/a -> 1/a */
    g->fwd = K(1.0f, 1.0) / ra;
#else
    g->fwd = K(2.0f, 2.0) / ra;
#endif
    g->inv = ra / K(4.0f, 4.0);
    g->off = g->c / ra;
    g->t0 = g->fwd * g->sc;
}

static inline real o_gat(const struct o_gat *g, real u)
{
    if (u >= K(0.0f, 0.0))
        return g->fwd * SQRTF(g->a * u + g->c);
#if ORACLE_MUTANT == 8          /* tangent slope 1/sc -> sc */
    return g->t0 + u * g->sc;
#else
    return g->t0 + u / g->sc;
#endif
}

static inline real o_igat(const struct o_gat *g, real t)
{
    if (t >= g->t0)
        return g->inv * t * t - g->off;
    return (t - g->t0) * g->sc;
}

/* ref_denoise.c:63-76: mirror in plane coordinates. */
static inline int o_dn_mirror(int v, int n)
{
    return o_mirror(v, n);
}

/* ref_denoise.c:82-97: the temporal weight of a 3x3 sum of squared
 * differences d2: alpha_max exp(-max(d2 / 9 / ht^2 - 2, 0)). */
static inline real o_alpha_weight(real d2, real alpha_max, real inv_h2)
{
    return alpha_max * EXPF(-OMAX(d2 * O_MUT_ALPHA_INV9 * inv_h2 - O_MUT_ALPHA_ALLOW,
                                  K(0.0f, 0.0)));
}

/* ref_denoise.c:82-97: the temporal weight, 3x3 of (prev - cur)^2 summed
 * row by row. */
static real o_alpha(const real *cur, const real *prev, int w2, int h2, int i, int k,
                    real alpha_max, real inv_h2)
{
    real d2 = K(0.0f, 0.0);

    for (int ky = -1; ky <= 1; ky++) {
        for (int kx = -1; kx <= 1; kx++) {
            size_t n = (size_t)o_dn_mirror(k + ky, h2) * w2 + o_dn_mirror(i + kx, w2);
            real e = prev[n] - cur[n];
            d2 += e * e;
        }
    }
    return o_alpha_weight(d2, alpha_max, inv_h2);
}

/* ref_denoise.c:268-296: the blended level and the variance it carries,
 * v = max((1 - a)^2 + a^2 v_s, 1e-4). */
static inline real o_blend(real tc, real tp, real a)
{
#if ORACLE_MUTANT == 9          /* blend towards the wrong frame */
    return tp * (K(1.0f, 1.0) - a) + tc * a;
#else
    return tc * (K(1.0f, 1.0) - a) + tp * a;
#endif
}

static inline real o_blend_var(real a, real v_s)
{
#if ORACLE_MUTANT == 10         /* (1 - a) for (1 - a)^2 */
    return OMAX((K(1.0f, 1.0) - a) + a * a * v_s, O_DN_VAR_MIN);
#else
    return OMAX((K(1.0f, 1.0) - a) * (K(1.0f, 1.0) - a) + a * a * v_s, O_DN_VAR_MIN);
#endif
}

/* ref_denoise.c:114-154: the NLM weight of a patch's sum of squared
 * differences d2. Without history (vc < 0): exp(-max(d2 / 9 / h^2 - 2, 0));
 * with it, the noise the two samples carry, vc + vq, in place of 2, and the
 * strength in units of the centre's noise: exp(-max(d2 / 9 / h^2 - (vc +
 * vq), 0) / vc). */
static inline real o_nlm_weight(real d2, real inv_h2, real vc, real vq)
{
    if (vc >= K(0.0f, 0.0)) {
#if ORACLE_MUTANT == 11         /* the centre's variance twice for vc + vq */
        real v2 = vc + vc;
        (void)vq;
#else
        real v2 = vc + vq;
#endif
#if ORACLE_MUTANT == 12         /* no division by the centre's variance */
        return EXPF(-OMAX(d2 * O_MUT_NLM_INV_NP * inv_h2 - v2, K(0.0f, 0.0)));
#else
        return EXPF(-OMAX(d2 * O_MUT_NLM_INV_NP * inv_h2 - v2, K(0.0f, 0.0)) / vc);
#endif
    }
    return EXPF(-OMAX(d2 * O_MUT_NLM_INV_NP * inv_h2 - O_MUT_NLM_ALLOW, K(0.0f, 0.0)));
}

/* ref_denoise.c:100-105: the variance as the state keeps it, its square
 * root in eight bits. */
static inline real o_var_store(real v)
{
    real q = RINTF(SQRTF(OCLAMP(v, K(0.0f, 0.0), K(1.0f, 1.0))) * K(255.0f, 255.0)) *
             K(1.0f / 255.0f, 1.0 / 255.0);
    return q * q;
}

/* ref_denoise.c:114-154: the non-local means at one sample, 7x7 search,
 * 3x3 patch, distances and weighted sums accumulated in raster order. */
static real o_nlm(const real *plane, const real *var, int w2, const int *mx, const int *my,
                  real inv_h2)
{
    const int sr = O_DN_SEARCH_R, pr = O_DN_PATCH_R;
    real cp[(2 * O_DN_PATCH_R + 1) * (2 * O_DN_PATCH_R + 1)];
    real acc = K(0.0f, 0.0), wsum = K(0.0f, 0.0);
    int n = 0;

    for (int ky = -pr; ky <= pr; ky++)
        for (int kx = -pr; kx <= pr; kx++)
            cp[n++] = plane[(size_t)my[ky] * w2 + mx[kx]];

    for (int dy = -sr; dy <= sr; dy++) {
        for (int dx = -sr; dx <= sr; dx++) {
            real d2 = K(0.0f, 0.0), wgt, sv;

            n = 0;
            for (int ky = -pr; ky <= pr; ky++) {
                for (int kx = -pr; kx <= pr; kx++) {
                    real e = cp[n++] - plane[(size_t)my[dy + ky] * w2 + mx[dx + kx]];
                    d2 += e * e;
                }
            }
            if (var)
                wgt = o_nlm_weight(d2, inv_h2, var[(size_t)my[0] * w2 + mx[0]],
                                   var[(size_t)my[dy] * w2 + mx[dx]]);
            else
                wgt = o_nlm_weight(d2, inv_h2, K(-1.0f, -1.0), K(-1.0f, -1.0));
            sv = plane[(size_t)my[dy] * w2 + mx[dx]];
            wsum += wgt;
            acc += wgt * sv;
        }
    }
    return acc / wsum;
}

/* ref_denoise.c:168-336: temporal average (when alpha_max > 0 and a
 * previous state exists), NLM (when strength > 0), inverse transform,
 * highlight guard and the 16-bit re-encode. `prev`/`next` are states of
 * 2 * w * h values indexed like the mosaic: the level, then the variance.
 * `out_level`, when not NULL, receives the value the re-encode rounds
 * (black + u * range, clamped, times 65535) for error statements. */
int PFX(denoise_temporal)(const uint16_t *bayer, uint32_t width, uint32_t height,
                          size_t stride, const CRAWProfile *profile,
                          const CRAWNoiseModel *model, float strength, float alpha_max,
                          const real *prev, real *next, uint16_t *out, size_t out_stride,
                          real *out_level)
{
    const int sr = O_DN_SEARCH_R, pr = O_DN_PATCH_R;
    const real guard_inv = K(1.0f, 1.0) / (O_DN_GUARD_HI - O_DN_GUARD_LO);
    struct o_gat g;
    real h, inv_h2, ralpha = (real)alpha_max;
    real *plane = NULL, *cur = NULL, *tprev = NULL, *var = NULL;
    int *mx = NULL, *my = NULL;
    int w2, h2, rad = sr + pr, ret = 0;

    if (!bayer || !profile || !model || !out)
        return -EINVAL;
    if (profile->cfa != CRAW_CFA_RGGB)
        return -ENOTSUP;
    if ((width & 1) || (height & 1) ||
        width < (uint32_t)(2 * (sr + pr + 2)) || height < (uint32_t)(2 * (sr + pr + 2)))
        return -EINVAL;
    if (stride < width || out_stride < width)
        return -EINVAL;
    if (!(model->a > 0.0f))
        return -EINVAL;
    if (!(strength >= 0.0f) || !(alpha_max >= 0.0f) || !(alpha_max <= 1.0f) ||
        !(strength > 0.0f || alpha_max > 0.0f))
        return -EINVAL;
    if (alpha_max > 0.0f && !next)
        return -EINVAL;
    if (!(alpha_max > 0.0f))
        prev = NULL;
    for (int s = 0; s < 4; s++)
        if (!(profile->white_level[s] > profile->black_level[s]))
            return -EINVAL;

    o_gat_init(&g, model->a, model->b);
    h = strength > 0.0f ? (real)strength : K(1.0f, 1.0);
    inv_h2 = K(1.0f, 1.0) / (h * h);

    w2 = (int)(width / 2);
    h2 = (int)(height / 2);
    plane = malloc((size_t)w2 * h2 * sizeof(*plane));
    mx = malloc((size_t)(w2 + 2 * rad) * sizeof(*mx));
    my = malloc((size_t)(h2 + 2 * rad) * sizeof(*my));
    if (prev) {
        cur = malloc((size_t)w2 * h2 * sizeof(*cur));
        tprev = malloc((size_t)w2 * h2 * sizeof(*tprev));
        var = malloc((size_t)w2 * h2 * sizeof(*var));
    }
    if (!plane || !mx || !my || (prev && (!cur || !tprev || !var))) {
        ret = -ENOMEM;
        goto done;
    }
    for (int i = 0; i < w2 + 2 * rad; i++)
        mx[i] = o_dn_mirror(i - rad, w2);
    for (int k = 0; k < h2 + 2 * rad; k++)
        my[k] = o_dn_mirror(k - rad, h2);

    for (int pl = 0; pl < 4; pl++) {
        int px = pl & 1, py = pl >> 1;
        real black = (real)profile->black_level[pl];
        real range = (real)profile->white_level[pl] - (real)profile->black_level[pl];
        real inv_range = K(1.0f, 1.0) / range;

        OMP_FOR
        for (int k = 0; k < h2; k++) {
            for (int i = 0; i < w2; i++) {
                size_t mp = (size_t)(2 * k + py) * width + 2 * i + px;
                real raw = (real)bayer[(size_t)(2 * k + py) * stride + 2 * i + px] *
                           K(1.0f / 65535.0f, 1.0 / 65535.0);
                real t = o_gat(&g, (raw - black) * inv_range);

                if (prev) {
                    cur[(size_t)k * w2 + i] = t;
                    tprev[(size_t)k * w2 + i] = o_gat(&g, prev[mp]);
                }
                plane[(size_t)k * w2 + i] = t;
            }
        }
        if (prev) {
            const real *prev_var = prev + (size_t)width * height;

            OMP_FOR
            for (int k = 0; k < h2; k++) {
                for (int i = 0; i < w2; i++) {
                    size_t c = (size_t)k * w2 + i;
                    size_t mp = (size_t)(2 * k + py) * width + 2 * i + px;
                    real a = o_alpha(cur, tprev, w2, h2, i, k, ralpha, inv_h2);

                    plane[c] = o_blend(cur[c], tprev[c], a);
                    var[c] = o_blend_var(a, prev_var[mp]);
                }
            }
        }
        if (alpha_max > 0.0f) {
            real *next_var = next + (size_t)width * height;

            OMP_FOR
            for (int k = 0; k < h2; k++) {
                for (int i = 0; i < w2; i++) {
                    size_t mp = (size_t)(2 * k + py) * width + 2 * i + px;

                    next[mp] = o_igat(&g, plane[(size_t)k * w2 + i]);
                    next_var[mp] = prev ? o_var_store(var[(size_t)k * w2 + i]) : K(1.0f, 1.0);
                }
            }
        }

        OMP_FOR
        for (int k = 0; k < h2; k++) {
            for (int i = 0; i < w2; i++) {
                real raw0, u0, u, gr, raw;

                if (strength > 0.0f)
                    u = o_igat(&g, o_nlm(plane, prev ? var : NULL, w2, mx + rad + i,
                                         my + rad + k, inv_h2));
                else
                    u = o_igat(&g, plane[(size_t)k * w2 + i]);

                raw0 = (real)bayer[(size_t)(2 * k + py) * stride + 2 * i + px] *
                       K(1.0f / 65535.0f, 1.0 / 65535.0);
                u0 = (raw0 - black) * inv_range;
                gr = OCLAMP((u0 - O_DN_GUARD_LO) * guard_inv, K(0.0f, 0.0), K(1.0f, 1.0));
                u = u + (u0 - u) * gr;

                raw = OCLAMP(black + u * range, K(0.0f, 0.0), K(1.0f, 1.0));
                out[(size_t)(2 * k + py) * out_stride + 2 * i + px] =
                    (uint16_t)RINTF(raw * K(65535.0f, 65535.0));
                if (out_level)
                    out_level[(size_t)(2 * k + py) * width + 2 * i + px] =
                        raw * K(65535.0f, 65535.0);
            }
        }
    }

done:
    free(var); free(tprev); free(cur); free(my); free(mx); free(plane);
    return ret;
}

/* Test hooks: the denoise arithmetic above, one value at a time, for the
 * closed-form checks of the tests (the same helpers the filters call). */
void PFX(gat_hook)(float a, float b, const real *u, real *t, const real *tin, real *uout,
                   size_t n, real consts[7])
{
    struct o_gat g;

    o_gat_init(&g, a, b);
    for (size_t i = 0; i < n; i++) {
        t[i] = o_gat(&g, u[i]);
        uout[i] = o_igat(&g, tin[i]);
    }
    consts[0] = g.a; consts[1] = g.c; consts[2] = g.sc; consts[3] = g.fwd;
    consts[4] = g.inv; consts[5] = g.off; consts[6] = g.t0;
}

real PFX(nlm_weight_hook)(real d2, real inv_h2, real vc, real vq)
{
    return o_nlm_weight(d2, inv_h2, vc, vq);
}

real PFX(alpha_weight_hook)(real d2, real alpha_max, real inv_h2)
{
    return o_alpha_weight(d2, alpha_max, inv_h2);
}

void PFX(blend_hook)(real tc, real tp, real a, real v_s, real out[3])
{
    out[0] = o_blend(tc, tp, a);
    out[1] = o_blend_var(a, v_s);
    out[2] = o_var_store(out[1]);
}

/* ------------------------------------------------------------- chroma */

/* ref_chroma.c:14-22: radius 4, sigma_s 2 quarter texels, K = Kc = 0.35,
 * eps 0.02 (reference DESIGN.md:1947-1967). */
#define O_CD_RADIUS    4
#if ORACLE_MUTANT == 13         /* 1 / (2 sigma_s^2) with sigma_s = 1 */
#define O_CD_INV_2S2   K(0.5f, 0.5)
#else
#define O_CD_INV_2S2   K(1.0f / (2.0f * 2.0f * 2.0f), 1.0 / 8.0)
#endif
#if ORACLE_MUTANT == 14         /* K 0.35 -> 0.25 */
#define O_CD_RANGE_K   K(0.25f, 0.25)
#else
#define O_CD_RANGE_K   K(0.35f, 0.35)
#endif
#if ORACLE_MUTANT == 15         /* Kc 0.35 -> 0.25 */
#define O_CD_CHROMA_K  K(0.25f, 0.25)
#else
#define O_CD_CHROMA_K  K(0.35f, 0.35)
#endif
#if ORACLE_MUTANT == 16         /* eps 0.02 -> 0.05 */
#define O_CD_RANGE_EPS K(0.05f, 0.05)
#else
#define O_CD_RANGE_EPS K(0.02f, 0.02)
#endif

/* ref_chroma.c:100-125: the bilateral's scales at a centre of level G:
 * 1 / (K L) and 1 / (Kc L), L = max(G, 0) + eps. */
static inline void o_cd_scales(real g, real *inv_range, real *inv_chroma)
{
    real level = OMAX(g, K(0.0f, 0.0)) + O_CD_RANGE_EPS;

    *inv_range = K(1.0f, 1.0) / (O_CD_RANGE_K * level);
    *inv_chroma = K(1.0f, 1.0) / (O_CD_CHROMA_K * level);
}

/* ref_chroma.c:100-125: the joint bilateral weight of a quarter-grid
 * offset (dx, dy) whose G, R - G and B - G differ from the centre's by
 * dg, dd0, dd1: exp(-(dx^2 + dy^2) / (2 sigma_s^2) - (dg / (K L))^2 -
 * (dd0 / (Kc L))^2 - (dd1 / (Kc L))^2). */
static inline real o_cd_weight(int dx, int dy, real dg, real dd0, real dd1, real inv_range,
                               real inv_chroma)
{
    real g = dg * inv_range, d0 = dd0 * inv_chroma, d1 = dd1 * inv_chroma;

    return EXPF(-(real)(dx * dx + dy * dy) * O_CD_INV_2S2 - g * g - d0 * d0 - d1 * d1);
}

/* ref_chroma.c:24-31: one reflection on the quarter grid. */
static inline int o_cd_mirror(int v, int n)
{
    if (v < 0)
        return -v;
    if (v >= n)
        return 2 * n - 2 - v;
    return v;
}

/* lib/ref/ref_chroma.c:33-158. */
int PFX(chroma_denoise)(real *rgb, uint32_t width, uint32_t height, float strength)
{
    real *down = NULL, *blur = NULL, rs = (real)strength;
    int w, h, wq, hq, ret = 0;

    if (!rgb || width < 4 * (O_CD_RADIUS + 1) || height < 4 * (O_CD_RADIUS + 1))
        return -EINVAL;
    if (!(strength > 0.0f) || !(strength <= 1.0f))
        return -EINVAL;

    w = (int)width;
    h = (int)height;
    wq = (w + 3) / 4;
    hq = (h + 3) / 4;
    down = malloc((size_t)wq * hq * 3 * sizeof(*down));
    blur = malloc((size_t)wq * hq * 3 * sizeof(*blur));
    if (!down || !blur) {
        ret = -ENOMEM;
        goto fail;
    }

    /* Pass 1: 4x4 block means of (R - G, B - G, G), row by row. */
    OMP_FOR
    for (int qy = 0; qy < hq; qy++) {
        for (int qx = 0; qx < wq; qx++) {
            real sum[3] = { K(0.0f, 0.0), K(0.0f, 0.0), K(0.0f, 0.0) }, n = K(0.0f, 0.0);
            real *o = &down[3 * ((size_t)qy * wq + qx)];

            for (int j = 0; j < 4; j++) {
                for (int i = 0; i < 4; i++) {
                    int x = 4 * qx + i, y = 4 * qy + j;
                    const real *px;

                    if (x >= w || y >= h)
                        continue;
                    px = &rgb[4 * ((size_t)y * w + x)];
                    sum[0] += px[0] - px[1];
                    sum[1] += px[2] - px[1];
                    sum[2] += px[1];
                    n += K(1.0f, 1.0);
                }
            }
            for (int c = 0; c < 3; c++)
                o[c] = sum[c] / n;
        }
    }

    /* Pass 2: 9x9 joint bilateral guided by G, mirrored. */
    OMP_FOR
    for (int qy = 0; qy < hq; qy++) {
        for (int qx = 0; qx < wq; qx++) {
            const real *c = &down[3 * ((size_t)qy * wq + qx)];
            real *o = &blur[3 * ((size_t)qy * wq + qx)];
            real inv_range, inv_chroma;
            real acc[2] = { K(0.0f, 0.0), K(0.0f, 0.0) }, wsum = K(0.0f, 0.0);

            o_cd_scales(c[2], &inv_range, &inv_chroma);

            for (int dy = -O_CD_RADIUS; dy <= O_CD_RADIUS; dy++) {
                int sy = o_cd_mirror(qy + dy, hq);

                for (int dx = -O_CD_RADIUS; dx <= O_CD_RADIUS; dx++) {
                    int sx = o_cd_mirror(qx + dx, wq);
                    const real *s = &down[3 * ((size_t)sy * wq + sx)];
                    real wgt = o_cd_weight(dx, dy, s[2] - c[2], s[0] - c[0], s[1] - c[1],
                                           inv_range, inv_chroma);

                    acc[0] += wgt * s[0];
                    acc[1] += wgt * s[1];
                    wsum += wgt;
                }
            }
            o[0] = acc[0] / wsum;
            o[1] = acc[1] / wsum;
            o[2] = c[2];
        }
    }

    /* Pass 3: bilinear up of (filtered - unfiltered), blended by strength. */
    OMP_FOR
    for (int y = 0; y < h; y++) {
        real fy = ((real)y + K(0.5f, 0.5)) * K(0.25f, 0.25) - K(0.5f, 0.5);
        real fly = FLOORF(fy), ty = fy - fly;
        int y0 = OCLAMP((int)fly, 0, hq - 1), y1 = OCLAMP((int)fly + 1, 0, hq - 1);

        for (int x = 0; x < w; x++) {
            real fx = ((real)x + K(0.5f, 0.5)) * K(0.25f, 0.25) - K(0.5f, 0.5);
            real flx = FLOORF(fx), tx = fx - flx;
            int x0 = OCLAMP((int)flx, 0, wq - 1), x1 = OCLAMP((int)flx + 1, 0, wq - 1);
            real w00 = (K(1.0f, 1.0) - tx) * (K(1.0f, 1.0) - ty), w10 = tx * (K(1.0f, 1.0) - ty);
            real w01 = (K(1.0f, 1.0) - tx) * ty, w11 = tx * ty;
            const real *d00 = &down[3 * ((size_t)y0 * wq + x0)], *d10 = &down[3 * ((size_t)y0 * wq + x1)];
            const real *d01 = &down[3 * ((size_t)y1 * wq + x0)], *d11 = &down[3 * ((size_t)y1 * wq + x1)];
            const real *b00 = &blur[3 * ((size_t)y0 * wq + x0)], *b10 = &blur[3 * ((size_t)y0 * wq + x1)];
            const real *b01 = &blur[3 * ((size_t)y1 * wq + x0)], *b11 = &blur[3 * ((size_t)y1 * wq + x1)];
            real *px = &rgb[4 * ((size_t)y * w + x)];
            real g = px[1];
            real d[2] = { px[0] - g, px[2] - g };

            for (int c = 0; c < 2; c++) {
                real lo = w00 * d00[c] + w10 * d10[c] + w01 * d01[c] + w11 * d11[c];
                real lf = w00 * b00[c] + w10 * b10[c] + w01 * b01[c] + w11 * b11[c];

                d[c] += rs * (lf - lo);
            }
            px[0] = g + d[0];
            px[2] = g + d[1];
        }
    }

fail:
    free(blur);
    free(down);
    return ret;
}

real PFX(cd_weight_hook)(int dx, int dy, real g, real dg, real dd0, real dd1)
{
    real inv_range, inv_chroma;

    o_cd_scales(g, &inv_range, &inv_chroma);
    return o_cd_weight(dx, dy, dg, dd0, dd1, inv_range, inv_chroma);
}

/* ------------------------------------------ highlight reconstruction */

#define O_HL_U_LO    K(0.92f, 0.92)
#define O_HL_U_HI    K(0.98f, 0.98)
#define O_HL_EPS     K(1e-4f, 1e-4)
#define O_HL_LOG_MAX K(2.0f, 2.0)
#define O_HL_W_SAT   K(1.0f, 1.0)
#define O_HL_DILATE  3
#define O_HL_LEVELS  16

/* The rebuilt mosaic's R16F store, rounded to nearest, ties to even, kept
 * as a float (lib/ref/ref.h:20-56): a subnormal half is a multiple of
 * 2^-24; a normal one keeps 10 mantissa bits. The double mode does not
 * round. */
static real o_hl_store(real v)
{
#if ORACLE_MODE == 3
    return v;
#else
    uint32_t b, a, sign;
    float f = v;

    memcpy(&b, &f, sizeof(b));
    sign = b & 0x80000000u;
    a = b & 0x7fffffffu;
    if (a == 0u || a >= 0x7f800000u)
        return f;
    if (a < 0x38800000u) {
        const float q = nearbyintf(fabsf(f) * 16777216.0f) * (1.0f / 16777216.0f);
        return sign ? -q : q;
    }
    a = (a + 0x0fffu + ((a >> 13) & 1u)) & ~0x1fffu;
    b = sign | a;
    memcpy(&f, &b, sizeof(f));
    return f;
#endif
}

static inline real o_hl_smoothstep(real u)
{
    real t = (u - O_HL_U_LO) / (O_HL_U_HI - O_HL_U_LO);

    t = OCLAMP(t, K(0.0f, 0.0), K(1.0f, 1.0));
    return t * t * (K(3.0f, 3.0) - K(2.0f, 2.0) * t);
}

/* One level of the colour field: (sum of w c_r, sum of w c_b, sum of w)
 * per texel, then (c_r, c_b, w) once pushed. */
struct o_hl_level {
    int w, h;
    real *t;
};

/* A texel's colour, its coordinates clamped to the level; the top level
 * still holds sums and is divided on the way. */
static void o_hl_texel(const struct o_hl_level *l, int x, int y, int top, real c[2])
{
    const real *t;

    x = x < 0 ? 0 : (x >= l->w ? l->w - 1 : x);
    y = y < 0 ? 0 : (y >= l->h ? l->h - 1 : y);
    t = l->t + 3 * ((size_t)y * l->w + x);
    c[0] = top ? (t[2] > 0 ? t[0] / t[2] : 0) : t[0];
    c[1] = top ? (t[2] > 0 ? t[1] / t[2] : 0) : t[1];
}

/* Bilinear in texel coordinates, x first then y, each as a (1 - f) + b f. */
static void o_hl_bilinear(const struct o_hl_level *l, real px, real py, int top, real out[2])
{
    const real x0 = FLOORF(px), y0 = FLOORF(py), fx = px - x0, fy = py - y0;
    const int ix = (int)x0, iy = (int)y0;
    real c00[2], c10[2], c01[2], c11[2];

    o_hl_texel(l, ix, iy, top, c00);
    o_hl_texel(l, ix + 1, iy, top, c10);
    o_hl_texel(l, ix, iy + 1, top, c01);
    o_hl_texel(l, ix + 1, iy + 1, top, c11);
    for (int i = 0; i < 2; i++) {
        const real a = c00[i] * (1 - fx) + c10[i] * fx;
        const real b = c01[i] * (1 - fx) + c11[i] * fx;

        out[i] = a * (1 - fy) + b * fy;
    }
}

/* lib/ref/ref_highlight.c:91-335: the clip ramp a = smoothstep(0.92, 0.98,
 * u) per site; the colour field's level 0 (a texel per 4x4 block: the
 * block's quads, top row first, each weighted (1 - a over the 7x7 quads
 * about it) min(u, 1)^2 with the clamped logs of R/G and B/G); the pull to
 * 1x1 (2x2 sums); the push back down (each level its own colour where it
 * has weight, blended towards the bilinear of the pushed level above by
 * min(w, 1)); and the fill: every ramped site raised (never lowered)
 * towards the level its unclipped neighbours and the field's colour
 * predict, blended by a. `out` is the balanced mosaic, R16F-rounded. */
int PFX(highlights)(const uint16_t *bayer, uint32_t width, uint32_t height, size_t stride,
                    const CRAWProfile *p, const float wb[3], real *out)
{
    const int w = (int)width, h = (int)height, qw = w / 2, qh = h / 2;
    struct o_hl_level lv[O_HL_LEVELS];
    int nl = 0, ret = 0;
    real *u = NULL, *a = NULL, *qa = NULL, rwb[3], inv_range[4];

    memset(lv, 0, sizeof(lv));
    if (!bayer || !p || !wb || !out)
        return -EINVAL;
    if (p->cfa != CRAW_CFA_RGGB)
        return -ENOTSUP;
    if (width < 8 || height < 8 || (width & 1) || (height & 1) || stride < width)
        return -EINVAL;
    for (int s = 0; s < 4; s++) {
        if (!(p->white_level[s] > p->black_level[s]))
            return -EINVAL;
        inv_range[s] = K(1.0f, 1.0) / ((real)p->white_level[s] - (real)p->black_level[s]);
    }
    for (int c = 0; c < 3; c++)
        rwb[c] = wb[c];

    u = malloc((size_t)w * h * sizeof(*u));
    a = malloc((size_t)w * h * sizeof(*a));
    qa = malloc((size_t)qw * qh * sizeof(*qa));
    if (!u || !a || !qa) {
        ret = -ENOMEM;
        goto done;
    }
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            const int s = (y & 1) * 2 + (x & 1);
            const real raw = (real)bayer[(size_t)y * stride + x] * K(1.0f / 65535.0f, 1.0 / 65535.0);
            const size_t o = (size_t)y * w + x;

            u[o] = (raw - (real)p->black_level[s]) * inv_range[s];
            a[o] = o_hl_smoothstep(u[o]);
        }
    }
    for (int qy = 0; qy < qh; qy++) {
        for (int qx = 0; qx < qw; qx++) {
            real m = 0;

            for (int s = 0; s < 4; s++)
                m = OMAX(m, a[(size_t)(2 * qy + (s >> 1)) * w + 2 * qx + (s & 1)]);
            qa[(size_t)qy * qw + qx] = m;
        }
    }

    /* the pyramid: level 0 a texel per 4x4 block, halved (rounding up) to 1x1 */
    for (int lw = (w + 3) / 4, lh = (h + 3) / 4;; lw = (lw + 1) / 2, lh = (lh + 1) / 2) {
        if (nl == O_HL_LEVELS) {
            ret = -EINVAL;
            goto done;
        }
        lv[nl].w = lw;
        lv[nl].h = lh;
        lv[nl].t = calloc((size_t)lw * lh * 3, sizeof(real));
        if (!lv[nl].t) {
            ret = -ENOMEM;
            goto done;
        }
        nl++;
        if (lw == 1 && lh == 1)
            break;
    }

    /* level 0 */
    for (int by = 0; by < lv[0].h; by++) {
        for (int bx = 0; bx < lv[0].w; bx++) {
            real *t = lv[0].t + 3 * ((size_t)by * lv[0].w + bx);

            for (int j = 0; j < 2; j++) {
                for (int i = 0; i < 2; i++) {
                    const int qx = 2 * bx + i, qy = 2 * by + j;
                    real amax = 0, umax = 0, rgb[3] = { 0, 0, 0 }, m, wq, g, r, b;

                    if (qx >= qw || qy >= qh)
                        continue;
                    for (int ny = qy - O_HL_DILATE; ny <= qy + O_HL_DILATE; ny++)
                        for (int nx = qx - O_HL_DILATE; nx <= qx + O_HL_DILATE; nx++)
                            if (nx >= 0 && ny >= 0 && nx < qw && ny < qh)
                                amax = OMAX(amax, qa[(size_t)ny * qw + nx]);
                    for (int s = 0; s < 4; s++) {
                        const real v = u[(size_t)(2 * qy + (s >> 1)) * w + 2 * qx + (s & 1)];

                        rgb[o_site_chan[s]] += v * rwb[o_site_chan[s]];
                        umax = OMAX(umax, v);
                    }
                    rgb[1] *= K(0.5f, 0.5);
                    m = OMIN(umax, K(1.0f, 1.0));
                    wq = (1 - amax) * m * m;
                    g = OMAX(rgb[1], O_HL_EPS);
                    r = OCLAMP(LOGF(OMAX(rgb[0], O_HL_EPS) / g), -O_HL_LOG_MAX, O_HL_LOG_MAX);
                    b = OCLAMP(LOGF(OMAX(rgb[2], O_HL_EPS) / g), -O_HL_LOG_MAX, O_HL_LOG_MAX);
                    t[0] += wq * r;
                    t[1] += wq * b;
                    t[2] += wq;
                }
            }
        }
    }

    /* pull */
    for (int l = 0; l + 1 < nl; l++) {
        const struct o_hl_level *f = &lv[l];
        struct o_hl_level *c = &lv[l + 1];

        for (int y = 0; y < c->h; y++)
            for (int x = 0; x < c->w; x++)
                for (int dy = 0; dy < 2; dy++)
                    for (int dx = 0; dx < 2; dx++) {
                        const int fx = 2 * x + dx, fy = 2 * y + dy;

                        if (fx < f->w && fy < f->h)
                            for (int k = 0; k < 3; k++)
                                c->t[3 * ((size_t)y * c->w + x) + k] +=
                                    f->t[3 * ((size_t)fy * f->w + fx) + k];
                    }
    }

    /* push */
    for (int l = nl - 2; l >= 0; l--) {
        struct o_hl_level *f = &lv[l];
        const int top = l + 2 == nl;

        for (int y = 0; y < f->h; y++) {
            for (int x = 0; x < f->w; x++) {
                real *t = f->t + 3 * ((size_t)y * f->w + x), up[2];

                o_hl_bilinear(&lv[l + 1], ((real)x + K(0.5f, 0.5)) * K(0.5f, 0.5) - K(0.5f, 0.5),
                              ((real)y + K(0.5f, 0.5)) * K(0.5f, 0.5) - K(0.5f, 0.5), top, up);
                if (t[2] > 0) {
                    const real sw = OMIN(t[2] / O_HL_W_SAT, K(1.0f, 1.0));

                    t[0] = up[0] + sw * (t[0] / t[2] - up[0]);
                    t[1] = up[1] + sw * (t[1] / t[2] - up[1]);
                } else {
                    t[0] = up[0];
                    t[1] = up[1];
                }
            }
        }
    }

    /* fill */
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            const size_t o = (size_t)y * w + x;
            const int c = o_site_chan[(y & 1) * 2 + (x & 1)];
            const real xc = u[o] * rwb[c], av = a[o];
            real res = xc;

            if (av > 0) {
                real sum[3] = { 0, 0, 0 }, n[3] = { 0, 0, 0 }, am[3] = { 0, 0, 0 };
                real lf[2], lrho[3], a1, a2, l1, l2, lc, A, lfull, lpart, L;
                int k1, k2;

                o_hl_bilinear(&lv[0], (real)(x >> 1) * K(0.5f, 0.5) - K(0.25f, 0.25),
                              (real)(y >> 1) * K(0.5f, 0.5) - K(0.25f, 0.25), 0, lf);
                lrho[0] = lf[0];
                lrho[1] = 0;
                lrho[2] = lf[1];
                for (int dy = -1; dy <= 1; dy++) {
                    for (int dx = -1; dx <= 1; dx++) {
                        const int nx = o_mirror(x + dx, w), ny = o_mirror(y + dy, h);
                        const int k = o_site_chan[(ny & 1) * 2 + (nx & 1)];
                        const size_t no = (size_t)ny * w + nx;

                        if (k == c)
                            continue;
                        sum[k] += u[no] * rwb[k];
                        n[k] += 1;
                        am[k] = OMAX(am[k], a[no]);
                    }
                }
                k1 = c == 0 ? 1 : 0;
                k2 = c == 2 ? 1 : 2;
                a1 = 1 - am[k1];
                a2 = 1 - am[k2];
                l1 = LOGF(OMAX(sum[k1] / n[k1], O_HL_EPS)) - lrho[k1];
                l2 = LOGF(OMAX(sum[k2] / n[k2], O_HL_EPS)) - lrho[k2];
                lc = LOGF(OMAX(xc, O_HL_EPS)) - lrho[c];
                A = a1 + a2;
                lfull = EXPF(OMAX(lc, OMAX(l1, l2)));
                lpart = A > 0 ? EXPF((a1 * l1 + a2 * l2) / A) : lfull;
                L = lfull + OMIN(A, K(1.0f, 1.0)) * (lpart - lfull);
                res = xc + av * OMAX(L * EXPF(lrho[c]) - xc, K(0.0f, 0.0));
            }
            out[o] = o_hl_store(res);
        }
    }

done:
    for (int i = 0; i < nl; i++)
        free(lv[i].t);
    free(u);
    free(a);
    free(qa);
    return ret;
}

/* -------------------------------------------------------- false colour */

/* lib/ref/ref_false_color.c:27-41: the median of nine as a selection, by
 * an insertion sort (so the result is one of the nine, bit for bit). */
static real o_fc_median9(real v[9])
{
    for (int i = 1; i < 9; i++) {
        const real x = v[i];
        int j = i;

        for (; j > 0 && v[j - 1] > x; j--)
            v[j] = v[j - 1];
        v[j] = x;
    }
    return v[4];
}

/* lib/ref/ref_false_color.c:46-54: the nine summed in stencil order (top
 * row left to right first), then scaled by the nearest 1/9. */
static real o_fc_mean9(const real v[9])
{
    real s = v[0];

    for (int i = 1; i < 9; i++)
        s += v[i];
    return s * K(1.0f / 9.0f, 1.0 / 9.0);
}

/* lib/ref/ref_false_color.c:56-134: `passes` passes, each a 3x3 median
 * sweep then a 3x3 box sweep over the whole frame, on the two colour
 * differences R - G and B - G (borders mirrored, one reflection); G and
 * alpha stay as they are and R, B become G plus the filtered differences.
 * Held here as two planes; every sweep reads the whole previous one. */
int PFX(false_color)(real *rgb, uint32_t width, uint32_t height, uint32_t passes)
{
    const int w = (int)width, h = (int)height;
    const size_t n = (size_t)width * height;
    real *a = NULL, *b = NULL;

    if (!rgb || width < 2 || height < 2)
        return -EINVAL;
    if (passes < 1 || passes > CRAW_FALSE_COLOR_MAX_PASSES)
        return -EINVAL;
    a = malloc(2 * n * sizeof(*a));
    b = malloc(2 * n * sizeof(*b));
    if (!a || !b) {
        free(a);
        free(b);
        return -ENOMEM;
    }
    for (size_t i = 0; i < n; i++) {
        a[i] = rgb[4 * i + 0] - rgb[4 * i + 1];
        a[n + i] = rgb[4 * i + 2] - rgb[4 * i + 1];
    }
    for (uint32_t sweep = 0; sweep < 2 * passes; sweep++) {
        const int blur = sweep & 1;
        real *t;

        OMP_FOR
        for (int y = 0; y < h; y++) {
            for (int x = 0; x < w; x++) {
                for (int c = 0; c < 2; c++) {
                    const real *src = a + c * n;
                    real v[9];

                    for (int j = 0; j < 3; j++)
                        for (int i = 0; i < 3; i++)
                            v[3 * j + i] = src[(size_t)o_mirror(y - 1 + j, h) * w +
                                               o_mirror(x - 1 + i, w)];
                    b[c * n + (size_t)y * w + x] = blur ? o_fc_mean9(v) : o_fc_median9(v);
                }
            }
        }
        t = a;
        a = b;
        b = t;
    }
    for (size_t i = 0; i < n; i++) {
        const real g = rgb[4 * i + 1];

        rgb[4 * i + 0] = g + a[i];
        rgb[4 * i + 2] = g + a[n + i];
    }
    free(a);
    free(b);
    return 0;
}

/* -------------------------------------------------------------- colour */

/* gamut.c:66-92: ACES 1.3 reference gamut compression of one AP1 triple. */
static inline void o_gamut_compress(const real thr[3], const real scl[3], real pwr, real inv_pwr,
                                    real rgb[3])
{
    real ach = OMAX(rgb[0], OMAX(rgb[1], rgb[2]));
    real mag;

    if (ach == K(0.0f, 0.0))
        return;
    mag = FABSF(ach);
    for (int i = 0; i < 3; i++) {
        real d = (ach - rgb[i]) / mag;
        real nd, p;

        if (d < thr[i])
            continue;
        nd = (d - thr[i]) / scl[i];
        p = POWF(nd, pwr);
        d = thr[i] + scl[i] * nd / POWF(K(1.0f, 1.0) + p, inv_pwr);
        rgb[i] = ach - d * mag;
    }
}

/* The operator alone on `n` AP1 triples in place, with the float (or,
 * in the double mode, double) constants: gamut.c:66-92. */
void PFX(gamut_compress)(real *rgb, size_t n)
{
    real thr[3], scl[3], pwr, inv_pwr;

#if ORACLE_MODE == 3
    oracle_gamut_params_d(thr, scl, &pwr, &inv_pwr);
#else
    oracle_gamut_params(thr, scl, &pwr, &inv_pwr);
#endif
    for (size_t i = 0; i < n; i++)
        o_gamut_compress(thr, scl, pwr, inv_pwr, &rgb[3 * i]);
}

/* ref_color.c:10-76: matrix, gain and exposure to XYZ, XYZ to the output
 * space (two matrix-vector products, as the reference does them), gamut
 * compression for ACEScg, optional clamp; alpha 1. */
int PFX(color)(real *rgb, uint32_t width, uint32_t height, const CRAWFrameParams *fp,
               int out_space)
{
    CRAWColorParams par;
    real m_out[3][3], cam_to_xyz[3][3], scale, thr[3], scl[3], pwr, inv_pwr;
    bool clamp_negative, compress;
    int ret;

    if (!rgb || !fp || !fp->profile)
        return -EINVAL;
    ret = oracle_profile_eval(fp->profile, fp->wb_mode, &fp->illuminant, &par);
    if (ret < 0)
        return ret;

#if ORACLE_MODE == 3
    {
        double md[3][3], td[3], sd[3], pd, ipd;

        ret = oracle_colorspace_from_xyz_d((enum CRAWColorSpace)out_space, md);
        if (ret < 0)
            return ret;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                m_out[i][j] = md[i][j];
        oracle_gamut_params_d(td, sd, &pd, &ipd);
        for (int i = 0; i < 3; i++) {
            thr[i] = td[i];
            scl[i] = sd[i];
        }
        pwr = pd;
        inv_pwr = ipd;
        scale = (double)par.gain * exp2((double)fp->exposure_stops);
    }
#else
    {
        float mf[3][3];

        ret = oracle_colorspace_from_xyz_d65((enum CRAWColorSpace)out_space, mf);
        if (ret < 0)
            return ret;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                m_out[i][j] = mf[i][j];
        oracle_gamut_params(thr, scl, &pwr, &inv_pwr);
        scale = par.gain * exp2f(fp->exposure_stops);
    }
#endif
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            cam_to_xyz[i][j] = par.cam_to_xyz[i][j];

    clamp_negative = (fp->flags & CRAW_FRAME_FLAG_CLAMP_NEGATIVE) != 0;
    compress = out_space == CRAW_CS_ACES_AP1_LINEAR &&
               !(fp->flags & CRAW_FRAME_FLAG_NO_GAMUT_COMPRESSION);

    OMP_FOR
    for (size_t i = 0; i < (size_t)width * height; i++) {
        real *px = &rgb[i * 4];
        real cam[3], xyz[3], out[3];

        for (int c = 0; c < 3; c++)
            cam[c] = px[c];
        for (int r = 0; r < 3; r++)
            xyz[r] = (cam_to_xyz[r][0] * cam[0] + cam_to_xyz[r][1] * cam[1]
                      + cam_to_xyz[r][2] * cam[2]) * scale;
        for (int r = 0; r < 3; r++)
            out[r] = m_out[r][0] * xyz[0] + m_out[r][1] * xyz[1] + m_out[r][2] * xyz[2];
        if (compress)
            o_gamut_compress(thr, scl, pwr, inv_pwr, out);
        if (clamp_negative)
            for (int r = 0; r < 3; r++)
                out[r] = OMAX(out[r], K(0.0f, 0.0));
        px[0] = out[0];
        px[1] = out[1];
        px[2] = out[2];
        px[3] = K(1.0f, 1.0);
    }
    return 0;
}

/* ------------------------------------------------------------------ AWB */

/* lib/awb.c:213-296: one sample per 8x8 block (partial blocks
 * at the edges average what they have), block means summed row by row,
 * clipped and near-black blocks skipped, ln(R/G) and ln(B/G) of the raw
 * response binned over [-2.5, 1] into 128 bins, weights in 16.16. */
int PFX(awb_histogram)(const real *rgb, uint32_t width, uint32_t height, size_t stride,
                       int channels, const float wb[3], uint32_t *count, uint32_t *weight,
                       uint32_t *bright)
{
    const uint32_t bins = CRAW_AWB_HIST_BINS;
    const real lmin = K(CRAW_AWB_HIST_LOG_MIN, -2.5), lmax = K(CRAW_AWB_HIST_LOG_MAX, 1.0);

    if (!rgb || !wb || !count || !weight || !bright)
        return -EINVAL;
    if (channels != 3 && channels != 4)
        return -EINVAL;
    if (stride < (size_t)width * (size_t)channels)
        return -EINVAL;
    if (!(wb[0] > 0.0f) || !(wb[2] > 0.0f))
        return -EINVAL;

    memset(count, 0, (size_t)bins * bins * sizeof(*count));
    memset(weight, 0, (size_t)bins * bins * sizeof(*weight));
    memset(bright, 0, (size_t)bins * bins * sizeof(*bright));

    for (uint32_t by = 0; by < height; by += 8) {
        uint32_t y1 = OMIN(by + 8, height);

        for (uint32_t bx = 0; bx < width; bx += 8) {
            uint32_t x1 = OMIN(bx + 8, width);
            real acc[3] = { K(0.0f, 0.0), K(0.0f, 0.0), K(0.0f, 0.0) };
            real n = (real)((y1 - by) * (x1 - bx));
            real r, g, b, rr, bb, x, y, w;
            uint32_t ix, iy;

            for (uint32_t py = by; py < y1; py++) {
                const real *row = rgb + (size_t)py * stride;
                for (uint32_t px = bx; px < x1; px++) {
                    acc[0] += row[(size_t)px * channels + 0];
                    acc[1] += row[(size_t)px * channels + 1];
                    acc[2] += row[(size_t)px * channels + 2];
                }
            }
            r = acc[0] / n;
            g = acc[1] / n;
            b = acc[2] / n;
            if (OMAX(OMAX(r, g), b) >= K(0.98f, 0.98) || g < K(0.002f, 0.002))
                continue;
            rr = (r / (real)wb[0]) / g;
            bb = (b / (real)wb[2]) / g;
            x = LOGF(rr);
            y = LOGF(bb);
            x = x > lmin ? (x < lmax ? x : lmax) : lmin;
            y = y > lmin ? (y < lmax ? y : lmax) : lmin;
            {
                real span = lmax - lmin;
                int bin_x = (int)FLOORF((x - lmin) / span * (real)bins);
                int bin_y = (int)FLOORF((y - lmin) / span * (real)bins);

                ix = (uint32_t)OCLAMP(bin_x, 0, (int)bins - 1);
                iy = (uint32_t)OCLAMP(bin_y, 0, (int)bins - 1);
            }
            w = OMIN(g, K(0.5f, 0.5));
            count[iy * bins + ix] += 1;
            weight[iy * bins + ix] += (uint32_t)ROUNDF(w * K(65536.0f, 65536.0));
            bright[iy * bins + ix] += (uint32_t)ROUNDF(g * g * g * g * K(65536.0f, 65536.0));
        }
    }
    return 0;
}

/* -------------------------------------------------------------- process */

/* ref_color.c:91-230 restricted to this path's stages: [temporal +]
 * [spatial denoise,] debayer with the frame's gains, [chroma denoise,]
 * [AWB histogram of that intermediate, with the frame's gains, as
 * pipeline.c:2235-2253 builds it,] colour. `prev`/`next` give the temporal
 * state as craw_ref_denoise_temporal() takes it (NULL prev: first frame);
 * `clean` (w*h), `inter` (w*h*4) and the histogram arrays are optional
 * outputs. */
int PFX(process)(const uint16_t *bayer, uint32_t width, uint32_t height, size_t stride,
                 const CRAWPipelineParams *pp, const CRAWFrameParams *fp, const real *prev,
                 real *next, uint16_t *clean_out, real *inter, uint32_t *count,
                 uint32_t *weight, uint32_t *bright, real *out)
{
    uint16_t *clean = NULL;
    real *state = NULL;
    float strength, alpha_max;
    CRAWColorParams par;
    int out_space, ret;

    if (!pp || !fp || !fp->profile || !out)
        return -EINVAL;
    if (pp->width && (pp->width != width || pp->height != height))
        return -EINVAL;
    if (pp->cfa != fp->profile->cfa)
        return -EINVAL;

    strength = (pp->flags & CRAW_PIPELINE_FLAG_DENOISE) ? fp->denoise_strength : 0.0f;
    alpha_max = (pp->flags & CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE)
              ? fp->denoise_temporal_strength : 0.0f;
    strength = OMAX(strength, 0.0f);
    alpha_max = OMAX(alpha_max, 0.0f);
    if (strength > 0.0f || alpha_max > 0.0f) {
        CRAWNoiseModel nm;

        ret = oracle_noise_model(fp, &nm);
        if (ret < 0)
            return ret;
        clean = malloc((size_t)width * height * sizeof(*clean));
        if (alpha_max > 0.0f && !next)
            state = malloc(2 * (size_t)width * height * sizeof(*state));
        if (!clean || (alpha_max > 0.0f && !next && !state)) {
            ret = -ENOMEM;
            goto done;
        }
        ret = PFX(denoise_temporal)(bayer, width, height, stride, fp->profile, &nm,
                                    strength, alpha_max, prev, next ? next : state,
                                    clean, width, NULL);
        if (ret < 0)
            goto done;
        if (clean_out)
            memcpy(clean_out, clean, (size_t)width * height * sizeof(*clean));
        bayer = clean;
        stride = width;
    }

    ret = oracle_profile_eval(fp->profile, fp->wb_mode, &fp->illuminant, &par);
    if (ret < 0)
        goto done;
    /* ref_color.c:148-171: the rebuilt mosaic replaces the prep and its
     * clip */
    if ((pp->flags & CRAW_PIPELINE_FLAG_HIGHLIGHT_RECONSTRUCTION) &&
        (fp->flags & CRAW_FRAME_FLAG_HIGHLIGHT_RECONSTRUCTION)) {
        real *v = malloc((size_t)width * height * sizeof(*v));

        if (!v) {
            ret = -ENOMEM;
            goto done;
        }
        ret = PFX(highlights)(bayer, width, height, stride, fp->profile, par.wb, v);
        if (ret >= 0)
            ret = PFX(debayer_mosaic)(v, width, height, pp->debayer, out);
        free(v);
    } else {
        ret = PFX(debayer_wb)(bayer, width, height, stride, fp->profile, par.wb, pp->debayer,
                              out);
    }
    if (ret < 0)
        goto done;
    /* ref_color.c:178-185: the false colour on the intermediate first */
    if ((pp->flags & CRAW_PIPELINE_FLAG_FALSE_COLOR) && fp->false_color_passes > 0) {
        ret = PFX(false_color)(out, width, height, fp->false_color_passes);
        if (ret < 0)
            goto done;
    }
    if ((pp->flags & CRAW_PIPELINE_FLAG_CHROMA_DENOISE) && fp->denoise_chroma_strength > 0.0f) {
        ret = PFX(chroma_denoise)(out, width, height, fp->denoise_chroma_strength);
        if (ret < 0)
            goto done;
    }
    if (inter)
        memcpy(inter, out, (size_t)width * height * 4 * sizeof(*out));
    if (count && weight && bright) {
        ret = PFX(awb_histogram)(out, width, height, (size_t)width * 4, 4, par.wb, count,
                                 weight, bright);
        if (ret < 0)
            goto done;
    }
    out_space = fp->out_space_set ? (int)fp->out_space : (int)pp->out_space;
    ret = PFX(color)(out, width, height, fp, out_space);

done:
    free(state);
    free(clean);
    return ret;
}
