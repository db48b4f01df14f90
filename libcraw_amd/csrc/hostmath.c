/*
 * libcraw for MI355X: host-side math of the path, the per-frame constants
 * the kernels are handed.
 *
 *   craw_profile_*            lib/profile.c:26-252, 255-415 of cyanreg/libcraw
 *   craw_noise_model_eval     lib/noise.c:37-110
 *   craw_colorspace_*, craw_planckian_xy, craw_bradford_adaptation
 *                             lib/colorspace.c:25-319
 *   craw_gamut_params         lib/gamut.c:20-64 (ACES 1.3 RGC constants)
 *   craw_awb_histogram_from_rgb  lib/awb.c:213-296
 *
 * Everything here runs once per frame or once per pipeline on the host
 * and keeps the reference's precision choices (double where it solves or
 * interpolates, float where it stores), so that the constants the kernels
 * receive are the reference's bit for bit. Built without floating-point
 * contraction for the same reason.
 */
#include <math.h>
#include <string.h>

#include <craw/pipeline.h>
#include "hostmath.h"

#define INV_SQRT2 0.70710678118654752440

void craw_profile_init(CRAWProfile *p)
{
    if (!p)
        return;
    memset(p, 0, sizeof(*p));
    p->cfa = CRAW_CFA_RGGB;
    for (int s = 0; s < 4; s++)
        p->white_level[s] = 1.0f;
    for (int c = 0; c < 3; c++) {
        p->as_shot_wb[c] = 1.0f;
        p->as_shot_matrix[c][c] = 1.0f;
    }
    p->gain = 1.0f;
}

static bool all_finite(const float *v, size_t n)
{
    for (size_t i = 0; i < n; i++)
        if (!isfinite(v[i]))
            return false;
    return true;
}

int craw_profile_validate(const CRAWProfile *p)
{
    if (!p || (unsigned)p->cfa > CRAW_CFA_GBRG)
        return CRAW_ERROR(EINVAL);

    for (int s = 0; s < 4; s++) {
        if (!isfinite(p->black_level[s]) || !isfinite(p->white_level[s]) ||
            p->black_level[s] < 0.0f || p->white_level[s] > 1.0f ||
            !(p->white_level[s] > p->black_level[s]))
            return CRAW_ERROR(EINVAL);
    }
    if (!all_finite(p->as_shot_wb, 3))
        return CRAW_ERROR(EINVAL);
    for (int c = 0; c < 3; c++)
        if (!(p->as_shot_wb[c] > 0.0f))
            return CRAW_ERROR(EINVAL);
    if (!all_finite(&p->as_shot_matrix[0][0], 9))
        return CRAW_ERROR(EINVAL);
    {
        const float (*m)[3] = p->as_shot_matrix;
        double det = (double)m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1])
                   - (double)m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
                   + (double)m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        if (!(fabs(det) > 1e-9))
            return CRAW_ERROR(EINVAL);
    }
    if (!isfinite(p->gain) || !(p->gain > 0.0f) ||
        !isfinite(p->as_shot_cct) || p->as_shot_cct < 0.0f ||
        !isfinite(p->as_shot_ei) || p->as_shot_ei < 0.0f ||
        !isfinite(p->exposure_index) || p->exposure_index < 0.0f)
        return CRAW_ERROR(EINVAL);
    if (p->nb_wb_knots > CRAW_PROFILE_MAX_KNOTS || p->nb_matrix_knots > CRAW_PROFILE_MAX_KNOTS ||
        p->nb_noise_knots > CRAW_PROFILE_MAX_NOISE_KNOTS)
        return CRAW_ERROR(EINVAL);
    for (uint32_t i = 0; i < p->nb_wb_knots; i++) {
        const CRAWWbKnot *k = &p->wb_knots[i];
        if (!isfinite(k->cct) || !(k->cct > 0.0f) ||
            (i && !(k->cct > p->wb_knots[i - 1].cct)) || !all_finite(k->wb, 3))
            return CRAW_ERROR(EINVAL);
        for (int c = 0; c < 3; c++)
            if (!(k->wb[c] > 0.0f))
                return CRAW_ERROR(EINVAL);
    }
    for (uint32_t i = 0; i < p->nb_matrix_knots; i++) {
        const CRAWMatrixKnot *k = &p->matrix_knots[i];
        if (!isfinite(k->cct) || !(k->cct > 0.0f) ||
            (i && !(k->cct > p->matrix_knots[i - 1].cct)) || !all_finite(&k->m[0][0], 9))
            return CRAW_ERROR(EINVAL);
    }
    if (p->name[sizeof(p->name) - 1] != '\0')
        return CRAW_ERROR(EINVAL);
    return 0;
}

/* The knot interval containing `cct` and the position inside it; knots
 * are interpolated linearly in Kelvin, clamped to the table. */
static uint32_t knot_find(const float *cct_of, size_t stride_floats, uint32_t n, double cct,
                          double *frac)
{
    uint32_t i = 0;
    double a, b, f = 0.0;

    while (i + 2 < n && cct >= cct_of[(i + 1) * stride_floats])
        i++;
    a = cct_of[i * stride_floats];
    b = cct_of[(i + 1) * stride_floats];
    if (b > a)
        f = (cct - a) / (b - a);
    *frac = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
    return i;
}

static double clampd(double v, double lo, double hi)
{
    v = v > lo ? v : lo;
    return v < hi ? v : hi;
}

/* The white balance gains at `cct` (two knots or more: linear in Kelvin
 * inside the interval, the position clamped to it). */
static void wb_gains_at(const CRAWProfile *p, double cct, double wb[3])
{
    double f;
    uint32_t i;

    if (p->nb_wb_knots == 1) {
        for (int c = 0; c < 3; c++)
            wb[c] = p->wb_knots[0].wb[c];
        return;
    }
    i = knot_find(&p->wb_knots[0].cct, sizeof(CRAWWbKnot) / sizeof(float), p->nb_wb_knots, cct,
                  &f);
    for (int c = 0; c < 3; c++) {
        double a = p->wb_knots[i].wb[c], b = p->wb_knots[i + 1].wb[c];
        wb[c] = a + (b - a) * f;
    }
}

int craw_profile_eval(const CRAWProfile *p, enum CRAWWbMode mode,
                      const CRAWIlluminant *illum, CRAWColorParams *out)
{
    if (!p || !out)
        return CRAW_ERROR(EINVAL);

    if (mode == CRAW_WB_AS_SHOT) {
        memcpy(out->wb, p->as_shot_wb, sizeof(out->wb));
        memcpy(out->cam_to_xyz, p->as_shot_matrix, sizeof(out->cam_to_xyz));
    } else if (mode == CRAW_WB_REBALANCE) {
        double wb[3], cct, f, tint_gain;
        uint32_t i;

        if (!illum)
            return CRAW_ERROR(EINVAL);
        if (!p->nb_wb_knots)
            return CRAW_ERROR(ENOTSUP);
        if (!isfinite(illum->cct) || !isfinite(illum->tint))
            return CRAW_ERROR(EINVAL);

        cct = clampd((double)illum->cct, (double)p->wb_knots[0].cct,
                     (double)p->wb_knots[p->nb_wb_knots - 1].cct);
        wb_gains_at(p, cct, wb);
        tint_gain = exp((double)illum->tint * INV_SQRT2);
        wb[0] *= tint_gain;
        wb[2] *= tint_gain;
        for (int c = 0; c < 3; c++)
            out->wb[c] = (float)wb[c];

        if (!p->nb_matrix_knots) {
            memcpy(out->cam_to_xyz, p->as_shot_matrix, sizeof(out->cam_to_xyz));
        } else if (p->nb_matrix_knots == 1) {
            memcpy(out->cam_to_xyz, p->matrix_knots[0].m, sizeof(out->cam_to_xyz));
        } else {
            uint32_t nb = p->nb_matrix_knots;
            double mcct = clampd(cct, (double)p->matrix_knots[0].cct,
                                 (double)p->matrix_knots[nb - 1].cct);

            i = knot_find(&p->matrix_knots[0].cct, sizeof(CRAWMatrixKnot) / sizeof(float),
                          nb, mcct, &f);
            for (int r = 0; r < 3; r++) {
                for (int c = 0; c < 3; c++) {
                    double a = p->matrix_knots[i].m[r][c], b = p->matrix_knots[i + 1].m[r][c];
                    out->cam_to_xyz[r][c] = (float)(a + (b - a) * f);
                }
            }
        }
    } else {
        return CRAW_ERROR(EINVAL);
    }

    out->gain = p->gain;
    if (p->as_shot_ei > 0.0f && p->exposure_index > 0.0f)
        out->gain *= p->exposure_index / p->as_shot_ei;
    return 0;
}

/* ------------------------------------------------ the white balance locus */

/* Resampling step of the locus walk, in mired (equal steps in mired: the
 * sensor's ratios and perceived balance are both near linear in it). */
#define LOCUS_STEP_MIRED 5.0

/* How far the neutral (lx, ly) = (ln R/G, ln B/G) lies across the green
 * direction from the locus point at `cct`: the residual r = neutral minus
 * locus (the locus point of gains wb is (-ln wb_R, -ln wb_B)), dotted with
 * (1, -1). Zero when r is along (1, 1), the direction a tint moves a
 * neutral. */
static double locus_across(const CRAWProfile *p, double lx, double ly, double cct, double r[2])
{
    double wb[3];

    wb_gains_at(p, cct, wb);
    r[0] = lx + log(wb[0]);
    r[1] = ly + log(wb[2]);
    return r[0] - r[1];
}

/* The locus point (lx, ly) lies from along the green direction
 * (lib/profile.c:285-365): the locus is sampled from the warm end to the
 * cool one at every knot and at most LOCUS_STEP_MIRED apart between them;
 * the first pair of samples the across-distance changes sign between is
 * bisected in mired (60 halvings), and a point beyond both ends goes to the
 * end nearer in that distance. The residual of the point found is left in
 * r. */
static void locus_green(const CRAWProfile *p, double lx, double ly, double *cct_out, double r[2])
{
    const uint32_t n = p->nb_wb_knots;
    double prev_cct = 0.0, prev_g = 0.0, next_cct = 0.0, ma, mb;
    int have_prev = 0, bracket = 0;

    for (uint32_t k = 0; k < n && !bracket; k++) {
        const double m0 = 1e6 / p->wb_knots[k].cct;
        const double span = k + 1 < n ? m0 - 1e6 / p->wb_knots[k + 1].cct : 0.0;
        int steps = k + 1 < n ? (int)ceil(span / LOCUS_STEP_MIRED) : 1;

        steps = steps > 0 ? steps : 1;
        for (int j = 0; j < steps; j++) {
            const double cct = 1e6 / (m0 - span * ((double)j / steps));
            const double g = locus_across(p, lx, ly, cct, r);

            if (g == 0.0) {
                *cct_out = cct;
                return;
            }
            if (have_prev && (g > 0.0) != (prev_g > 0.0)) {
                next_cct = cct;
                bracket = 1;
                break;
            }
            prev_cct = cct;
            prev_g = g;
            have_prev = 1;
        }
    }

    if (!bracket) {
        const double warm = p->wb_knots[0].cct, cool = p->wb_knots[n - 1].cct;
        double rw[2], rc[2];
        const int to_warm = fabs(locus_across(p, lx, ly, warm, rw)) <=
                            fabs(locus_across(p, lx, ly, cool, rc));

        *cct_out = to_warm ? warm : cool;
        r[0] = to_warm ? rw[0] : rc[0];
        r[1] = to_warm ? rw[1] : rc[1];
        return;
    }

    ma = 1e6 / prev_cct;
    mb = 1e6 / next_cct;
    for (int k = 0; k < 60; k++) {
        const double m = 0.5 * (ma + mb);
        const double g = locus_across(p, lx, ly, 1e6 / m, r);

        if (g == 0.0) {
            ma = mb = m;
            break;
        }
        if ((g > 0.0) == (prev_g > 0.0)) {
            ma = m;
            prev_g = g;
        } else {
            mb = m;
        }
    }
    *cct_out = 1e6 / (0.5 * (ma + mb));
    locus_across(p, lx, ly, *cct_out, r);
}

int craw_profile_neutral_to_illuminant(const CRAWProfile *p, float raw_rg, float raw_bg,
                                       CRAWIlluminant *out)
{
    double cct, r[2];

    if (!out || !p)
        return CRAW_ERROR(EINVAL);
    if (!(raw_rg > 0.0f) || !(raw_bg > 0.0f) || !isfinite(raw_rg) || !isfinite(raw_bg))
        return CRAW_ERROR(EINVAL);
    if (p->nb_wb_knots < 2)
        return CRAW_ERROR(ENOTSUP);

    locus_green(p, log((double)raw_rg), log((double)raw_bg), &cct, r);
    out->cct = (float)cct;
    /* the residual along (1, 1) / sqrt(2), as a tint (its axis is
     * (-1, -1) / sqrt(2)) */
    out->tint = (float)(-(r[0] + r[1]) * INV_SQRT2);
    return 0;
}

int craw_profile_illuminant_to_neutral(const CRAWProfile *p, const CRAWIlluminant *illum,
                                       float *raw_rg, float *raw_bg)
{
    double wb[3], cct, g;

    if (!p || !illum || !raw_rg || !raw_bg)
        return CRAW_ERROR(EINVAL);
    if (!p->nb_wb_knots)
        return CRAW_ERROR(ENOTSUP);
    if (!isfinite(illum->cct) || !isfinite(illum->tint))
        return CRAW_ERROR(EINVAL);

    cct = clampd((double)illum->cct, (double)p->wb_knots[0].cct,
                 (double)p->wb_knots[p->nb_wb_knots - 1].cct);
    wb_gains_at(p, cct, wb);
    /* what the gains of craw_profile_eval() for this illuminant undo */
    g = exp(-(double)illum->tint * INV_SQRT2);
    *raw_rg = (float)(g / wb[0]);
    *raw_bg = (float)(g / wb[2]);
    return 0;
}

/* Power laws in the gain between knots (log-log lines), the gain's own
 * scaling outside; lib/noise.c:37-110. */
static float loglin(float a0, float a1, float t)
{
    if (a0 > 0.0f && a1 > 0.0f)
        return expf(logf(a0) + t * (logf(a1) - logf(a0)));
    return a0 + t * (a1 - a0);
}

int craw_noise_model_eval(const CRAWNoiseKnot *k, uint32_t n, float iso, CRAWNoiseModel *out)
{
    uint32_t i;
    float t;

    if (!k || !n || !out)
        return CRAW_ERROR(EINVAL);
    if (!(iso > 0.0f) || !isfinite(iso)) {
        out->a = k[0].a;
        out->b = k[0].b;
        return 0;
    }
    if (iso <= k[0].iso) {
        out->a = k[0].a * (iso / k[0].iso);
        out->b = k[0].b;
        return 0;
    }
    if (iso >= k[n - 1].iso) {
        float g = iso / k[n - 1].iso;
        out->a = k[n - 1].a * g;
        out->b = k[n - 1].b * g * g;
        return 0;
    }
    for (i = 0; i + 1 < n && !(iso <= k[i + 1].iso); i++)
        ;
    if (!(k[i + 1].iso > k[i].iso)) {
        out->a = k[i].a;
        out->b = k[i].b;
        return 0;
    }
    t = (logf(iso) - logf(k[i].iso)) / (logf(k[i + 1].iso) - logf(k[i].iso));
    out->a = loglin(k[i].a, k[i + 1].a, t);
    out->b = loglin(k[i].b, k[i + 1].b, t);
    return 0;
}

int craw_profile_noise_eval(const CRAWProfile *p, float iso, CRAWNoiseModel *out)
{
    if (!p || !out)
        return CRAW_ERROR(EINVAL);
    if (!p->nb_noise_knots)
        return CRAW_ERROR(ENOTSUP);
    return craw_noise_model_eval(p->noise_knots, p->nb_noise_knots,
                                 iso > 0.0f ? iso : p->as_shot_ei, out);
}

/* ------------------------------------------------------- colour spaces */

typedef double mat3[3][3];

static const double cs_prim[CRAW_CS_NB][3][2] = {
    [CRAW_CS_BT709_LINEAR]    = { { 0.64, 0.33 }, { 0.30, 0.60 }, { 0.15, 0.06 } },
    [CRAW_CS_BT2020_LINEAR]   = { { 0.708, 0.292 }, { 0.170, 0.797 }, { 0.131, 0.046 } },
    [CRAW_CS_P3D65_LINEAR]    = { { 0.680, 0.320 }, { 0.265, 0.690 }, { 0.150, 0.060 } },
    [CRAW_CS_V_GAMUT_LINEAR]  = { { 0.730, 0.280 }, { 0.165, 0.840 }, { 0.100, -0.030 } },
    [CRAW_CS_ACES_AP0_LINEAR] = { { 0.7347, 0.2653 }, { 0.0, 1.0 }, { 0.0001, -0.0770 } },
    [CRAW_CS_ACES_AP1_LINEAR] = { { 0.713, 0.293 }, { 0.165, 0.830 }, { 0.128, 0.044 } },
};
static const double white_d65[2] = { 0.3127, 0.3290 };
static const double white_aces[2] = { 0.32168, 0.33767 };

static void xy_to_xyz(const double xy[2], double out[3])
{
    out[0] = xy[0] / xy[1];
    out[1] = 1.0;
    out[2] = (1.0 - xy[0] - xy[1]) / xy[1];
}

static void mat_mul(const mat3 a, const mat3 b, mat3 out)
{
    mat3 t;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            t[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
    memcpy(out, t, sizeof(t));
}

static void mat_vec(const mat3 m, const double v[3], double out[3])
{
    double t[3];
    for (int i = 0; i < 3; i++)
        t[i] = m[i][0] * v[0] + m[i][1] * v[1] + m[i][2] * v[2];
    memcpy(out, t, sizeof(t));
}

/* Inverse by cofactors over the determinant. */
static int mat_inv(const mat3 m, mat3 out)
{
    double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    double c01 = m[1][0] * m[2][2] - m[1][2] * m[2][0];
    double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    double det = m[0][0] * c00 - m[0][1] * c01 + m[0][2] * c02;
    mat3 t;

    if (!(fabs(det) > 1e-12))
        return CRAW_ERROR(EINVAL);
    t[0][0] = c00 / det;
    t[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
    t[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
    t[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
    t[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
    t[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
    t[2][0] = c02 / det;
    t[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
    t[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
    memcpy(out, t, sizeof(t));
    return 0;
}

static int bradford(const double src[2], const double dst[2], mat3 out)
{
    static const mat3 cone = {
        {  0.8951,  0.2664, -0.1614 },
        { -0.7502,  1.7135,  0.0367 },
        {  0.0389, -0.0685,  1.0296 },
    };
    mat3 cone_inv, gain = { { 0 } }, t;
    double ws[3], wd[3], ls[3], ld[3];

    xy_to_xyz(src, ws);
    xy_to_xyz(dst, wd);
    mat_vec(cone, ws, ls);
    mat_vec(cone, wd, ld);
    for (int i = 0; i < 3; i++) {
        if (!(fabs(ls[i]) > 1e-12))
            return CRAW_ERROR(EINVAL);
        gain[i][i] = ld[i] / ls[i];
    }
    if (mat_inv(cone, cone_inv) < 0)
        return CRAW_ERROR(EINVAL);
    mat_mul(gain, cone, t);
    mat_mul(cone_inv, t, out);
    return 0;
}

int craw_colorspace_to_xyz_d(enum CRAWColorSpace cs, double out[3][3])
{
    mat3 p, pinv, m;
    double w[3], s[3];
    bool aces;

    if ((unsigned)cs >= CRAW_CS_NB)
        return CRAW_ERROR(EINVAL);
    if (cs == CRAW_CS_XYZ_D65) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                out[i][j] = i == j;
        return 0;
    }
    aces = cs == CRAW_CS_ACES_AP0_LINEAR || cs == CRAW_CS_ACES_AP1_LINEAR;

    /* Columns along the primaries, scaled so that they sum to the white. */
    for (int c = 0; c < 3; c++) {
        double col[3];
        xy_to_xyz(cs_prim[cs][c], col);
        for (int i = 0; i < 3; i++)
            p[i][c] = col[i];
    }
    xy_to_xyz(aces ? white_aces : white_d65, w);
    if (mat_inv(p, pinv) < 0)
        return CRAW_ERROR(EINVAL);
    mat_vec(pinv, w, s);
    for (int i = 0; i < 3; i++)
        for (int c = 0; c < 3; c++)
            m[i][c] = p[i][c] * s[c];

    if (aces) {
        mat3 cat;
        if (bradford(white_aces, white_d65, cat) < 0)
            return CRAW_ERROR(EINVAL);
        mat_mul(cat, m, m);
    }
    memcpy(out, m, sizeof(m));
    return 0;
}

static void store_f(const mat3 in, float out[3][3])
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            out[i][j] = (float)in[i][j];
}

int craw_colorspace_from_xyz_d65(enum CRAWColorSpace cs, float m[3][3])
{
    mat3 to, from;

    if (!m || craw_colorspace_to_xyz_d(cs, to) < 0 || mat_inv(to, from) < 0)
        return CRAW_ERROR(EINVAL);
    store_f(from, m);
    return 0;
}

int craw_colorspace_to_xyz_d65(enum CRAWColorSpace cs, float m[3][3])
{
    mat3 to;

    if (!m || craw_colorspace_to_xyz_d(cs, to) < 0)
        return CRAW_ERROR(EINVAL);
    store_f(to, m);
    return 0;
}

int craw_planckian_xy(float cct, float xy[2])
{
    double t, u, v, d;

    if (!xy || !isfinite(cct))
        return CRAW_ERROR(EINVAL);
    t = fmin(fmax((double)cct, 1000.0), 15000.0);
    /* Krystek 1985, Color Research & Application 10(1). */
    u = (0.860117757 + 1.54118254e-4 * t + 1.28641212e-7 * t * t) /
        (1.0 + 8.42420235e-4 * t + 7.08145163e-7 * t * t);
    v = (0.317398726 + 4.22806245e-5 * t + 4.20481691e-8 * t * t) /
        (1.0 - 2.89741816e-5 * t + 1.61456053e-7 * t * t);
    d = 2.0 * u - 8.0 * v + 4.0;
    xy[0] = (float)(3.0 * u / d);
    xy[1] = (float)(2.0 * v / d);
    return 0;
}

int craw_bradford_adaptation(const float src[2], const float dst[2], float m[3][3])
{
    double s[2], d[2];
    mat3 cat;

    if (!src || !dst || !m)
        return CRAW_ERROR(EINVAL);
    s[0] = src[0]; s[1] = src[1];
    d[0] = dst[0]; d[1] = dst[1];
    if (!(s[1] > 0.0) || !(d[1] > 0.0) || bradford(s, d, cat) < 0)
        return CRAW_ERROR(EINVAL);
    store_f(cat, m);
    return 0;
}

/* ------------------------------------------------------ gamut compress */

void craw_gamut_params(CRAWGamutParams *g)
{
    /* ACES 1.3 RGC: distance limits and thresholds for cyan, magenta,
     * yellow, and the curve's power. */
    static const double lim[3] = { 1.147, 1.264, 1.312 };
    static const double thr[3] = { 0.815, 0.803, 0.880 };
    const double pwr = 1.2;

    for (int i = 0; i < 3; i++) {
        /* The scale that sends a distance of lim to exactly 1. */
        double s = (lim[i] - thr[i]) /
                   pow(pow((1.0 - thr[i]) / (lim[i] - thr[i]), -pwr) - 1.0, 1.0 / pwr);
        g->thr[i] = (float)thr[i];
        g->scl[i] = (float)s;
    }
    g->pwr = (float)pwr;
    g->inv_pwr = (float)(1.0 / pwr);
}

/* ---------------------------------------------------------- histogram */

int craw_awb_histogram_from_rgb(const float *rgb, uint32_t width, uint32_t height,
                                size_t stride, int channels, const float wb[3],
                                uint32_t *count, uint32_t *weight, uint32_t *bright)
{
    const uint32_t bins = CRAW_AWB_HIST_BINS;

    if (!rgb || !wb || !count || !weight || !bright)
        return CRAW_ERROR(EINVAL);
    if (channels != 3 && channels != 4)
        return CRAW_ERROR(EINVAL);
    if (stride < (size_t)width * (size_t)channels)
        return CRAW_ERROR(EINVAL);
    if (!(wb[0] > 0.0f) || !(wb[2] > 0.0f))
        return CRAW_ERROR(EINVAL);

    memset(count, 0, (size_t)bins * bins * sizeof(*count));
    memset(weight, 0, (size_t)bins * bins * sizeof(*weight));
    memset(bright, 0, (size_t)bins * bins * sizeof(*bright));

    for (uint32_t by = 0; by < height; by += 8) {
        for (uint32_t bx = 0; bx < width; bx += 8) {
            CRAWAwbSample s;
            float acc[3] = { 0.0f, 0.0f, 0.0f };
            uint32_t y1 = by + 8 < height ? by + 8 : height;
            uint32_t x1 = bx + 8 < width ? bx + 8 : width;

            for (uint32_t y = by; y < y1; y++) {
                const float *row = rgb + (size_t)y * stride;
                for (uint32_t x = bx; x < x1; x++) {
                    acc[0] += row[(size_t)x * channels + 0];
                    acc[1] += row[(size_t)x * channels + 1];
                    acc[2] += row[(size_t)x * channels + 2];
                }
            }
            if (!craw_awb_sample(acc, (float)((y1 - by) * (x1 - bx)), wb[0], wb[2], &s))
                continue;
            count[s.bin] += 1;
            weight[s.bin] += s.weight;
            bright[s.bin] += s.bright;
        }
    }
    return 0;
}
