/*
 * The oracle's host math: the per-frame constants of the path, restated.
 * TEST INFRASTRUCTURE, see oracle.h. Each function is pinned bit for bit
 * against the reference's own sources built under oracle/_ref.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

#define O_INV_SQRT2 0.70710678118654752440

/* profile.c:26-40 of the reference: identity profile. */
void oracle_profile_init(CRAWProfile *p)
{
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

/* profile.c:133-145: interpolation factor inside [a, b], clamped. */
static double o_frac(double a, double b, double cct)
{
    double f = 0.0;

    if (b > a)
        f = (cct - a) / (b - a);
    return f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
}

/* profile.c:147-166: gains along the wb table, linear in Kelvin. */
static void o_wb_at(const CRAWProfile *p, double cct, double wb[3])
{
    uint32_t i = 0;
    double f;

    if (p->nb_wb_knots == 1) {
        for (int c = 0; c < 3; c++)
            wb[c] = p->wb_knots[0].wb[c];
        return;
    }
    while (i + 2 < p->nb_wb_knots && cct >= p->wb_knots[i + 1].cct)
        i++;
    f = o_frac(p->wb_knots[i].cct, p->wb_knots[i + 1].cct, cct);
    for (int c = 0; c < 3; c++) {
        double a = p->wb_knots[i].wb[c], b = p->wb_knots[i + 1].wb[c];
        wb[c] = a + (b - a) * f;
    }
}

/* profile.c:168-190: matrix along the matrix table. */
static void o_matrix_at(const CRAWProfile *p, double cct, float m[3][3])
{
    uint32_t i = 0;
    double f;

    if (p->nb_matrix_knots == 1) {
        memcpy(m, p->matrix_knots[0].m, sizeof(float) * 9);
        return;
    }
    while (i + 2 < p->nb_matrix_knots && cct >= p->matrix_knots[i + 1].cct)
        i++;
    f = o_frac(p->matrix_knots[i].cct, p->matrix_knots[i + 1].cct, cct);
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++) {
            double a = p->matrix_knots[i].m[r][c], b = p->matrix_knots[i + 1].m[r][c];
            m[r][c] = (float)(a + (b - a) * f);
        }
    }
}

static double o_clampd(double v, double lo, double hi)
{
    v = v > lo ? v : lo;
    return v < hi ? v : hi;
}

/* profile.c:198-252: as shot, or re-balanced along the tables. */
int oracle_profile_eval(const CRAWProfile *p, enum CRAWWbMode mode,
                        const CRAWIlluminant *illum, CRAWColorParams *out)
{
    if (!p || !out)
        return -EINVAL;

    if (mode == CRAW_WB_AS_SHOT) {
        memcpy(out->wb, p->as_shot_wb, sizeof(out->wb));
        memcpy(out->cam_to_xyz, p->as_shot_matrix, sizeof(out->cam_to_xyz));
    } else if (mode == CRAW_WB_REBALANCE) {
        double wb[3], cct, tg;

        if (!illum)
            return -EINVAL;
        if (!p->nb_wb_knots)
            return -ENOTSUP;
        if (!isfinite(illum->cct) || !isfinite(illum->tint))
            return -EINVAL;
        cct = o_clampd((double)illum->cct, (double)p->wb_knots[0].cct,
                       (double)p->wb_knots[p->nb_wb_knots - 1].cct);
        o_wb_at(p, cct, wb);
        tg = exp((double)illum->tint * O_INV_SQRT2);
        wb[0] *= tg;
        wb[2] *= tg;
        for (int c = 0; c < 3; c++)
            out->wb[c] = (float)wb[c];
        if (p->nb_matrix_knots) {
            uint32_t nb = p->nb_matrix_knots;
            double mcct = o_clampd(cct, (double)p->matrix_knots[0].cct,
                                   (double)p->matrix_knots[nb - 1].cct);
            o_matrix_at(p, mcct, out->cam_to_xyz);
        } else {
            memcpy(out->cam_to_xyz, p->as_shot_matrix, sizeof(out->cam_to_xyz));
        }
    } else {
        return -EINVAL;
    }

    out->gain = p->gain;
    if (p->as_shot_ei > 0.0f && p->exposure_index > 0.0f)
        out->gain *= p->exposure_index / p->as_shot_ei;
    return 0;
}

/* noise.c:37-110: log-log between knots, gain law outside. */
int oracle_noise_model_eval(const CRAWNoiseKnot *k, uint32_t n, float iso,
                            CRAWNoiseModel *out)
{
    uint32_t i;
    float t, g;

    if (!k || !n || !out)
        return -EINVAL;
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
        g = iso / k[n - 1].iso;
        out->a = k[n - 1].a * g;
        out->b = k[n - 1].b * g * g;
        return 0;
    }
    for (i = 0; i + 1 < n; i++) {
        if (iso <= k[i + 1].iso)
            break;
    }
    if (!(k[i + 1].iso > k[i].iso)) {
        out->a = k[i].a;
        out->b = k[i].b;
        return 0;
    }
    t = (logf(iso) - logf(k[i].iso)) / (logf(k[i + 1].iso) - logf(k[i].iso));
    if (k[i].a > 0.0f && k[i + 1].a > 0.0f)
        out->a = expf(logf(k[i].a) + t * (logf(k[i + 1].a) - logf(k[i].a)));
    else
        out->a = k[i].a + t * (k[i + 1].a - k[i].a);
    if (k[i].b > 0.0f && k[i + 1].b > 0.0f)
        out->b = expf(logf(k[i].b) + t * (logf(k[i + 1].b) - logf(k[i].b)));
    else
        out->b = k[i].b + t * (k[i + 1].b - k[i].b);
    return 0;
}

/* noise.c:112-123 */
int oracle_profile_noise_eval(const CRAWProfile *p, float iso, CRAWNoiseModel *out)
{
    if (!p || !out)
        return -EINVAL;
    if (!p->nb_noise_knots)
        return -ENOTSUP;
    if (!(iso > 0.0f))
        iso = p->as_shot_ei;
    return oracle_noise_model_eval(p->noise_knots, p->nb_noise_knots, iso, out);
}

/* ref_color.c:78-89: the override when a > 0, else the table. */
int oracle_noise_model(const CRAWFrameParams *fp, CRAWNoiseModel *out)
{
    if (!fp || !fp->profile || !out)
        return -EINVAL;
    if (fp->noise_override.a > 0.0f) {
        *out = fp->noise_override;
        return 0;
    }
    return oracle_profile_noise_eval(fp->profile, fp->profile->as_shot_ei, out);
}

/* colorspace.c:25-63: primaries per space; the ACES pair against the ACES
 * white. */
static const double o_prim[CRAW_CS_NB][3][2] = {
    [CRAW_CS_BT709_LINEAR]    = { { 0.64, 0.33 }, { 0.30, 0.60 }, { 0.15, 0.06 } },
    [CRAW_CS_BT2020_LINEAR]   = { { 0.708, 0.292 }, { 0.170, 0.797 }, { 0.131, 0.046 } },
    [CRAW_CS_P3D65_LINEAR]    = { { 0.680, 0.320 }, { 0.265, 0.690 }, { 0.150, 0.060 } },
    [CRAW_CS_V_GAMUT_LINEAR]  = { { 0.730, 0.280 }, { 0.165, 0.840 }, { 0.100, -0.030 } },
    [CRAW_CS_ACES_AP0_LINEAR] = { { 0.7347, 0.2653 }, { 0.0, 1.0 }, { 0.0001, -0.0770 } },
    [CRAW_CS_ACES_AP1_LINEAR] = { { 0.713, 0.293 }, { 0.165, 0.830 }, { 0.128, 0.044 } },
};
static const double o_d65[2] = { 0.3127, 0.3290 };
static const double o_aces[2] = { 0.32168, 0.33767 };

static void o_xyz(double x, double y, double out[3])
{
    out[0] = x / y;
    out[1] = 1.0;
    out[2] = (1.0 - x - y) / y;
}

static void o_mul(const double a[3][3], const double b[3][3], double out[3][3])
{
    double t[3][3];

    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            t[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
    memcpy(out, t, sizeof(t));
}

static void o_mulv(const double m[3][3], const double v[3], double out[3])
{
    double t[3];

    for (int i = 0; i < 3; i++)
        t[i] = m[i][0] * v[0] + m[i][1] * v[1] + m[i][2] * v[2];
    memcpy(out, t, sizeof(t));
}

/* colorspace.c:89-116: cofactor inverse. */
static int o_inv(const double m[3][3], double out[3][3])
{
    double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1])
               - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
               + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    double t[3][3];

    if (!(fabs(det) > 1e-12))
        return -EINVAL;
    t[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
    t[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
    t[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
    t[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
    t[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
    t[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
    t[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
    t[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
    t[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
    memcpy(out, t, sizeof(t));
    return 0;
}

/* colorspace.c:118-145: columns from the primaries, scaled to the white. */
static int o_rgb_to_xyz(const double prim[3][2], const double white[2], double out[3][3])
{
    double p[3][3], inv[3][3], w[3], s[3];

    for (int c = 0; c < 3; c++) {
        double col[3];
        o_xyz(prim[c][0], prim[c][1], col);
        for (int i = 0; i < 3; i++)
            p[i][c] = col[i];
    }
    o_xyz(white[0], white[1], w);
    if (o_inv(p, inv) < 0)
        return -EINVAL;
    o_mulv(inv, w, s);
    for (int i = 0; i < 3; i++)
        for (int c = 0; c < 3; c++)
            out[i][c] = p[i][c] * s[c];
    return 0;
}

/* colorspace.c:147-180: von Kries in the Bradford cone space. */
static int o_bradford(const double src[2], const double dst[2], double out[3][3])
{
    static const double bfd[3][3] = {
        {  0.8951,  0.2664, -0.1614 },
        { -0.7502,  1.7135,  0.0367 },
        {  0.0389, -0.0685,  1.0296 },
    };
    double bi[3][3], sc[3][3] = { { 0 } }, tmp[3][3], ws[3], wd[3], cs[3], cd[3];

    o_xyz(src[0], src[1], ws);
    o_xyz(dst[0], dst[1], wd);
    o_mulv(bfd, ws, cs);
    o_mulv(bfd, wd, cd);
    for (int i = 0; i < 3; i++) {
        if (!(fabs(cs[i]) > 1e-12))
            return -EINVAL;
        sc[i][i] = cd[i] / cs[i];
    }
    if (o_inv(bfd, bi) < 0)
        return -EINVAL;
    o_mul(sc, bfd, tmp);
    o_mul(bi, tmp, out);
    return 0;
}

/* colorspace.c:185-220: RGB to XYZ D65 in double. */
int oracle_colorspace_to_xyz_d(enum CRAWColorSpace cs, double out[3][3])
{
    double m[3][3];
    bool aces;

    if ((unsigned)cs >= CRAW_CS_NB)
        return -EINVAL;
    if (cs == CRAW_CS_XYZ_D65) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                out[i][j] = i == j ? 1.0 : 0.0;
        return 0;
    }
    aces = cs == CRAW_CS_ACES_AP0_LINEAR || cs == CRAW_CS_ACES_AP1_LINEAR;
    if (o_rgb_to_xyz(o_prim[cs], aces ? o_aces : o_d65, m) < 0)
        return -EINVAL;
    if (aces) {
        double cat[3][3];
        if (o_bradford(o_aces, o_d65, cat) < 0)
            return -EINVAL;
        o_mul(cat, m, m);
    }
    memcpy(out, m, sizeof(m));
    return 0;
}

int oracle_colorspace_from_xyz_d(enum CRAWColorSpace cs, double out[3][3])
{
    double m[3][3];

    if (oracle_colorspace_to_xyz_d(cs, m) < 0)
        return -EINVAL;
    return o_inv(m, out);
}

int oracle_colorspace_from_xyz_d65(enum CRAWColorSpace cs, float m[3][3])
{
    double d[3][3];

    if (!m || oracle_colorspace_from_xyz_d(cs, d) < 0)
        return -EINVAL;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            m[i][j] = (float)d[i][j];
    return 0;
}

int oracle_colorspace_to_xyz_d65(enum CRAWColorSpace cs, float m[3][3])
{
    double d[3][3];

    if (!m || oracle_colorspace_to_xyz_d(cs, d) < 0)
        return -EINVAL;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            m[i][j] = (float)d[i][j];
    return 0;
}

/* colorspace.c:268-290: Krystek's rational approximation in CIE 1960 uv. */
int oracle_planckian_xy(float cct, float xy[2])
{
    double t, u, v, d;

    if (!xy || !isfinite(cct))
        return -EINVAL;
    t = fmin(fmax((double)cct, 1000.0), 15000.0);
    u = (0.860117757 + 1.54118254e-4 * t + 1.28641212e-7 * t * t) /
        (1.0 + 8.42420235e-4 * t + 7.08145163e-7 * t * t);
    v = (0.317398726 + 4.22806245e-5 * t + 4.20481691e-8 * t * t) /
        (1.0 - 2.89741816e-5 * t + 1.61456053e-7 * t * t);
    d = 2.0 * u - 8.0 * v + 4.0;
    xy[0] = (float)(3.0 * u / d);
    xy[1] = (float)(2.0 * v / d);
    return 0;
}

int oracle_bradford_adaptation(const float src[2], const float dst[2], float m[3][3])
{
    double s[2], d[2], cat[3][3];

    if (!src || !dst || !m)
        return -EINVAL;
    for (int i = 0; i < 2; i++) {
        s[i] = src[i];
        d[i] = dst[i];
    }
    if (!(s[1] > 0.0) || !(d[1] > 0.0))
        return -EINVAL;
    if (o_bradford(s, d, cat) < 0)
        return -EINVAL;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            m[i][j] = (float)cat[i][j];
    return 0;
}

/* gamut.c:20-64: ACES 1.3 RGC limits and thresholds; the scale solved so
 * that a distance of `lim` compresses to exactly 1. */
static const double o_lim[3] = { 1.147, 1.264, 1.312 };
static const double o_thr[3] = { 0.815, 0.803, 0.880 };
#define O_PWR 1.2

static double o_rgc_scale(double lim, double thr, double pwr)
{
    return (lim - thr) / pow(pow((1.0 - thr) / (lim - thr), -pwr) - 1.0, 1.0 / pwr);
}

void oracle_gamut_params(float thr[3], float scl[3], float *pwr, float *inv_pwr)
{
    for (int i = 0; i < 3; i++) {
        thr[i] = (float)o_thr[i];
        scl[i] = (float)o_rgc_scale(o_lim[i], o_thr[i], O_PWR);
    }
    *pwr = (float)O_PWR;
    *inv_pwr = (float)(1.0 / O_PWR);
}

void oracle_gamut_params_d(double thr[3], double scl[3], double *pwr, double *inv_pwr)
{
    for (int i = 0; i < 3; i++) {
        thr[i] = o_thr[i];
        scl[i] = o_rgc_scale(o_lim[i], o_thr[i], O_PWR);
    }
    *pwr = O_PWR;
    *inv_pwr = 1.0 / O_PWR;
}
