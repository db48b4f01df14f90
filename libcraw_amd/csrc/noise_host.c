/*
 * libcraw for MI355X: the noise fit of a frame in host memory
 * (lib/noise.c:126-504 of cyanreg/libcraw), and the fit both paths share.
 *
 * Measurement: each CFA plane, normalised in double and stored as float,
 * is high-passed with [1 -2 1; -2 4 -2; 1 -2 1] / 6 (zero on any bilinear
 * signal, unit gain on white noise variance) and every sample is binned by
 * its 9x9 plane mean (summed in raster order), which is exactly
 * uncorrelated with that kernel, into bins equally spaced in sqrt(mean).
 * A bin's noise is the median |residual| over 0.6745. The bins' means are
 * summed in the frame's order (planes, rows, columns) in double, as the
 * reference sums them, so this path reproduces it bit for bit; a bin's
 * median is found by selection rather than by sorting it.
 *
 * Fit: least trimmed squares over every pair of bins seeds the line (the
 * pair whose two-thirds smallest squared log residuals sum least, a > 0,
 * b >= 0), then 40 steps at most of Gauss-Newton on ln(a u + b) with a and
 * b as logarithms, a Tukey biweight on the log residual (its scale the
 * residuals' MAD every eighth step, floored at 0.02) and bins weighted by
 * ln(1 + samples), damped and clamped to half a log unit a step.
 */
#include <math.h>
#include <pthread.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

#include "noise_fit.h"

#define MAD_SCALE 0.674489750196081746f   /* Phi^-1(0.75) */
#define TUKEY_C 4.685
#define SCALE_FLOOR 0.02

void craw_noise_estimate_params_init(CRAWNoiseEstimateParams *p)
{
    if (!p)
        return;
    memset(p, 0, sizeof(*p));
    p->nb_bins = 128;
    p->min_bin_samples = 20000;
    p->max_level = 0.98f;
}

int craw_noise_resolve(const CRAWNoiseEstimateParams *params, uint32_t *nb_bins,
                       uint32_t *min_samples, float *max_level)
{
    CRAWNoiseEstimateParams def;

    if (!params) {
        craw_noise_estimate_params_init(&def);
        params = &def;
    }
    *nb_bins = params->nb_bins ? params->nb_bins : 128;
    *min_samples = params->min_bin_samples ? params->min_bin_samples : 20000;
    *max_level = params->max_level > 0.0f ? params->max_level : 0.98f;
    if (*nb_bins < 8 || *nb_bins > CRAW_NOISE_MAX_BINS)
        return CRAW_ERROR(EINVAL);
    return 0;
}

int craw_noise_check(uint32_t width, uint32_t height, size_t stride, const CRAWProfile *profile)
{
    if (profile->cfa != CRAW_CFA_RGGB)
        return CRAW_ERROR(ENOTSUP);
    if ((width & 1) || (height & 1) || width < 32 || height < 32 || stride < width)
        return CRAW_ERROR(EINVAL);
    for (int s = 0; s < 4; s++)
        if (!(profile->white_level[s] > profile->black_level[s]))
            return CRAW_ERROR(EINVAL);
    if (width / 2 < 2 * CRAW_NOISE_MEAN_R + 1 || height / 2 < 2 * CRAW_NOISE_MEAN_R + 1)
        return CRAW_ERROR(EINVAL);
    return 0;
}

/* The element of rank k (0-based, ascending) of v[0..n), reordering v:
 * Hoare's selection with a median-of-three pivot. */
static double select_rank(double *v, size_t n, size_t k)
{
    size_t lo = 0, hi = n - 1;

    while (hi > lo) {
        const size_t mid = lo + (hi - lo) / 2;
        double a = v[lo], b = v[mid], c = v[hi], pivot, t;
        size_t i = lo, j = hi;

        pivot = a < b ? (b < c ? b : (a < c ? c : a)) : (a < c ? a : (b < c ? c : b));
        while (i <= j) {
            while (v[i] < pivot)
                i++;
            while (v[j] > pivot)
                j--;
            if (i <= j) {
                t = v[i];
                v[i] = v[j];
                v[j] = t;
                i++;
                if (j == 0)
                    break;
                j--;
            }
        }
        if (k <= j)
            hi = j;
        else if (k >= i)
            lo = i;
        else
            return v[k];
    }
    return v[k];
}

/* Least trimmed squares seed over every pair of bins (i < j): the line
 * through the two, kept when a > 0 and b >= 0, costed by the sum of the
 * smallest two thirds of the squared log residuals in ascending order; the
 * first pair of least cost in (i, j) order wins. The pairs' costs are
 * independent, so rows of i are costed on several threads into a table and
 * the table is then scanned in order: the same pair and the same bits as a
 * single loop. */
struct LtsJob {
    const double *m, *v, *lv;
    int nb, keep, first, step;
    double *cost;               /* nb x nb, NAN for a pair not kept */
};

static void lts_rows(const struct LtsJob *jb)
{
    const int nb = jb->nb;
    double *sq = malloc((size_t)nb * sizeof(*sq));

    for (int i = jb->first; i < nb; i += jb->step) {
        for (int j = i + 1; j < nb; j++) {
            double a, b, cost = 0.0;

            jb->cost[(size_t)i * nb + j] = NAN;
            if (!sq || jb->m[j] - jb->m[i] < 1e-5)
                continue;
            a = (jb->v[j] - jb->v[i]) / (jb->m[j] - jb->m[i]);
            b = jb->v[i] - a * jb->m[i];
            if (!(a > 0.0) || b < 0.0)
                continue;
            /* an insertion sort: any sort gives the same ascending array */
            for (int k = 0; k < nb; k++) {
                const double d = jb->lv[k] - log(a * jb->m[k] + b), x = d * d;
                int q = k;

                for (; q > 0 && sq[q - 1] > x; q--)
                    sq[q] = sq[q - 1];
                sq[q] = x;
            }
            for (int k = 0; k < jb->keep; k++)
                cost += sq[k];
            jb->cost[(size_t)i * nb + j] = cost;
        }
    }
    free(sq);
}

static void *lts_thread(void *arg)
{
    lts_rows((const struct LtsJob *)arg);
    return NULL;
}

#define LTS_THREADS 8

static int lts_seed(const double *m, const double *v, int nb, double *a_out, double *b_out,
                    double *scale)
{
    const int keep = (nb * 2) / 3 + 1;
    const int nt = nb >= 48 ? LTS_THREADS : 1;
    double best = INFINITY, ba = 0.0, bb = 0.0;
    double *lv = malloc((size_t)nb * sizeof(*lv));
    double *cost = malloc((size_t)nb * nb * sizeof(*cost));
    struct LtsJob jobs[LTS_THREADS];
    pthread_t th[LTS_THREADS];
    bool started[LTS_THREADS] = { false };

    if (!lv || !cost) {
        free(lv);
        free(cost);
        return 0;
    }
    /* ln v once (the same value every pair would compute) */
    for (int k = 0; k < nb; k++)
        lv[k] = log(v[k]);
    for (int t = 0; t < nt; t++) {
        jobs[t] = (struct LtsJob) { m, v, lv, nb, keep, t, nt, cost };
        if (t > 0)
            started[t] = pthread_create(&th[t], NULL, lts_thread, &jobs[t]) == 0;
    }
    lts_rows(&jobs[0]);
    for (int t = 1; t < nt; t++) {
        if (started[t])
            pthread_join(th[t], NULL);
        else
            lts_rows(&jobs[t]);
    }
    for (int i = 0; i < nb; i++) {
        for (int j = i + 1; j < nb; j++) {
            const double c = cost[(size_t)i * nb + j];

            if (c < best) {
                best = c;
                ba = (v[j] - v[i]) / (m[j] - m[i]);
                bb = v[i] - ba * m[i];
            }
        }
    }
    free(lv);
    free(cost);
    if (!(ba > 0.0))
        return 0;
    *a_out = ba;
    *b_out = bb;
    *scale = sqrt(best / keep);
    return 1;
}

/* Reweighted Gauss-Newton on ln(a m + b) in (ln a, ln b). */
static void refine(const double *m, const double *v, const double *cnt, int nb, double scale,
                   double *a_io, double *b_io)
{
    double la = log(*a_io), lb = log(*b_io > 1e-16 ? *b_io : 1e-16);
    double *res = malloc((size_t)nb * sizeof(*res)), *tmp = malloc((size_t)nb * sizeof(*tmp));

    if (!res || !tmp) {
        free(res);
        free(tmp);
        return;
    }
    scale = scale < SCALE_FLOOR ? SCALE_FLOOR : scale;

    for (int it = 0; it < 40; it++) {
        const double ea = exp(la), eb = exp(lb);
        double h_aa = 0.0, h_ab = 0.0, h_bb = 0.0, g_a = 0.0, g_b = 0.0;
        double det, da, db, cut;

        for (int k = 0; k < nb; k++)
            res[k] = log(v[k]) - log(ea * m[k] + eb);
        /* the residuals' robust scale, every eighth step */
        if ((it & 7) == 0) {
            for (int k = 0; k < nb; k++)
                tmp[k] = fabs(res[k]);
            scale = 1.4826 * select_rank(tmp, (size_t)nb, (size_t)(nb / 2));
            scale = scale < SCALE_FLOOR ? SCALE_FLOOR : scale;
        }
        cut = TUKEY_C * scale;

        for (int k = 0; k < nb; k++) {
            const double den = ea * m[k] + eb;
            const double ja = ea * m[k] / den, jb = eb / den;
            const double u = res[k] / cut;
            double w = fabs(u) < 1.0 ? (1.0 - u * u) * (1.0 - u * u) : 0.0;

            w *= log(1.0 + cnt[k]);
            h_aa += w * ja * ja;
            h_ab += w * ja * jb;
            h_bb += w * jb * jb;
            g_a += w * ja * res[k];
            g_b += w * jb * res[k];
        }
        /* damped: a frame without shadows leaves b undetermined */
        det = (h_aa + 1e-9) * (h_bb + 1e-9) - h_ab * h_ab;
        if (!(fabs(det) > 1e-30))
            break;
        da = ((h_bb + 1e-9) * g_a - h_ab * g_b) / det;
        db = (-h_ab * g_a + (h_aa + 1e-9) * g_b) / det;
        da = CRAW_CLAMP(da, -0.5, 0.5);
        db = CRAW_CLAMP(db, -0.5, 0.5);
        la += da;
        lb += db;
        lb = lb < -40.0 ? -40.0 : lb;
        if (fabs(da) < 1e-9 && fabs(db) < 1e-9)
            break;
    }
    free(res);
    free(tmp);
    *a_io = exp(la);
    *b_io = lb <= -40.0 ? 0.0 : exp(lb);
}

int craw_noise_fit(const CRAWNoiseBin *bins, uint32_t nb_bins, uint32_t min_samples,
                   const CRAWProfile *profile, CRAWNoiseEstimate *out)
{
    double *m = malloc(nb_bins * sizeof(*m)), *v = malloc(nb_bins * sizeof(*v));
    double *c = malloc(nb_bins * sizeof(*c));
    double a = 0.0, b = 0.0, scale = 0.0, rms = 0.0, range0;
    int nb = 0, nin = 0, ret = 0;

    if (!m || !v || !c) {
        ret = CRAW_ERROR(ENOMEM);
        goto done;
    }
    for (uint32_t i = 0; i < nb_bins; i++) {
        double sigma;

        if (bins[i].n < min_samples)
            continue;
        sigma = bins[i].median / MAD_SCALE;
        if (!(sigma > 0.0))
            continue;
        m[nb] = bins[i].mean_sum / (double)bins[i].n;
        v[nb] = sigma * sigma;
        c[nb] = (double)bins[i].n;
        nb++;
    }
    if (nb < 3 || !lts_seed(m, v, nb, &a, &b, &scale)) {
        ret = CRAW_ERROR(ENODATA);
        goto done;
    }
    refine(m, v, c, nb, scale, &a, &b);

    /* a bin the model misses by more than about 40% describes the scene */
    for (int k = 0; k < nb; k++) {
        const double d = log(v[k]) - log(a * m[k] + b);
        if (fabs(d) < 0.35) {
            rms += d * d;
            nin++;
        }
    }
    rms = sqrt(rms / (nin ? nin : 1));
    range0 = profile->white_level[0] - profile->black_level[0];

    memset(out, 0, sizeof(*out));
    out->model.a = (float)a;
    out->model.b = (float)b;
    out->nb_bins = (uint32_t)nb;
    out->nb_inliers = (uint32_t)nin;
    out->fit_rms = (float)rms;
    out->read_noise_lsb = (float)(sqrt(b) * range0 * 65535.0);
done:
    free(m);
    free(v);
    free(c);
    return ret;
}

/* ------------------------------------------------- host measurement */

struct HostBin {
    float *r;
    size_t n, cap;
    double mean_sum;
};

static int bin_push(struct HostBin *b, float r)
{
    if (b->n == b->cap) {
        const size_t cap = b->cap ? 2 * b->cap : 4096;
        float *p = realloc(b->r, cap * sizeof(*p));
        if (!p)
            return CRAW_ERROR(ENOMEM);
        b->r = p;
        b->cap = cap;
    }
    b->r[b->n++] = r;
    return 0;
}

/* The element of rank k of v[0..n) as the float it is (see select_rank). */
static float select_rank_f(float *v, size_t n, size_t k)
{
    size_t lo = 0, hi = n - 1;

    while (hi > lo) {
        const size_t mid = lo + (hi - lo) / 2;
        float a = v[lo], b = v[mid], c = v[hi], pivot, t;
        size_t i = lo, j = hi;

        pivot = a < b ? (b < c ? b : (a < c ? c : a)) : (a < c ? a : (b < c ? c : b));
        while (i <= j) {
            while (v[i] < pivot)
                i++;
            while (v[j] > pivot)
                j--;
            if (i <= j) {
                t = v[i];
                v[i] = v[j];
                v[j] = t;
                i++;
                if (j == 0)
                    break;
                j--;
            }
        }
        if (k <= j)
            hi = j;
        else if (k >= i)
            lo = i;
        else
            return v[k];
    }
    return v[k];
}

int craw_noise_estimate(const uint16_t *bayer, uint32_t width, uint32_t height, size_t stride,
                        const CRAWProfile *profile, const CRAWNoiseEstimateParams *params,
                        CRAWNoiseEstimate *out)
{
    const int R = CRAW_NOISE_MEAN_R;
    uint32_t nb_bins, min_samples;
    float max_level;
    struct HostBin *hb = NULL;
    CRAWNoiseBin *bins = NULL;
    float *plane = NULL;
    int w2, h2, ret;

    if (!bayer || !profile || !out)
        return CRAW_ERROR(EINVAL);
    if ((ret = craw_noise_check(width, height, stride, profile)) < 0)
        return ret;
    if ((ret = craw_noise_resolve(params, &nb_bins, &min_samples, &max_level)) < 0)
        return ret;

    w2 = (int)(width / 2);
    h2 = (int)(height / 2);
    plane = malloc((size_t)w2 * h2 * sizeof(*plane));
    hb = calloc(nb_bins, sizeof(*hb));
    bins = calloc(nb_bins, sizeof(*bins));
    if (!plane || !hb || !bins) {
        ret = CRAW_ERROR(ENOMEM);
        goto done;
    }

    for (int s = 0; s < 4; s++) {
        const int px = s & 1, py = s >> 1;
        const double black = profile->black_level[s];
        /* the range is formed in float, as the reference forms it */
        const double range = profile->white_level[s] - profile->black_level[s];

        for (int k = 0; k < h2; k++)
            for (int i = 0; i < w2; i++) {
                const double raw = bayer[(size_t)(2 * k + py) * stride + 2 * i + px] *
                                   (1.0 / 65535.0);
                plane[(size_t)k * w2 + i] = (float)((raw - black) / range);
            }

        for (int k = R; k < h2 - R; k++) {
            for (int i = R; i < w2 - R; i++) {
                const float *p = plane + (size_t)k * w2 + i;
                float r, mean = 0.0f;
                int bi;

                r = (p[-w2 - 1] + p[-w2 + 1] + p[w2 - 1] + p[w2 + 1] -
                     2.0f * (p[-w2] + p[w2] + p[-1] + p[1]) + 4.0f * p[0]) * (1.0f / 6.0f);
                for (int dy = -R; dy <= R; dy++)
                    for (int dx = -R; dx <= R; dx++)
                        mean += p[(ptrdiff_t)dy * w2 + dx];
                mean *= 1.0f / ((2 * R + 1) * (2 * R + 1));
                if (!(mean > 0.0f) || mean > max_level)
                    continue;
                bi = (int)(sqrtf(mean) * (float)nb_bins);
                bi = CRAW_CLAMP(bi, 0, (int)nb_bins - 1);
                if ((ret = bin_push(&hb[bi], fabsf(r))) < 0)
                    goto done;
                hb[bi].mean_sum += mean;
            }
        }
    }

    for (uint32_t i = 0; i < nb_bins; i++) {
        bins[i].n = hb[i].n;
        bins[i].mean_sum = hb[i].mean_sum;
        bins[i].median = hb[i].n ? select_rank_f(hb[i].r, hb[i].n, hb[i].n / 2) : 0.0f;
    }
    ret = craw_noise_fit(bins, nb_bins, min_samples, profile, out);

done:
    if (hb)
        for (uint32_t i = 0; i < nb_bins; i++)
            free(hb[i].r);
    free(hb);
    free(bins);
    free(plane);
    return ret;
}
