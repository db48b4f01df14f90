/*
 * libcraw for MI355X: the host's reading of the AWB histogram a recording
 * leaves (lib/awb.c:15-211 and lib/awb_smoother.c:21-172 of
 * cyanreg/libcraw), in double, as the reference keeps it.
 *
 * The estimator maps every occupied bin's centre, as a raw neutral, to the
 * locus point it lies from along the green direction; the bin's distance
 * from the locus sample at that Kelvin decides whether it is a neutral
 * candidate (inside `band`) and how much it counts (a Gaussian of width
 * `sigma`). Two weightings of the same bins are read: the blocks' clipped
 * green (zone) and its fourth power (bright, which favours light sources
 * and glare). Each gives the weighted mean mired and tint of its
 * candidates, their spread along the locus and the share of its mass
 * inside the band; the bright reading is preferred when it is valid, its
 * confidence the product of the in-band share (saturating at 0.3) and how
 * tight the spread is (zero at 60 mired).
 *
 * The smoother keeps the track in mired: an exponential approach with a
 * time constant, a hysteresis below which nothing moves, and a slew limit
 * that scales the step's fraction so the tint moves by the same share.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <craw/awb.h>

/* In-band share at which a reading is fully trusted, and the spread along
 * the locus at which it is not trusted at all. */
#define CONF_FULL_ZONE_FRACTION 0.3
#define CONF_ZERO_SPREAD_MIRED 60.0

void craw_awb_params_init(CRAWAwbParams *p)
{
    if (!p)
        return;
    memset(p, 0, sizeof(*p));
    p->band = 0.08f;
    p->sigma = 0.05f;
    p->min_zone_fraction = 0.02f;
    p->ambience = 0.0f;
    p->ambience_knee_cct = 4000.0f;
    p->cct_min = 0.0f;
    p->cct_max = 0.0f;
    p->prefer_bright = true;
    p->bright_max_spread_mired = 60.0f;
}

/* Sums of one weighting: its mass over all bins and inside the band, and
 * the falloff-weighted moments of the candidates' mired and tint. */
struct Reading {
    double mass, in_band;
    double w, w_mired, w_mired2, w_tint;
};

static void reading_add(struct Reading *r, double mass, bool in_band, double falloff,
                        double mired, double tint)
{
    double w;

    r->mass += mass;
    if (!in_band)
        return;
    r->in_band += mass;
    w = mass * falloff;
    if (!(w > 0.0))
        return;
    r->w += w;
    r->w_mired += w * mired;
    r->w_mired2 += w * mired * mired;
    r->w_tint += w * tint;
}

static void reading_result(const struct Reading *r, double min_zone_fraction,
                           double max_spread, CRAWAwbCandidate *c)
{
    double share, mean, var, spread;

    memset(c, 0, sizeof(*c));
    if (!(r->mass > 0.0) || !(r->w > 0.0))
        return;
    share = r->in_band / r->mass;
    mean = r->w_mired / r->w;
    /* the moments' variance can round a hair below zero when every
     * candidate sits in one bin */
    var = r->w_mired2 / r->w - mean * mean;
    spread = var > 0.0 ? sqrt(var) : 0.0;

    c->valid = share >= min_zone_fraction && (max_spread <= 0.0 || spread <= max_spread);
    c->illuminant.cct = (float)(1e6 / mean);
    c->illuminant.tint = (float)(r->w_tint / r->w);
    c->in_zone_fraction = (float)share;
    c->spread_mired = (float)spread;
}

int craw_awb_estimate(const CRAWProfile *p, const CRAWAwbHistogram *hist,
                      const CRAWAwbParams *params, CRAWAwbEstimate *out)
{
    CRAWAwbParams defaults;
    struct Reading zone, bright;
    const CRAWAwbCandidate *pick;
    double span, cct, conf, lo, hi;
    const uint32_t nb = hist ? hist->bins : 0;

    if (!p || !hist || !out)
        return CRAW_ERROR(EINVAL);
    if (!hist->count || !hist->weight || nb < 2 || !(hist->log_max > hist->log_min))
        return CRAW_ERROR(EINVAL);
    if (p->nb_wb_knots < 2)
        return CRAW_ERROR(ENOTSUP);
    if (!params) {
        craw_awb_params_init(&defaults);
        params = &defaults;
    }
    if (!(params->band > 0.0f) || !(params->sigma > 0.0f))
        return CRAW_ERROR(EINVAL);

    memset(out, 0, sizeof(*out));
    out->source = CRAW_AWB_SOURCE_NONE;
    memset(&zone, 0, sizeof(zone));
    memset(&bright, 0, sizeof(bright));
    span = (double)hist->log_max - hist->log_min;

    for (uint32_t iy = 0; iy < nb; iy++) {
        const double ly = hist->log_min + span * (iy + 0.5) / nb;

        for (uint32_t ix = 0; ix < nb; ix++) {
            const uint32_t bin = iy * nb + ix;
            CRAWIlluminant il;
            CRAWIlluminant on_locus;
            float rg, bg;
            double lx, ex, ey, d, q, falloff, mired;
            bool in_band;
            int ret;

            if (!hist->count[bin])
                continue;
            lx = hist->log_min + span * (ix + 0.5) / nb;

            /* the bin centre as a raw neutral, its Kelvin on the locus, and
             * its distance from the locus sample at that Kelvin */
            ret = craw_profile_neutral_to_illuminant(p, (float)exp(lx), (float)exp(ly), &il);
            if (ret < 0)
                return ret;
            on_locus.cct = il.cct;
            on_locus.tint = 0.0f;
            ret = craw_profile_illuminant_to_neutral(p, &on_locus, &rg, &bg);
            if (ret < 0)
                return ret;
            ex = lx - log((double)rg);
            ey = ly - log((double)bg);
            d = sqrt(ex * ex + ey * ey);

            in_band = d < params->band;
            q = d / params->sigma;
            falloff = exp(-q * q);
            mired = 1e6 / il.cct;

            reading_add(&zone, hist->weight[bin] / 65536.0, in_band, falloff, mired, il.tint);
            if (hist->weight_bright)
                reading_add(&bright, hist->weight_bright[bin] / 65536.0, in_band, falloff, mired,
                            il.tint);
        }
    }

    reading_result(&zone, params->min_zone_fraction, 0.0, &out->zone);
    reading_result(&bright, params->min_zone_fraction, params->bright_max_spread_mired,
                   &out->bright);

    /* a light source in frame is the illuminant itself; a surface only a
     * guess at it */
    if (params->prefer_bright && out->bright.valid) {
        pick = &out->bright;
        out->source = CRAW_AWB_SOURCE_BRIGHT;
    } else if (out->zone.in_zone_fraction > 0.0f || out->zone.valid) {
        pick = &out->zone;
        out->source = CRAW_AWB_SOURCE_ZONE;
    } else {
        return 0;
    }
    if (!(pick->illuminant.cct > 0.0f))
        return 0;

    conf = CRAW_CLAMP(pick->in_zone_fraction / CONF_FULL_ZONE_FRACTION, 0.0, 1.0) *
           CRAW_CLAMP(1.0 - pick->spread_mired / CONF_ZERO_SPREAD_MIRED, 0.0, 1.0);

    /* ambience keeps some of a warm light's warmth */
    cct = pick->illuminant.cct;
    if (params->ambience > 0.0f && cct < params->ambience_knee_cct)
        cct += (double)params->ambience * (params->ambience_knee_cct - cct);

    lo = params->cct_min > 0.0f ? params->cct_min : p->wb_knots[0].cct;
    hi = params->cct_max > 0.0f ? params->cct_max : p->wb_knots[p->nb_wb_knots - 1].cct;
    hi = hi < lo ? lo : hi;
    cct = CRAW_CLAMP(cct, lo, hi);

    out->valid = pick->valid;
    out->illuminant.cct = (float)cct;
    out->illuminant.tint = pick->illuminant.tint;
    out->confidence = (float)conf;
    out->in_zone_fraction = pick->in_zone_fraction;
    out->spread_mired = pick->spread_mired;
    return 0;
}

/* ------------------------------------------------------------ smoother */

struct CRAWAwbSmoother {
    CRAWAwbSmootherParams params;
    bool tracking;
    double mired, tint;     /* the track, in mired */
    double pts;             /* of the last push or seed */
};

void craw_awb_smoother_params_init(CRAWAwbSmootherParams *p)
{
    if (!p)
        return;
    memset(p, 0, sizeof(*p));
    p->time_constant = 1.0f;
    p->hysteresis_mired = 3.0f;
    p->max_slew_mired_per_s = 30.0f;
    p->min_confidence = 0.2f;
}

int craw_awb_smoother_create(const CRAWAwbSmootherParams *params, CRAWAwbSmoother **sm)
{
    CRAWAwbSmoother *s;
    const CRAWAwbSmootherParams *q;

    if (!sm)
        return CRAW_ERROR(EINVAL);
    *sm = NULL;
    s = calloc(1, sizeof(*s));
    if (!s)
        return CRAW_ERROR(ENOMEM);
    if (params)
        s->params = *params;
    else
        craw_awb_smoother_params_init(&s->params);
    q = &s->params;
    if (!(q->time_constant > 0.0f) || !isfinite(q->time_constant) ||
        !(q->max_slew_mired_per_s > 0.0f) || !isfinite(q->max_slew_mired_per_s) ||
        !(q->hysteresis_mired >= 0.0f) || !isfinite(q->hysteresis_mired) ||
        !isfinite(q->min_confidence)) {
        free(s);
        return CRAW_ERROR(EINVAL);
    }
    *sm = s;
    return 0;
}

void craw_awb_smoother_free(CRAWAwbSmoother **sm)
{
    if (!sm || !*sm)
        return;
    free(*sm);
    *sm = NULL;
}

void craw_awb_smoother_reset(CRAWAwbSmoother *sm)
{
    if (!sm)
        return;
    sm->tracking = false;
    sm->mired = sm->tint = sm->pts = 0.0;
}

void craw_awb_smoother_set(CRAWAwbSmoother *sm, double pts_s, const CRAWIlluminant *illum)
{
    if (!sm || !illum || !isfinite(illum->cct) || !(illum->cct > 0.0f) ||
        !isfinite(illum->tint))
        return;
    sm->mired = 1e6 / (double)illum->cct;
    sm->tint = illum->tint;
    sm->tracking = true;
    sm->pts = pts_s;
}

int craw_awb_smoother_push(CRAWAwbSmoother *sm, double pts_s, const CRAWAwbEstimate *est,
                           bool scene_cut, CRAWIlluminant *out)
{
    double dt, to_mired = 0.0, to_tint = 0.0;
    bool usable, confident;

    if (!sm || !est || !out)
        return CRAW_ERROR(EINVAL);

    usable = est->valid && isfinite(est->illuminant.cct) && est->illuminant.cct > 0.0f &&
             isfinite(est->illuminant.tint);
    if (usable) {
        to_mired = 1e6 / (double)est->illuminant.cct;
        to_tint = est->illuminant.tint;
    }
    confident = usable && est->confidence >= sm->params.min_confidence;

    /* the time since the last push, in [0, 1] s: a seek or a gap moves the
     * track by at most a second's worth, a step back not at all */
    dt = sm->tracking ? pts_s - sm->pts : 0.0;
    dt = dt > 0.0 ? dt : 0.0;
    dt = dt < 1.0 ? dt : 1.0;

    if (usable && (!sm->tracking || (scene_cut && confident))) {
        /* the first usable estimate starts the track, confident or not; a
         * cut restarts it from a confident one */
        sm->mired = to_mired;
        sm->tint = to_tint;
        sm->tracking = true;
    } else if (!sm->tracking) {
        return CRAW_ERROR(EAGAIN);
    } else if (confident) {
        const double delta = to_mired - sm->mired;

        if (fabs(delta) >= (double)sm->params.hysteresis_mired) {
            double frac = 1.0 - exp(-dt / (double)sm->params.time_constant);
            const double limit = (double)sm->params.max_slew_mired_per_s * dt;
            const double step = delta * frac;

            /* the slew limit scales the fraction, so the tint moves by the
             * same share of its own error */
            if (fabs(step) > limit)
                frac *= limit / fabs(step);
            sm->mired += delta * frac;
            sm->tint += (to_tint - sm->tint) * frac;
        }
    }
    sm->pts = pts_s;
    out->cct = (float)(1e6 / sm->mired);
    out->tint = (float)sm->tint;
    return 0;
}
