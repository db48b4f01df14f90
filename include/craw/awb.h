/*
 * libcraw for MI355X: auto white balance histogram.
 *
 * Interface of include/craw/awb.h in cyanreg/libcraw: the histogram
 * geometry, the readback struct and the CPU builder (awb.h:44-60, 134-142),
 * and the host estimator and temporal smoother that consume the
 * histogram a recording leaves (awb.h:62-176), in double on the host.
 * Same names, structs and fields as the reference.
 */
#ifndef CRAW_AWB_H
#define CRAW_AWB_H

#include <stdint.h>
#include <stdbool.h>

#include <craw/profile.h>

CRAW_BEGIN_DECLS

#define CRAW_AWB_HIST_BINS     128
#define CRAW_AWB_HIST_LOG_MIN -2.5f
#define CRAW_AWB_HIST_LOG_MAX  1.0f

/* ln(R/G) across, ln(B/G) down, of the raw response (the frame's gains
 * divided back out), one sample per 8x8 block; index = iy * bins + ix. */
typedef struct CRAWAwbHistogram {
    uint32_t bins;
    float log_min, log_max;
    const uint32_t *count;          /* samples per bin */
    const uint32_t *weight;         /* sum of min(g, 0.5), 16.16 */
    const uint32_t *weight_bright;  /* sum of g^4, 16.16 */
} CRAWAwbHistogram;

/* CPU histogram with the GPU pass's semantics (awb.c:213-296 of the
 * reference): `rgb` is camera RGB after the gains `wb_applied`, `channels`
 * 3 or 4, `stride` in floats; partial blocks at the right and bottom edges
 * average the pixels they have. The three outputs hold bins^2 entries and
 * are cleared first. */
CRAW_API int craw_awb_histogram_from_rgb(const float *rgb, uint32_t width,
                                         uint32_t height, size_t stride,
                                         int channels, const float wb_applied[3],
                                         uint32_t *count, uint32_t *weight,
                                         uint32_t *weight_bright);

/* Estimator parameters (awb.h:62-88 of the reference). */
typedef struct CRAWAwbParams {
    float band;                     /* ln-ratio distance from the locus; default 0.08 */
    float sigma;                    /* Gaussian falloff with that distance; default 0.05 */
    float min_zone_fraction;        /* share of weight inside the band; default 0.02 */
    float ambience;                 /* warmth retained below the knee, 0..1; default 0 */
    float ambience_knee_cct;        /* default 4000 */
    float cct_min, cct_max;         /* 0: the profile's table range */
    bool prefer_bright;             /* the g^4-weighted reading first; default true */
    float bright_max_spread_mired;  /* its trust limit along the locus; default 60 */
} CRAWAwbParams;

CRAW_API void craw_awb_params_init(CRAWAwbParams *p);

/* One reading of the histogram under one weighting. */
typedef struct CRAWAwbCandidate {
    bool valid;
    CRAWIlluminant illuminant;      /* before ambience and clamping */
    float in_zone_fraction;
    float spread_mired;             /* weighted standard deviation along the locus */
} CRAWAwbCandidate;

enum CRAWAwbSource {
    CRAW_AWB_SOURCE_NONE = 0,
    CRAW_AWB_SOURCE_ZONE,           /* the linearly weighted candidates */
    CRAW_AWB_SOURCE_BRIGHT,         /* the g^4 weighted ones */
};

typedef struct CRAWAwbEstimate {
    bool valid;
    CRAWIlluminant illuminant;      /* the chosen reading after ambience and clamping */
    float confidence;               /* 0..1 */
    float in_zone_fraction;
    float spread_mired;
    enum CRAWAwbSource source;
    CRAWAwbCandidate zone;
    CRAWAwbCandidate bright;
} CRAWAwbEstimate;

/* The illuminant of one frame from its histogram (awb.c:86-211 of the
 * reference). Needs a profile with a white balance table of two knots or
 * more; `params` NULL for the defaults. */
CRAW_API int craw_awb_estimate(const CRAWProfile *p, const CRAWAwbHistogram *hist,
                               const CRAWAwbParams *params, CRAWAwbEstimate *out);

/* Temporal smoothing of the estimates into an illuminant track
 * (awb.h:144-176, awb_smoother.c:21-172 of the reference). */
typedef struct CRAWAwbSmoother CRAWAwbSmoother;

typedef struct CRAWAwbSmootherParams {
    float time_constant;            /* seconds, exponential in mired; default 1.0 */
    float hysteresis_mired;         /* smaller changes are ignored; default 3.0 */
    float max_slew_mired_per_s;     /* rate limit; default 30.0 */
    float min_confidence;           /* below it an estimate only holds the track; default 0.2 */
} CRAWAwbSmootherParams;

CRAW_API void craw_awb_smoother_params_init(CRAWAwbSmootherParams *p);
CRAW_API int craw_awb_smoother_create(const CRAWAwbSmootherParams *params,
                                      CRAWAwbSmoother **sm);
CRAW_API void craw_awb_smoother_free(CRAWAwbSmoother **sm);
/* Forgets the track; the next push starts from its estimate. */
CRAW_API void craw_awb_smoother_reset(CRAWAwbSmoother *sm);
/* Puts the track at `illum` as of `pts_s` (a seed); ignored unless the
 * illuminant is finite with a positive cct. */
CRAW_API void craw_awb_smoother_set(CRAWAwbSmoother *sm, double pts_s,
                                    const CRAWIlluminant *illum);
/* Feeds one frame's estimate at `pts_s` and gives the illuminant to apply:
 * the first usable estimate starts the track; then a confident one moves
 * it (or with `scene_cut` replaces it), others hold it. CRAW_ERROR(EAGAIN)
 * with `out` untouched while there is no track and no usable estimate. */
CRAW_API int craw_awb_smoother_push(CRAWAwbSmoother *sm, double pts_s,
                                    const CRAWAwbEstimate *est, bool scene_cut,
                                    CRAWIlluminant *out);

CRAW_END_DECLS

#endif /* CRAW_AWB_H */
