/*
 * libcraw for MI355X: sensor noise model.
 *
 * Interface of include/craw/noise.h in cyanreg/libcraw (noise.h:53-161):
 * the variance of a sensel is var(u) = a * u + b, u the signal above black
 * normalised so that black is 0 and white 1, before any white balance gain;
 * its evaluation from a profile's table and its fit to a frame.
 */
#ifndef CRAW_NOISE_H
#define CRAW_NOISE_H

#include <stdint.h>
#include <stdbool.h>

#include <craw/utils.h>

CRAW_BEGIN_DECLS

struct CRAWProfile;

typedef struct CRAWNoiseModel {
    float a;    /* shot term */
    float b;    /* read term */
} CRAWNoiseModel;

/* The model measured at one exposure index; tables ascend in iso. */
typedef struct CRAWNoiseKnot {
    float iso;
    float a, b;
} CRAWNoiseKnot;

#define CRAW_PROFILE_MAX_NOISE_KNOTS 16

/* Evaluates a knot table at `iso`: log-log interpolation between knots, a
 * proportional to iso outside the table, b proportional to iso^2 above it
 * and constant below it, the first knot for an iso that is not positive
 * and finite. CRAW_ERROR(EINVAL) for an empty table. Follows
 * lib/noise.c:37-110 of the reference. */
CRAW_API int craw_noise_model_eval(const CRAWNoiseKnot *knots, uint32_t nb_knots,
                                   float iso, CRAWNoiseModel *out);

/* ------------------------------------------------------------ estimation */

/* Options of the noise fit (noise.h:84-104 of the reference). Zero fields
 * take the defaults. */
typedef struct CRAWNoiseEstimateParams {
    uint32_t nb_bins;           /* intensity bins in sqrt(u), 8..512; 0: 128 */
    uint32_t min_bin_samples;   /* smaller bins are dropped; 0: 20000 */
    float max_level;            /* local means above it are ignored; 0: 0.98 */
    uint8_t padding[64 - 4 * 3];
} CRAWNoiseEstimateParams;

/* The fitted model and the evidence for it (noise.h:106-129). */
typedef struct CRAWNoiseEstimate {
    CRAWNoiseModel model;
    uint32_t nb_bins;           /* bins with enough samples */
    uint32_t nb_inliers;        /* of those, bins the fit explains (|ln error| < 0.35) */
    float fit_rms;              /* RMS ln error over the inliers */
    float read_noise_lsb;       /* sqrt(b) in 16-bit codes of the first site's range */
    uint8_t padding[64 - 4 * 6];
} CRAWNoiseEstimate;

#define CRAW_NOISE_MAX_BINS 512

CRAW_API void craw_noise_estimate_params_init(CRAWNoiseEstimateParams *p);

/* Fits var = a u + b to one RGGB frame in host memory (noise.c:338-504 of
 * the reference): each CFA plane high-passed with [1 -2 1; -2 4 -2;
 * 1 -2 1] / 6, every sample binned by its 9x9 plane mean (bins equal in
 * sqrt(mean)), each bin's noise the median |residual| / 0.6745, and the
 * line through the bins fitted robustly (least trimmed squares over bin
 * pairs, then reweighted Gauss-Newton on ln(var) with a Tukey biweight).
 * Both dimensions even and at least 32, `stride` in samples.
 * CRAW_ERROR(ENOTSUP) for another CFA phase, CRAW_ERROR(ENODATA) when
 * fewer than three bins survive. */
CRAW_API int craw_noise_estimate(const uint16_t *bayer, uint32_t width, uint32_t height,
                                 size_t stride, const struct CRAWProfile *profile,
                                 const CRAWNoiseEstimateParams *params,
                                 CRAWNoiseEstimate *out);

CRAW_END_DECLS

#endif /* CRAW_NOISE_H */
