/*
 * libcraw for MI355X: sensor noise model.
 *
 * Interface of include/craw/noise.h in cyanreg/libcraw (noise.h:53-161):
 * the variance of a sensel is var(u) = a * u + b, u the signal above black
 * normalised so that black is 0 and white 1, before any white balance gain.
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

CRAW_END_DECLS

#endif /* CRAW_NOISE_H */
