/*
 * libcraw for MI355X: the noise fit shared by the host and device
 * measurements of a frame (noise_host.c, noise_dev.cpp).
 */
#ifndef CRAW_NOISE_FIT_H
#define CRAW_NOISE_FIT_H

#include <stdint.h>

#include <craw/noise.h>
#include <craw/profile.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The 9x9 box mean's radius a sample is binned by. */
#define CRAW_NOISE_MEAN_R 4

/* What a frame's measurement leaves for one intensity bin: its sample
 * count, the sum of its samples' local means, and the median of their
 * |residual| (the element of rank n / 2 in ascending order). */
typedef struct CRAWNoiseBin {
    uint64_t n;
    double mean_sum;
    float median;
} CRAWNoiseBin;

/* Resolved options; CRAW_ERROR(EINVAL) for a bin count out of range. */
int craw_noise_resolve(const CRAWNoiseEstimateParams *params, uint32_t *nb_bins,
                       uint32_t *min_samples, float *max_level);

/* Checks a frame and profile as the estimate requires. */
int craw_noise_check(uint32_t width, uint32_t height, size_t stride,
                     const CRAWProfile *profile);

/* The robust line through the bins holding at least `min_samples`
 * samples, and the estimate's fields. */
int craw_noise_fit(const CRAWNoiseBin *bins, uint32_t nb_bins, uint32_t min_samples,
                   const CRAWProfile *profile, CRAWNoiseEstimate *out);

#ifdef __cplusplus
}
#endif

#endif
