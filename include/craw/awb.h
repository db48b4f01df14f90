/*
 * libcraw for MI355X: auto white balance histogram.
 *
 * Interface of include/craw/awb.h in cyanreg/libcraw (awb.h:44-60, 134-142):
 * the histogram geometry, the readback struct and the CPU builder. The
 * estimator and the temporal smoother that consume it (awb.h:62-176) run on
 * the host after the path and are not part of this library.
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

CRAW_END_DECLS

#endif /* CRAW_AWB_H */
