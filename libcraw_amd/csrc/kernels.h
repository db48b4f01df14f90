/*
 * libcraw for MI355X: the kernel launchers the host pipeline calls.
 */
#ifndef CRAW_KERNELS_H
#define CRAW_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "k_common.h"

/* The chroma stage's buffers: camera RGB as three float planes of the
 * frame, the quarter grid (ceil(W/4) x ceil(H/4)) of block means as three
 * planes (R - G, B - G, G), and its bilateral result as two. */
struct KInter {
    float *rgb;
    float *down;
    float *blur;
    int wq, hq;
};

/* Constants of the raw-domain denoise (lib/pipeline.c:1414-1467 and
 * lib/ref/ref_denoise.c:26-46 of the reference), all formed on the host
 * in float exactly as the reference forms them. */
struct KDenoise {
    int w2, h2;                  /* plane size */
    int src_pitch;               /* uint16 elements */
    int dst_pitch;               /* uint16 elements */
    float black[4], range[4], inv_range[4];
    float gat_a, gat_c, gat_sc, gat_fwd, gat_inv, gat_off, gat_t0;
    float inv_h2;
    float guard_lo, guard_inv;
    float alpha_max;
};

/* Which filters run: spatial NLM, temporal average, and whether the
 * temporal one has a previous state to read. */
#define KDN_SPATIAL  1
#define KDN_TEMPORAL 2
#define KDN_HISTORY  4

struct KDenoiseState {
    const float *prev_level;     /* 4 planes of w2 x h2, or NULL */
    const uint8_t *prev_var;
    float *next_level;           /* or NULL without the temporal pass */
    uint8_t *next_var;
};

hipError_t craw_launch_zero_u32(uint32_t *p, size_t n, hipStream_t s);

/* method: 0 bilinear, 1 MHC, 2 RCD. to_inter: write the chroma stage's
 * inputs instead of finishing the frame. */
hipError_t craw_launch_debayer(int method, bool to_inter, const uint16_t *src, int src_pitch,
                               int W, int H, const KPrep &prep, const KColor &col,
                               const KOut &out, const KInter &inter, hipStream_t s);

hipError_t craw_launch_chroma_blur(const KInter &it, hipStream_t s);

hipError_t craw_launch_chroma_up_finish(const KInter &it, float strength, int W, int H,
                                        const KColor &col, const KOut &out, hipStream_t s);

/* mode: KDN_* bits. fp16: the packed-FP16 distance form. */
hipError_t craw_launch_denoise(int mode, bool fp16, const KDenoise &p, const uint16_t *src,
                               uint16_t *dst, const KDenoiseState &st, hipStream_t s);

#endif /* CRAW_KERNELS_H */
