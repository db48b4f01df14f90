/*
 * libcraw for MI355X: internal host math declarations, and the AWB
 * histogram's per-block arithmetic, which the CPU builder and the GPU
 * kernels share so that the two bin every block identically.
 */
#ifndef CRAW_HOSTMATH_H
#define CRAW_HOSTMATH_H

#include <stdint.h>
#include <stdbool.h>
#include "detmath.h"

#ifdef __cplusplus
extern "C" {
#endif

#include <craw/profile.h>

/* ACES 1.3 reference gamut compression constants, per channel cyan,
 * magenta, yellow (lib/gamut.h:16-21 of the reference). */
typedef struct CRAWGamutParams {
    float thr[3];
    float scl[3];
    float pwr;
    float inv_pwr;
} CRAWGamutParams;

void craw_gamut_params(CRAWGamutParams *g);

/* A colour space's RGB to XYZ D65 in double; not exported. */
int craw_colorspace_to_xyz_d(enum CRAWColorSpace cs, double out[3][3]);

#ifdef __cplusplus
}
#endif

/* One histogram sample: bin index, 16.16 weights. */
typedef struct CRAWAwbSample {
    uint32_t bin;
    uint32_t weight;
    uint32_t bright;
} CRAWAwbSample;

/* lib/awb.c:258-291 of the reference, for one block whose channel sums
 * (taken in raster order) are `acc` over `n` pixels: the block mean,
 * the clip and black skips, ln(R/G) and ln(B/G) of the raw response
 * with the frame's gains divided out, the bin on [-2.5, 1] in 128 steps
 * (dividing by the span, then multiplying by the bins, as the CPU
 * reference does; a NaN lands in bin 0), and the 16.16 weights
 * min(g, 0.5) and g^4 rounded half away from zero. The logarithm is the
 * deterministic one of detmath.h. Returns false for a skipped block. */
CRAW_DM_FN bool craw_awb_sample_log(const float acc[3], float n, float wb0, float wb2,
                                    bool float_log, CRAWAwbSample *s);

CRAW_DM_FN bool craw_awb_sample(const float acc[3], float n, float wb0, float wb2,
                                CRAWAwbSample *s)
{
    return craw_awb_sample_log(acc, n, wb0, wb2, false, s);
}

/* The same with the logarithm chosen: float_log (device code only) takes
 * it from the hardware's float log2 (v_log_f32) times ln 2, with no double
 * arithmetic: the precision form's histogram, whose bins then differ from
 * the correctly rounded logarithm's only for ratios within an ulp or two
 * of a bin edge. */
CRAW_DM_FN bool craw_awb_sample_log(const float acc[3], float n, float wb0, float wb2,
                                    bool float_log, CRAWAwbSample *s)
{
    const float lmin = -2.5f, lmax = 1.0f, span = lmax - lmin;
    float r = acc[0] / n, g = acc[1] / n, b = acc[2] / n;
    float mx = r > g ? r : g;
    float rr, bb, x, y, w;
    int bx, by;

    mx = mx > b ? mx : b;
    if (mx >= 0.98f || g < 0.002f)
        return false;

    rr = (r / wb0) / g;
    bb = (b / wb2) / g;
#if defined(__HIP_DEVICE_COMPILE__)
    if (float_log) {
        x = __builtin_amdgcn_logf(rr) * 0.693147180559945309f;
        y = __builtin_amdgcn_logf(bb) * 0.693147180559945309f;
    } else {
        x = craw_dm_logf(rr);
        y = craw_dm_logf(bb);
    }
#else
    (void)float_log;
    x = craw_dm_logf(rr);
    y = craw_dm_logf(bb);
#endif
    x = x > lmin ? (x < lmax ? x : lmax) : lmin;
    y = y > lmin ? (y < lmax ? y : lmax) : lmin;

    bx = (int)floorf((x - lmin) / span * 128.0f);
    by = (int)floorf((y - lmin) / span * 128.0f);
    bx = bx > 0 ? (bx < 127 ? bx : 127) : 0;
    by = by > 0 ? (by < 127 ? by : 127) : 0;

    w = g < 0.5f ? g : 0.5f;
    s->bin = (uint32_t)by * 128u + (uint32_t)bx;
    s->weight = (uint32_t)roundf(w * 65536.0f);
    s->bright = (uint32_t)roundf(g * g * g * g * 65536.0f);
    return true;
}

#endif /* CRAW_HOSTMATH_H */
