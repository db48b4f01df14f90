/*
 * libcraw for MI355X: device-side helpers and the argument blocks the
 * kernels are launched with.
 *
 * Every kernel in this library that produces path values is built without
 * floating-point contraction and with correctly rounded float division
 * and square root, and evaluates each expression in the order
 * cyanreg/libcraw's CPU reference (lib/ref) writes it, so that its float
 * results are the CPU reference's bit for bit (transcendentals through
 * detmath.h).
 */
#ifndef CRAW_K_COMMON_H
#define CRAW_K_COMMON_H

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "detmath.h"
#include "hostmath.h"

#define KDEV __device__ __forceinline__

/* Mirror without repeating the edge sample (lib/ref/ref_debayer.c:20-33,
 * lib/ref/ref_denoise.c:63-76): as many reflections as it takes, 0 for a
 * dimension below 2. */
KDEV int kmirror(int v, int n)
{
    if (n < 2)
        return 0;
    while (v < 0 || v >= n) {
        if (v < 0)
            v = -v;
        if (v >= n)
            v = 2 * n - 2 - v;
    }
    return v;
}

/* Per-site value out of four, by selects rather than an indexed load
 * (the site varies per lane; an indexed kernel argument would go through
 * scratch memory). */
KDEV float ksel4(int s, float a0, float a1, float a2, float a3)
{
    float lo = (s & 1) ? a1 : a0;
    float hi = (s & 1) ? a3 : a2;
    return (s & 2) ? hi : lo;
}

KDEV double ksel4d(int s, double a0, double a1, double a2, double a3)
{
    double lo = (s & 1) ? a1 : a0;
    double hi = (s & 1) ? a3 : a2;
    return (s & 2) ? hi : lo;
}

/* Black/white normalisation, white balance and clip of one sample,
 * as lib/ref/ref_debayer.c:42-63 computes it: divide by the
 * range (white - black, formed on the host in float, exactly as the
 * reference forms it), multiply by the channel's gain, clip at 1.0;
 * below black is kept. With `half` the mosaic is the highlight stage's
 * rebuilt one (R16F bits, already balanced and unclipped) and the prep is
 * the identity (lib/pipeline.c:2052-2062). */
struct KPrep {
    float black[4];
    float range[4];
    float wb[4];      /* per quad site: R, G, G, B gains */
    double rrange[4]; /* 1 / range[s] in double, correctly rounded */
    int half;
    int fdiv;         /* divide in float (the precision form: no double arithmetic) */
    /* The frame's black levels, ranges and gains bound the debayer's
     * quotients to operands kdiv_rn divides exactly (prep_kdiv_ok in
     * plan.cpp). */
    int kdiv;
};

/* n / d correctly rounded, for d in [2^-40, 2^40] and n +0 or in
 * magnitude in [2^-80, 2^40], where neither operand nor the quotient is
 * near the ends of the float range (a -0 dividend gives +0 where the
 * division gives -0). A reciprocal estimate refined once by Newton's
 * iteration, then the quotient corrected twice by its exact residual
 * (Markstein, IBM J. Res. Dev. 34(1), 1990): the steps of the compiler's
 * correctly rounded float division without its scaling and fix-up
 * instructions (v_div_scale, v_div_fmas, v_div_fixup), which leave such
 * operands unchanged; eight instructions against eleven. Checked against
 * IEEE division on the GPU over operands across that range
 * (tests/test_gpu_kdiv.py). */
KDEV float kdiv_rn(float n, float d)
{
    float y = __builtin_amdgcn_rcpf(d);
    const float e = __builtin_fmaf(-d, y, 1.0f);
    float q, r;

    y = __builtin_fmaf(e, y, y);
    q = n * y;
    r = __builtin_fmaf(-d, q, n);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-d, q, n);
    return __builtin_fmaf(r, y, q);
}

/* The division of x = raw / 65535 - black by the range is x times the
 * range's reciprocal in double, rounded once to float: the product is
 * within 2^-52 (relative) of x / range, and a quotient of two floats lies
 * at least 2^-49 from every midpoint of the float grid, so its rounding is
 * the correctly rounded float quotient, the float division's value (three
 * instructions against eleven). The precision form divides in float,
 * through kdiv_rn where the frame's constants keep x and the range in its
 * range (kdiv: x is 0 or at least 2^-40 in magnitude, the range in
 * [2^-16, 2]). */
KDEV float kprep(uint32_t raw, int s, const KPrep &p)
{
    if (p.half)
        return __half2float(__ushort_as_half((unsigned short)raw));
    float r = (float)raw * (1.0f / 65535.0f);
    const float x = r - ksel4(s, p.black[0], p.black[1], p.black[2], p.black[3]);
    const float range = ksel4(s, p.range[0], p.range[1], p.range[2], p.range[3]);
    float t = p.fdiv ? (p.kdiv ? kdiv_rn(x, range) : x / range)
                     : (float)((double)x * ksel4d(s, p.rrange[0], p.rrange[1], p.rrange[2],
                                                  p.rrange[3]));
    t *= ksel4(s, p.wb[0], p.wb[1], p.wb[2], p.wb[3]);
    return t < 1.0f ? t : 1.0f;
}

/* The colour stage's constants (lib/ref/ref_color.c:10-76): the frame's
 * cam_to_xyz, the scale gain * 2^stops, XYZ to the output space, and the
 * gamut compression. The two products are applied one after the other,
 * as the reference applies them. */
#define KCOLOR_CLAMP_NEGATIVE 1u
#define KCOLOR_GAMUT_COMPRESS 2u
/* The precision form (CRAW_PIPELINE_FLAG_NLM_FP16): the transcendentals of
 * the colour, chroma and histogram stages in float on the hardware's
 * v_exp_f32 / v_log_f32 (2^x and log2 x, about an ulp each), with no
 * double arithmetic, instead of the correctly rounded ones of detmath.h. */
#define KCOLOR_FLOAT_MATH     4u

KDEV float kf_exp(float x)
{
    return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
}

KDEV float kf_log(float x)
{
    return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
}

/* x^y for x >= 0: 2^(y log2 x), 0 at x = 0 for y > 0. */
KDEV float kf_pow(float x, float y)
{
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

struct KColor {
    float cam[3][3];
    float scale;
    float out[3][3];
    float thr[3], scl[3], pwr, inv_pwr;
    float thr_lo[3];     /* kgamut's screen: kgamut_screen_bound(thr[i]) */
    uint32_t flags;
};

/* A bound below a gamut threshold for kgamut's screen: thr (1 - 2^-20),
 * or -inf (the screen never passes) for a threshold that is not a
 * positive float in [2^-100, 2^100]. Rounding the double product to
 * float moves it by at most 2^-24 relative, which the margin covers. */
__host__ __device__ static inline float kgamut_screen_bound(float thr)
{
    if (!(thr >= 0x1p-100f && thr <= 0x1p100f))
        return -__builtin_inff();
    return (float)((double)thr * (1.0 - 0x1p-20));
}

/* lib/gamut.c:66-92: ACES 1.3 RGC of one AP1 triple. Each channel's
 * distance comes from the input triple; the compression of a channel past
 * its threshold is one loop body (not unrolled), so its two powf
 * evaluations appear once in the code.
 *
 * Most triples have every distance below its threshold and leave the
 * compression unchanged. A screen finds them without the three correctly
 * rounded divisions: with y the hardware reciprocal of mag (1 ulp) and
 * q = (ach - x) y rounded, q is within 2^-22 (relative) of the quotient
 * (ach - x) / mag for mag in [2^-64, 2^64] (y a normal float; a q that
 * underflows is within 2^-149 of it). q < thr_lo = thr (1 - 2^-20) then
 * puts the quotient below thr (1 - 2^-21), under the midpoint between thr
 * and its predecessor (at least thr (1 - 2^-24)), so its correct rounding,
 * the distance d, is below thr and the channel is left as it is. A NaN or
 * infinite operand fails the screen and takes the exact path. */
KDEV void kgamut(const KColor &c, float &r, float &g, float &b)
{
    const float ach = r > (g > b ? g : b) ? r : (g > b ? g : b);
    /* the constants read once, before the loop: indexed by the loop's
     * counter inside it they would be a scalar load and a wait a pass */
    const float t0 = c.thr[0], t1 = c.thr[1], t2 = c.thr[2];
    const float s0 = c.scl[0], s1 = c.scl[1], s2 = c.scl[2];
    const float pwr = c.pwr, inv_pwr = c.inv_pwr;
    const bool float_math = (c.flags & KCOLOR_FLOAT_MATH) != 0;
    float mag, d0, d1, d2;

    if (ach == 0.0f)
        return;
    mag = fabsf(ach);
    {
        const float y = __builtin_amdgcn_rcpf(mag);

        if (mag >= 0x1p-64f && mag <= 0x1p64f && (ach - r) * y < c.thr_lo[0] &&
            (ach - g) * y < c.thr_lo[1] && (ach - b) * y < c.thr_lo[2])
            return;
    }
    d0 = (ach - r) / mag;
    d1 = (ach - g) / mag;
    d2 = (ach - b) / mag;
#pragma unroll 1
    for (int i = 0; i < 3; i++) {
        const float d = i == 0 ? d0 : (i == 1 ? d1 : d2);
        const float thr = i == 0 ? t0 : (i == 1 ? t1 : t2);
        const float scl = i == 0 ? s0 : (i == 1 ? s1 : s2);
        float nd, p, v;

        if (d < thr)
            continue;
        nd = (d - thr) / scl;
        if (float_math) {
            p = kf_pow(nd, pwr);
            v = ach - (thr + scl * nd / kf_pow(1.0f + p, inv_pwr)) * mag;
        } else {
            p = craw_dm_powf(nd, pwr);
            v = ach - (thr + scl * nd / craw_dm_powf(1.0f + p, inv_pwr)) * mag;
        }
        if (i == 0)
            r = v;
        else if (i == 1)
            g = v;
        else
            b = v;
    }
}

/* GAMUT false: a caller that never sees the gamut compression (its code
 * left out). */
template <bool GAMUT = true>
KDEV void kcolor(const KColor &c, float cr, float cg, float cb, float o[3])
{
    float x[3];

#pragma unroll
    for (int r = 0; r < 3; r++)
        x[r] = (c.cam[r][0] * cr + c.cam[r][1] * cg + c.cam[r][2] * cb) * c.scale;
#pragma unroll
    for (int r = 0; r < 3; r++)
        o[r] = c.out[r][0] * x[0] + c.out[r][1] * x[1] + c.out[r][2] * x[2];
    if (GAMUT && (c.flags & KCOLOR_GAMUT_COMPRESS))
        kgamut(c, o[0], o[1], o[2]);
    if (c.flags & KCOLOR_CLAMP_NEGATIVE) {
#pragma unroll
        for (int r = 0; r < 3; r++)
            o[r] = o[r] > 0.0f ? o[r] : 0.0f;
    }
}

/* Where a tile's finished camera RGB goes. */
struct KOut {
    void *dst;           /* RGBA16F or RGBA32F */
    size_t pitch;        /* bytes */
    int half;            /* 1: RGBA16F */
    /* AWB (hist NULL: none): the finish stage writes one packed sample
     * per 8x8 block into `samples` (row-major over the frame's
     * ceil(W/8) x ceil(H/8) blocks); the histogram stage adds them into
     * `hist` (count, weight, bright, each 128^2). */
    uint32_t *hist;
    uint2 *samples;
    float awb_wb0, awb_wb2;
    int float_log;       /* the precision form's float logarithm (craw_awb_sample_log) */
};

/* A block's histogram sample, packed: x = bin | weight << 14 (bin < 2^14,
 * weight <= 2^15), y = bright (<= 2^16); x = KAWB_NONE for a block the
 * estimator skips. */
#define KAWB_NONE 0xffffffffu

/* A pointer to global memory as such (address space 1), for loads: a
 * pointer read from memory rather than from the kernel's arguments has no
 * known address space, and its loads would be flat ones, which every LDS
 * wait then waits for too. */
#define KGLOBAL(T) T __attribute__((address_space(1)))
template <class T>
KDEV KGLOBAL(T) *kglobal(T *p)
{
    return (KGLOBAL(T) *)p;
}

KDEV void kstore_rgba(const KOut &o, int x, int y, const float v[3])
{
    char *row = (char *)o.dst + (size_t)y * o.pitch;

    if (o.half) {
        __half2 lo = __halves2half2(__float2half_rn(v[0]), __float2half_rn(v[1]));
        __half2 hi = __halves2half2(__float2half_rn(v[2]), __float2half_rn(1.0f));
        uint2 u;
        __builtin_memcpy(&u.x, &lo, 4);
        __builtin_memcpy(&u.y, &hi, 4);
        ((uint2 *)row)[x] = u;
    } else {
        ((float4 *)row)[x] = make_float4(v[0], v[1], v[2], 1.0f);
    }
}

#endif /* CRAW_K_COMMON_H */
