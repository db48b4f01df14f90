/*
 * Deterministic single-precision exp, log and pow for host and device.
 *
 * The CPU reference of libcraw calls glibc's expf (NLM weights, temporal
 * alpha, chroma bilateral), logf (AWB histogram) and powf (ACES gamut
 * compression). The GPU's own libraries round those differently, and a one
 * ulp difference in a weight moves a 16-bit re-encode or a histogram bin.
 * These functions are one source compiled into both the HIP kernels and
 * the C oracle, so that the two agree bit for bit by construction.
 *
 * Method: the argument is taken to double, the function evaluated there to
 * about 2^-51 relative, and the result rounded once to float. On CDNA4 a
 * double FMA issues at the same rate as a float one, so this costs the
 * GPU about what a float polynomial would. Because the double result is
 * within ~2^-51 of the true value, the float result is the correctly
 * rounded one except for arguments within that distance of a rounding
 * midpoint (for exp, none: every float argument in [-200, 90] rounds
 * correctly), which is also where glibc's (nearly correctly rounded)
 * versions would be the ones to differ.
 *
 *   exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, with ln2 split
 *        into a 32-bit head and a tail (the fdlibm split) and k rounded by
 *        the 1.5 * 2^52 shift, degree-10
 *        Chebyshev polynomial of exp(r) in Horner form with FMA, scaled by
 *        2^k through the exponent field.
 *   log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
 *        s = (m - 1) / (m + 1), odd series to s^19.
 *   pow: exp(y log x) for x >= 0, both halves in double.
 *
 * Every operation is an IEEE double add, multiply, divide or fused
 * multiply-add, or an exact integer manipulation, so host and device
 * produce identical bits. The translation units that include this must
 * be built without floating-point contraction (-ffp-contract=off) so that
 * the compiler fuses nothing on its own, and the host with FMA support
 * (-mfma), so that fma() is the hardware instruction rather than a libm
 * emulation (both are correctly rounded; the flag is for speed).
 */
#ifndef CRAW_DETMATH_H
#define CRAW_DETMATH_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__) || defined(__HIP_DEVICE_COMPILE__)
#define CRAW_DM_FN __host__ __device__ static inline
#else
#define CRAW_DM_FN static inline
#endif

/* lo <= x <= hi for a number; any value in that range for a NaN (which
 * callers replace). The device form is one max and one min. */
#if defined(__HIP_DEVICE_COMPILE__)
#define CRAW_DM_CLAMP(x, lo, hi) __builtin_fmin(__builtin_fmax((x), (lo)), (hi))
#else
#define CRAW_DM_CLAMP(x, lo, hi) ((x) < (lo) ? (lo) : ((x) > (hi) ? (hi) : (x)))
#endif

CRAW_DM_FN uint64_t craw_dm_bits(double x)
{
    uint64_t u;
    __builtin_memcpy(&u, &x, sizeof(u));
    return u;
}

CRAW_DM_FN double craw_dm_from_bits(uint64_t u)
{
    double x;
    __builtin_memcpy(&x, &u, sizeof(x));
    return x;
}

/* exp(x) in double, for |x| <= 700. */
CRAW_DM_FN double craw_dm_exp_d(double x)
{
    const double inv_ln2 = 0x1.71547652b82fep+0;
    const double ln2_hi  = 0x1.62e42fee00000p-1;
    const double ln2_lo  = 0x1.a39ef35793c76p-33;
    /* k = the integer nearest x / ln2: x / ln2 + 1.5 * 2^52 rounded once
     * (one FMA) holds k in its low mantissa bits, and taking the shift
     * back off is exact. Over every float argument in [-200, 90] this k
     * equals rint(x * inv_ln2) (checked exhaustively). */
    const double shift = 0x1.8p52;
    double t = __builtin_fma(x, inv_ln2, shift);
    double k = t - shift;
    double r = __builtin_fma(-k, ln2_hi, x);
    double p;
    int32_t ki;

    r = __builtin_fma(-k, ln2_lo, r);

    /* Degree-10 Chebyshev fit of exp on [-ln2/2, ln2/2] (mpmath chebyfit,
     * coefficients rounded to double): relative error 2^-51.4 with the
     * rounded coefficients, and over every float argument in [-200, 90]
     * its float results equal the correctly rounded exp (checked
     * exhaustively against long double expl). */
    p = 0x1.28a2c0c0079ddp-22;
    p = __builtin_fma(p, r, 0x1.72faf04a14bf2p-19);
    p = __builtin_fma(p, r, 0x1.a019a66a5defap-16);
    p = __builtin_fma(p, r, 0x1.a01978c697102p-13);
    p = __builtin_fma(p, r, 0x1.6c16c17f43ad0p-10);
    p = __builtin_fma(p, r, 0x1.1111112dd687ap-7);
    p = __builtin_fma(p, r, 0x1.55555555520afp-5);
    p = __builtin_fma(p, r, 0x1.555555554b756p-3);
    p = __builtin_fma(p, r, 0x1.0000000000005p-1);
    p = __builtin_fma(p, r, 0x1.000000000001ep+0);
    p = __builtin_fma(p, r, 1.0);

    /* p in [0.70, 1.42]; 2^k is a normal double for |k| <= 1010, so the
     * scaling is exact (one ldexp on the device, a multiply by the power
     * built in the exponent field on the host: the same value). */
    ki = (int32_t)(uint32_t)craw_dm_bits(t);
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_ldexp(p, ki);
#else
    return p * craw_dm_from_bits((uint64_t)(uint32_t)(ki + 1023) << 52);
#endif
}

/* exp(x) in double through a table of 2^(j/N), for -200 <= x <= 0: the
 * kernels' form of craw_dm_exp_d, whose float rounding it equals (below).
 *
 * x = m ln2/N + r with m the integer nearest x N/ln2 (the 1.5 * 2^52 shift
 * again: t's low word is m), |r| <= ln2/(2N) by the same two-part
 * Cody-Waite reduction (m has at most 19 bits, so m times the 32-bit head
 * is exact); exp(x) = 2^(m >> log2 N) * T[m mod N] * exp(r), with exp(r) - 1
 * its Taylor polynomial, degree 4 for N = 256 (truncation 2^-54.6) and
 * degree 3 for N = 1024 (2^-50.7). T is the table of detmath_tab.h, 2^(j/N)
 * rounded to double; on the GPU a copy of it in LDS, indexed by the low bits
 * of m. Over every float argument in [-200, 0] the float rounding of either
 * result equals that of craw_dm_exp_d, i.e. the correctly rounded exp
 * (checked exhaustively, both table sizes). Against craw_dm_exp_d's 13
 * FMAs it spends 7 (N = 256) or 6 (N = 1024) double operations and an LDS
 * read. */
CRAW_DM_FN double craw_dm_exp_t256_d(double x, const double *tab)
{
    const double shift = 0x1.8p52;
    double t = __builtin_fma(x, 0x1.71547652b82fep+8, shift);      /* 256 / ln2 */
    double m = t - shift;
    double r = __builtin_fma(-m, 0x1.62e42fee00000p-9, x);         /* ln2/256 head */
    uint32_t lo;
    double T, c;

    r = __builtin_fma(-m, 0x1.a39ef35793c76p-41, r);               /* and tail */
    lo = (uint32_t)craw_dm_bits(t);
    T = tab[lo & 255u];
    c = __builtin_fma(r, 0x1.5555555555555p-5, 0x1.5555555555555p-3);
    c = __builtin_fma(c, r, 0.5);
    c = __builtin_fma(c, r, 1.0);
    c = c * r;
    T = __builtin_fma(T, c, T);
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_ldexp(T, (int32_t)lo >> 8);
#else
    return T * craw_dm_from_bits((uint64_t)(uint32_t)(((int32_t)lo >> 8) + 1023) << 52);
#endif
}

CRAW_DM_FN double craw_dm_exp_t1024_d(double x, const double *tab)
{
    const double shift = 0x1.8p52;
    double t = __builtin_fma(x, 0x1.71547652b82fep+10, shift);     /* 1024 / ln2 */
    double m = t - shift;
    double r = __builtin_fma(-m, 0x1.62e42fee00000p-11, x);        /* ln2/1024 head */
    uint32_t lo;
    double T, c;

    r = __builtin_fma(-m, 0x1.a39ef35793c76p-43, r);               /* and tail */
    lo = (uint32_t)craw_dm_bits(t);
    T = tab[lo & 1023u];
    c = __builtin_fma(r, 0x1.5555555555555p-3, 0.5);
    c = __builtin_fma(c, r, 1.0);
    c = c * r;
    T = __builtin_fma(T, c, T);
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_ldexp(T, (int32_t)lo >> 10);
#else
    return T * craw_dm_from_bits((uint64_t)(uint32_t)(((int32_t)lo >> 10) + 1023) << 52);
#endif
}

/* log(x) in double, for finite x > 0 (double normal). */
CRAW_DM_FN double craw_dm_log_d(double x)
{
    const double ln2_hi = 0x1.62e42fee00000p-1;
    const double ln2_lo = 0x1.a39ef35793c76p-33;
    uint64_t u = craw_dm_bits(x);
    int64_t e = (int64_t)((u >> 52) & 0x7ff) - 1023;
    double m = craw_dm_from_bits((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    double f, s, z, rr, s2, logm, ed;

    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e = e + 1;
    }

    f = m - 1.0;                  /* exact: m is within a factor 2 of 1 */
    s = f / (2.0 + f);
    z = s * s;

    /* sum_{n>=1} z^n / (2n + 1) up to z^9 */
    rr = 0x1.af286bca1af28p-5;
    rr = __builtin_fma(rr, z, 0x1.e1e1e1e1e1e1ep-5);
    rr = __builtin_fma(rr, z, 0x1.1111111111111p-4);
    rr = __builtin_fma(rr, z, 0x1.3b13b13b13b14p-4);
    rr = __builtin_fma(rr, z, 0x1.745d1745d1746p-4);
    rr = __builtin_fma(rr, z, 0x1.c71c71c71c71cp-4);
    rr = __builtin_fma(rr, z, 0x1.2492492492492p-3);
    rr = __builtin_fma(rr, z, 0x1.999999999999ap-3);
    rr = __builtin_fma(rr, z, 0x1.5555555555555p-2);
    rr = rr * z;

    s2 = s + s;
    logm = __builtin_fma(s2, rr, s2);

    ed = (double)e;
    return ed * ln2_hi + __builtin_fma(ed, ln2_lo, logm);
}

/* expf with the reference's semantics: NaN in, NaN out; overflow to
 * +inf and underflow through the subnormals to +0 follow from the single
 * rounding of the double result. */
CRAW_DM_FN float craw_dm_expf(float x)
{
    float r;

    /* branch-free: a NaN is clamped like any argument, and its result
     * replaced at the end */
    r = (float)craw_dm_exp_d(CRAW_DM_CLAMP((double)x, -200.0, 200.0));
    return x != x ? x + x : r;
}

/* craw_dm_expf for an argument that is <= 0 and not a NaN (the weights of
 * the denoise and chroma filters): the same value, without the checks
 * that cannot fire there. */
CRAW_DM_FN float craw_dm_expf_nonpos(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    /* the clamp in float (one max; -200 is a float), the same value */
    return (float)craw_dm_exp_d((double)__builtin_fmaxf(x, -200.0f));
#else
    const double xd = (double)x;

    return (float)craw_dm_exp_d(xd < -200.0 ? -200.0 : xd);
#endif
}

/* exp(-max(a, 0)), the weight of the denoise filters, for any float a:
 * equal to craw_dm_expf(-(a > 0 ? a : 0)) for every a, a NaN included
 * (it gives 1). The argument is clamped at 200 in float (exp(-200) rounds
 * to 0 like any smaller value), which the device does with one max and
 * one min. */
CRAW_DM_FN float craw_dm_expf_negmax0(float a)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float c = __builtin_fminf(__builtin_fmaxf(a, 0.0f), 200.0f);
#else
    const float c = a > 0.0f ? (a < 200.0f ? a : 200.0f) : 0.0f;
#endif

    return (float)craw_dm_exp_d(-(double)c);
}

/* craw_dm_expf_negmax0 and craw_dm_expf_nonpos through the tables: the
 * same float for every argument (the clamps are theirs, and over the
 * clamped range [-200, 0] the table forms round like craw_dm_exp_d). */
CRAW_DM_FN float craw_dm_expf_negmax0_t256(float a, const double *tab)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float c = __builtin_fminf(__builtin_fmaxf(a, 0.0f), 200.0f);
#else
    const float c = a > 0.0f ? (a < 200.0f ? a : 200.0f) : 0.0f;
#endif

    return (float)craw_dm_exp_t256_d(-(double)c, tab);
}

CRAW_DM_FN float craw_dm_expf_negmax0_t1024(float a, const double *tab)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float c = __builtin_fminf(__builtin_fmaxf(a, 0.0f), 200.0f);
#else
    const float c = a > 0.0f ? (a < 200.0f ? a : 200.0f) : 0.0f;
#endif

    return (float)craw_dm_exp_t1024_d(-(double)c, tab);
}

CRAW_DM_FN float craw_dm_expf_nonpos_t1024(float x, const double *tab)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (float)craw_dm_exp_t1024_d((double)__builtin_fmaxf(x, -200.0f), tab);
#else
    const double xd = (double)x;

    return (float)craw_dm_exp_t1024_d(xd < -200.0 ? -200.0 : xd, tab);
#endif
}

/* logf: NaN for NaN and negatives, -inf at zero, +inf at +inf. Float
 * subnormals are normal doubles, so they need no special path. */
CRAW_DM_FN float craw_dm_logf(float x)
{
    if (x != x)
        return x + x;
    if (x < 0.0f)
        return __builtin_nanf("");
    if (x == 0.0f)
        return -__builtin_inff();
    if (x == __builtin_inff())
        return x;

    return (float)craw_dm_log_d((double)x);
}

/* powf for x >= 0 (the only bases the path raises): 0^y = 0 and inf^y =
 * inf for y > 0, 1^y = 1, NaN for a negative or NaN base. */
CRAW_DM_FN float craw_dm_powf(float x, float y)
{
    double t;

    if (x != x || y != y)
        return x + y;
    if (x < 0.0f)
        return __builtin_nanf("");
    if (x == 1.0f || y == 0.0f)
        return 1.0f;
    if (x == 0.0f)
        return y > 0.0f ? 0.0f : __builtin_inff();
    if (x == __builtin_inff())
        return y > 0.0f ? x : 0.0f;

    t = (double)y * craw_dm_log_d((double)x);
    t = t < -700.0 ? -700.0 : t;
    t = t > 700.0 ? 700.0 : t;

    return (float)craw_dm_exp_d(t);
}

#endif /* CRAW_DETMATH_H */
