/*
 * The oracle: a CPU restatement of cyanreg/libcraw's float32 CPU reference
 * (lib/ref) for the raw development path, used as the checker of this
 * repository's HIP kernels.
 *
 * TEST INFRASTRUCTURE. Only the tests, the smoke check and the benchmark's
 * CPU baseline may load it, and only to check or to time it; the library
 * never calls it.
 *
 * oracle_core.c is compiled three times, giving three sets of the same
 * functions under three prefixes:
 *
 *   oracle_f32l_*  float, glibc's expf/logf/powf: lib/ref as written;
 *   oracle_f32d_*  float, the deterministic expf/logf/powf of detmath.h
 *                  that the kernels also use: what the GPU must reproduce
 *                  bit for bit;
 *   oracle_f64_*   double throughout: the FP64 rerun that error statements
 *                  are measured against.
 *
 * The two float modes differ only where glibc's expf/logf/powf misround
 * (about one argument in 10^4), which the tests count. oracle_host.c holds
 * the mode-independent host math (profile evaluation, colour spaces, gamut
 * compression constants, noise model) as oracle_* functions.
 *
 * Parity pins (each function's header in oracle_core.c names its
 * reference file:line):
 *   prep, bilinear, MHC, RCD: the known-answer tests of
 *       tests/test_ref_debayer.c (constants, ramps, directional step,
 *       hostile inputs, 4x4 frame, prep arithmetic) and brute-force
 *       restatements on small frames in tests/;
 *   GAT and inverse: the closed forms of the reference's DESIGN.md
 *       (c = 3a^2/8 + b, T(u) = (2/a) sqrt(a u + c), its tangent below 0,
 *       the algebraic inverse) at several (a, b, u), u < 0 included;
 *   NLM: the weight exp(-max(d2/9/h^2 - 2, 0)) against its closed form;
 *       noise-free two-level planes whose output is the count-weighted
 *       mean of the two levels; a numpy direct summation over the 7x7
 *       window on noisy frames; constant planes are fixed points, h -> 0
 *       is the identity, flat and checkerboard noise ratios;
 *   temporal: a uniform offset D between the previous level and this
 *       frame's gives alpha_max exp(-max(D^2/ht^2 - 2, 0)), the blend and
 *       the stored variance; with a history, a two-level plane of two
 *       known variances gives the weight exp(-max(d2/9/h^2 - (vp + vq),
 *       0) / vp); static-scene variance tends to (1 - a) / (1 + a), first
 *       frame equals the spatial filter;
 *   chroma: a numpy statement of the reference's definition (block means,
 *       9x9 bilateral with sigma_s 2, K = Kc = 0.35, eps 0.02, mirrored,
 *       bilinear up) on a G edge crossing a colour edge; G untouched, flat
 *       colour preserved, blotch reduction;
 *   each of those closed-form checks is shown, by tests/test_oracle_pins.py,
 *       to fail on a variant of this file with one constant planted wrong
 *       (ORACLE_MUTANT);
 *   false colour: each sweep against scipy's median filter and a double
 *       box mean (mirror mode), G and alpha untouched, and the reference's
 *       spike known answers (tests/test_pipeline.c:3994-4050: a site spike's
 *       colour at most 65% after one pass through bilinear, 10% after five);
 *   AWB histogram: oracle/_ref (the reference's own lib/awb.c, built from
 *       its sources) on random and structured inputs, plus test_awb.c's
 *       flat-neutral, skip and partial-block cases;
 *   colour spaces, profile evaluation, gamut constants, noise model:
 *       oracle/_ref (lib/colorspace.c, profile.c, gamut.c, noise.c).
 * lib/ref itself cannot be built here (it needs the Vulkan headers and a
 * generated version header), so no function of lib/ref's C files is compared
 * against the reference's own binary.
 */
#ifndef CRAW_ORACLE_H
#define CRAW_ORACLE_H

#include <stdint.h>
#include <stddef.h>
#include <craw/pipeline.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Host math, mode independent (oracle_host.c). */
void oracle_profile_init(CRAWProfile *p);
int oracle_profile_eval(const CRAWProfile *p, enum CRAWWbMode mode,
                        const CRAWIlluminant *illum, CRAWColorParams *out);
int oracle_noise_model_eval(const CRAWNoiseKnot *knots, uint32_t nb_knots,
                            float iso, CRAWNoiseModel *out);
int oracle_profile_noise_eval(const CRAWProfile *p, float iso, CRAWNoiseModel *out);
int oracle_noise_model(const CRAWFrameParams *fp, CRAWNoiseModel *out);
int oracle_colorspace_to_xyz_d(enum CRAWColorSpace cs, double out[3][3]);
int oracle_colorspace_from_xyz_d(enum CRAWColorSpace cs, double out[3][3]);
int oracle_colorspace_from_xyz_d65(enum CRAWColorSpace cs, float m[3][3]);
int oracle_colorspace_to_xyz_d65(enum CRAWColorSpace cs, float m[3][3]);
int oracle_planckian_xy(float cct, float xy[2]);
int oracle_bradford_adaptation(const float src[2], const float dst[2], float m[3][3]);
/* ACES 1.3 RGC constants: thr[3], scl[3], pwr, inv_pwr (float and double). */
void oracle_gamut_params(float thr[3], float scl[3], float *pwr, float *inv_pwr);
void oracle_gamut_params_d(double thr[3], double scl[3], double *pwr, double *inv_pwr);

/* One mode's set (oracle_core.c). `real` is float for f32l/f32d and
 * double for f64. */
#define ORACLE_DECLARE(P, real)                                                      \
    void P##set_threads(int n);                                                      \
    void P##prep(const uint16_t *bayer, uint32_t w, uint32_t h, size_t stride,       \
                 const CRAWProfile *p, const float wb[3], real *v);                  \
    int P##debayer_mosaic(const real *v, uint32_t w, uint32_t h, int method,         \
                          real *rgb);                                                \
    int P##debayer_wb(const uint16_t *bayer, uint32_t w, uint32_t h, size_t stride,  \
                      const CRAWProfile *p, const float wb[3], int method, real *rgb);\
    int P##denoise_temporal(const uint16_t *bayer, uint32_t w, uint32_t h,           \
                            size_t stride, const CRAWProfile *p,                     \
                            const CRAWNoiseModel *model, float strength,             \
                            float alpha_max, const real *prev, real *next,           \
                            uint16_t *out, size_t out_stride, real *out_level);      \
    int P##chroma_denoise(real *rgb, uint32_t w, uint32_t h, float strength);        \
    int P##false_color(real *rgb, uint32_t w, uint32_t h, uint32_t passes);          \
    int P##highlights(const uint16_t *bayer, uint32_t w, uint32_t h, size_t stride,  \
                      const CRAWProfile *p, const float wb[3], real *out);           \
    int P##color(real *rgb, uint32_t w, uint32_t h, const CRAWFrameParams *fp,       \
                 int out_space);                                                     \
    void P##gamut_compress(real *rgb, size_t n);                                     \
    int P##awb_histogram(const real *rgb, uint32_t w, uint32_t h, size_t stride,     \
                         int channels, const float wb[3], uint32_t *count,           \
                         uint32_t *weight, uint32_t *bright);                        \
    int P##process(const uint16_t *bayer, uint32_t w, uint32_t h, size_t stride,     \
                   const CRAWPipelineParams *pp, const CRAWFrameParams *fp,          \
                   const real *prev, real *next, uint16_t *clean, real *inter,       \
                   uint32_t *count, uint32_t *weight, uint32_t *bright, real *out);  \
    /* test hooks: the denoise and chroma weights one value at a time */            \
    void P##gat_hook(float a, float b, const real *u, real *t, const real *tin,      \
                     real *uout, size_t n, real consts[7]);                          \
    real P##nlm_weight_hook(real d2, real inv_h2, real vc, real vq);                 \
    real P##alpha_weight_hook(real d2, real alpha_max, real inv_h2);                 \
    void P##blend_hook(real tc, real tp, real a, real v_s, real out[3]);             \
    real P##cd_weight_hook(int dx, int dy, real g, real dg, real dd0, real dd1);

ORACLE_DECLARE(oracle_f32l_, float)
ORACLE_DECLARE(oracle_f32d_, float)
ORACLE_DECLARE(oracle_f64_, double)

#ifdef __cplusplus
}
#endif

#endif /* CRAW_ORACLE_H */
