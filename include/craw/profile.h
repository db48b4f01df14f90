/*
 * libcraw for MI355X: camera profile, white balance and colour spaces.
 *
 * Interface of include/craw/profile.h in cyanreg/libcraw (profile.h:89-246):
 * the same types, field order, enum values and functions, so that a
 * profile filled for the reference is a profile here.
 */
#ifndef CRAW_PROFILE_H
#define CRAW_PROFILE_H

#include <stdint.h>
#include <stdbool.h>

#include <craw/utils.h>
#include <craw/noise.h>

CRAW_BEGIN_DECLS

/* CFA phase, named by the 2x2 quad read row by row from even (x, y). Only
 * RGGB is developed; the pipeline refuses the others with ENOTSUP. */
enum CRAWCfa {
    CRAW_CFA_RGGB = 0,
    CRAW_CFA_BGGR,
    CRAW_CFA_GRBG,
    CRAW_CFA_GBRG,
};

/* Per-site arrays are indexed by (y & 1) * 2 + (x & 1). */
#define CRAW_SITE_TL 0
#define CRAW_SITE_TR 1
#define CRAW_SITE_BL 2
#define CRAW_SITE_BR 3

/* An illuminant on the profile's locus: Kelvin plus a tint across it. */
typedef struct CRAWIlluminant {
    float cct;
    float tint;
} CRAWIlluminant;

#define CRAW_PROFILE_MAX_KNOTS 16

typedef struct CRAWWbKnot {
    float cct;
    float wb[3];
} CRAWWbKnot;

/* Balanced camera RGB to XYZ D65, row major. */
typedef struct CRAWMatrixKnot {
    float cct;
    float m[3][3];
} CRAWMatrixKnot;

typedef struct CRAWProfile {
    enum CRAWCfa cfa;

    /* Per quad site, in UNORM units of the input samples. */
    float black_level[4];
    float white_level[4];

    float as_shot_wb[3];
    float as_shot_matrix[3][3];
    float as_shot_cct;

    float gain;

    float as_shot_ei;
    float exposure_index;

    CRAWWbKnot     wb_knots[CRAW_PROFILE_MAX_KNOTS];
    uint32_t       nb_wb_knots;
    CRAWMatrixKnot matrix_knots[CRAW_PROFILE_MAX_KNOTS];
    uint32_t       nb_matrix_knots;

    char name[64];

    CRAWNoiseKnot noise_knots[CRAW_PROFILE_MAX_NOISE_KNOTS];
    uint32_t      nb_noise_knots;
} CRAWProfile;

enum CRAWWbMode {
    CRAW_WB_AS_SHOT = 0,
    CRAW_WB_REBALANCE,
};

/* What a frame's colour stage uses, as evaluated from a profile. */
typedef struct CRAWColorParams {
    float wb[3];
    float cam_to_xyz[3][3];
    float gain;
} CRAWColorParams;

/* Output colour spaces, linear, scene referred. Same values as the
 * reference (profile.h:218-227). */
enum CRAWColorSpace {
    CRAW_CS_XYZ_D65 = 0,
    CRAW_CS_BT709_LINEAR,
    CRAW_CS_BT2020_LINEAR,
    CRAW_CS_P3D65_LINEAR,
    CRAW_CS_V_GAMUT_LINEAR,
    CRAW_CS_ACES_AP0_LINEAR,
    CRAW_CS_ACES_AP1_LINEAR,
    CRAW_CS_NB,
};

/* Identity profile: RGGB, black 0, white 1, unit gains and matrix. */
CRAW_API void craw_profile_init(CRAWProfile *p);

/* Levels, table order and matrix sanity; 0 or CRAW_ERROR(EINVAL). */
CRAW_API int craw_profile_validate(const CRAWProfile *p);

/* The frame's gains, matrix and gain (profile.c:198-252 of the reference):
 * as shot, or re-balanced to `illum` along the tables, linear in Kelvin. */
CRAW_API int craw_profile_eval(const CRAWProfile *p, enum CRAWWbMode mode,
                               const CRAWIlluminant *illum,
                               CRAWColorParams *out);

/* The point of the white balance locus a neutral surface's raw ratios
 * (R/G, B/G before any balance) lie from along the green direction, and
 * the signed distance as a tint (profile.c:367-389 of the reference): the
 * inverse of craw_profile_illuminant_to_neutral() wherever the table
 * reaches; a point past an end of the locus goes to that end. Needs at
 * least two white balance knots (CRAW_ERROR(ENOTSUP)). */
CRAW_API int craw_profile_neutral_to_illuminant(const CRAWProfile *p,
                                                float raw_rg, float raw_bg,
                                                CRAWIlluminant *out);

/* The raw ratios a neutral surface gives under `illum`
 * (profile.c:391-415 of the reference). */
CRAW_API int craw_profile_illuminant_to_neutral(const CRAWProfile *p,
                                                const CRAWIlluminant *illum,
                                                float *raw_rg, float *raw_bg);

/* The profile's noise table at `iso` (as_shot_ei when not positive);
 * CRAW_ERROR(ENOTSUP) without a table. */
CRAW_API int craw_profile_noise_eval(const CRAWProfile *p, float iso,
                                     CRAWNoiseModel *out);

/* Row-major XYZ (D65) to `cs` and back, derived in double from the
 * primaries, with a Bradford adaptation for the ACES white
 * (colorspace.c:183-253 of the reference). */
CRAW_API int craw_colorspace_from_xyz_d65(enum CRAWColorSpace cs, float m[3][3]);
CRAW_API int craw_colorspace_to_xyz_d65(enum CRAWColorSpace cs, float m[3][3]);

/* Planckian locus (Krystek 1985) and Bradford adaptation. */
CRAW_API int craw_planckian_xy(float cct, float xy[2]);
CRAW_API int craw_bradford_adaptation(const float src[2], const float dst[2],
                                      float m[3][3]);

CRAW_END_DECLS

#endif /* CRAW_PROFILE_H */
