/*
 * libcraw for MI355X: planning one frame's recording on the host, shared
 * by the plain pipeline and the batched form; temporal state; profiling.
 *
 * craw_frame_plan() is the host phase of lib/pipeline.c:1716-1931 of
 * cyanreg/libcraw for this path: validation, the frame's colour
 * parameters, the prep and denoise constants (pipeline.c:1360-1467), the
 * temporal read/write set selection (1784-1818), the output matrix and the
 * gamut compression constants (1870-1931). It fails without side effects;
 * craw_frame_commit() moves the temporal indices once the frame's work has
 * been enqueued (2303-2313).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define DN_SEARCH_R 3
#define DN_PATCH_R  1
#define DN_GUARD_LO 0.90f
#define DN_GUARD_HI 1.00f
#define CD_MIN_SIZE 20

static const uint32_t unsupported_flags =
    CRAW_PIPELINE_FLAG_LENS | CRAW_PIPELINE_FLAG_DEFRINGE |
    CRAW_PIPELINE_FLAG_SHARPEN;

static const uint32_t known_flags =
    CRAW_PIPELINE_FLAG_AWB_HISTOGRAM | CRAW_PIPELINE_FLAG_DENOISE |
    CRAW_PIPELINE_FLAG_CHROMA_DENOISE | CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE |
    CRAW_PIPELINE_FLAG_PROFILE | CRAW_PIPELINE_FLAG_NLM_FP16 | CRAW_PIPELINE_FLAG_NLM_FAST |
    CRAW_PIPELINE_FLAG_FALSE_COLOR | CRAW_PIPELINE_FLAG_HIGHLIGHT_RECONSTRUCTION |
    unsupported_flags;

int craw_stage_config(CRAWContext *ctx, uint32_t flags, enum CRAWDebayer debayer,
                      enum CRAWColorSpace out_space, enum CRAWFormat out_format,
                      CRAWStageConfig *cfg)
{
    if (flags & unsupported_flags) {
        craw_log(ctx, CRAW_LOG_ERROR, "Pipeline flags 0x%x name stages this library does not "
                 "have\n", flags & unsupported_flags);
        return CRAW_ERROR(ENOTSUP);
    }
    if (flags & ~known_flags)
        return CRAW_ERROR(EINVAL);
    if (out_format != CRAW_FORMAT_R16G16B16A16_SFLOAT &&
        out_format != CRAW_FORMAT_R32G32B32A32_SFLOAT) {
        craw_log(ctx, CRAW_LOG_ERROR, "Output format must be RGBA16F or RGBA32F\n");
        return CRAW_ERROR(EINVAL);
    }
    if ((unsigned)debayer >= CRAW_DEBAYER_NB || (unsigned)out_space >= CRAW_CS_NB)
        return CRAW_ERROR(EINVAL);
    cfg->debayer = debayer;
    cfg->out_space = out_space;
    cfg->out_format = out_format;
    cfg->awb = !!(flags & CRAW_PIPELINE_FLAG_AWB_HISTOGRAM);
    cfg->has_denoise = !!(flags & CRAW_PIPELINE_FLAG_DENOISE);
    cfg->has_temporal = !!(flags & CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE);
    cfg->has_chroma = !!(flags & CRAW_PIPELINE_FLAG_CHROMA_DENOISE);
    cfg->has_fc = !!(flags & CRAW_PIPELINE_FLAG_FALSE_COLOR);
    cfg->has_hl = !!(flags & CRAW_PIPELINE_FLAG_HIGHLIGHT_RECONSTRUCTION);
    cfg->fp16 = !!(flags & CRAW_PIPELINE_FLAG_NLM_FP16);
    cfg->fast = !!(flags & CRAW_PIPELINE_FLAG_NLM_FAST);
    return 0;
}

static int fmt_bytes(enum CRAWFormat f)
{
    switch (f) {
    case CRAW_FORMAT_R16_UNORM:           return 2;
    case CRAW_FORMAT_R16G16B16A16_SFLOAT: return 8;
    case CRAW_FORMAT_R32G32B32A32_SFLOAT: return 16;
    default:                              return 0;
    }
}

static bool overlaps(const CRAWImage *a, const CRAWImage *b)
{
    uintptr_t a0 = (uintptr_t)a->data, a1 = a0 + a->pitch * a->height;
    uintptr_t b0 = (uintptr_t)b->data, b1 = b0 + b->pitch * b->height;

    return a0 < b1 && b0 < a1;
}

int craw_hl_levels(uint32_t w, uint32_t h, int lw[KHL_LEVELS], int lh[KHL_LEVELS],
                   size_t off[KHL_LEVELS], size_t *total)
{
    /* lib/pipeline.c:885-912: level 0 a texel per 4x4 block, halved
     * (rounding up) to 1x1 */
    uint32_t x = (w + 3) / 4, y = (h + 3) / 4;
    size_t at = 0;
    int n = 0;

    while (1) {
        if (n == KHL_LEVELS)
            return CRAW_ERROR(EINVAL);
        lw[n] = (int)x;
        lh[n] = (int)y;
        off[n] = at;
        at += (size_t)x * y;
        n++;
        if (x == 1 && y == 1)
            break;
        x = (x + 1) / 2;
        y = (y + 1) / 2;
    }
    *total = at;
    return n;
}

/* Whether the debayer's quotients stay where kdiv_rn divides exactly
 * (divisor in [2^-40, 2^40], dividend +0 or in [2^-80, 2^40]; the
 * dividends are differences and sums of non-negative values, never -0).
 * With every
 * site's black level 0 or in [2^-16, 1], range in [2^-16, 2] and gain in
 * [2^-4, 2^4], a prepped sample is 0 or of magnitude in [2^-45, 2^20]: raw
 * / 65535 - black is 0 or at least 2^-40 (two floats of at least 2^-16,
 * or raw 0), the range at most 2 and the gain at least 2^-4. Every such
 * float is a multiple of 2^-68, and so is every sum and difference the
 * debayer forms from them with power-of-two weights (a rounded multiple of
 * 2^-68 above it in magnitude stays one); the low pass divides by 16 and
 * the gradient means multiply by a ninth, so a nonzero dividend is at least
 * 2^-76, while no value exceeds 2^26. A rebuilt R16F mosaic (the highlight
 * stage) holds multiples of 2^-24 below 2^16, inside the same bounds. The
 * precision form's prep divides raw / 65535 - black (0, never -0, or at
 * least 2^-40 in magnitude, at most 2) by the range through kdiv_rn under
 * the same condition. */
static bool prep_kdiv_ok(const KPrep &p)
{
    for (int s = 0; s < 4; s++) {
        const float b = p.black[s], r = p.range[s], g = p.wb[s];

        if (!(b == 0.0f || (b >= 0x1p-16f && b <= 1.0f)))
            return false;
        if (!(r >= 0x1p-16f && r <= 2.0f) || !(g >= 0x1p-4f && g <= 0x1p4f))
            return false;
    }
    return true;
}

int craw_frame_plan(CRAWContext *ctx, const CRAWStageConfig &cfg, const CRAWFrameBuffers &buf,
                    uint32_t w, uint32_t h, const CRAWImage *src, const CRAWImage *dst,
                    const CRAWFrameParams *fp, CRAWFramePlan *plan)
{
    CRAWColorParams frame;
    CRAWGamutParams gp;
    CRAWTemporalState *ts = buf.ts;
    enum CRAWColorSpace out_space;
    float m_out[3][3];
    bool spatial, temporal, replay;
    int t_read, t_write, ret;

    if (!src || !dst || !fp || !fp->profile || !plan)
        return CRAW_ERROR(EINVAL);
    memset(plan, 0, sizeof(*plan));

    if (src->format != CRAW_FORMAT_R16_UNORM) {
        craw_log(ctx, CRAW_LOG_ERROR, "Source must be CRAW_FORMAT_R16_UNORM (got %d)\n",
                 (int)src->format);
        return CRAW_ERROR(EINVAL);
    }
    if (dst->format != cfg.out_format) {
        craw_log(ctx, CRAW_LOG_ERROR, "Destination format %d does not match the pipeline's "
                 "%d\n", (int)dst->format, (int)cfg.out_format);
        return CRAW_ERROR(EINVAL);
    }
    if (src->width != w || src->height != h || dst->width != w || dst->height != h) {
        craw_log(ctx, CRAW_LOG_ERROR, "Frame size mismatch: src %ux%u dst %ux%u, expected "
                 "%ux%u\n", src->width, src->height, dst->width, dst->height, w, h);
        return CRAW_ERROR(EINVAL);
    }
    if (!w || !h || (w & 1) || (h & 1) || w < 8 || h < 8)
        return CRAW_ERROR(EINVAL);
    if (!src->data || !dst->data)
        return CRAW_ERROR(EINVAL);
    if (src->pitch < (size_t)w * 2 || (src->pitch & 1) ||
        dst->pitch < (size_t)w * fmt_bytes(dst->format) ||
        (dst->pitch % (size_t)fmt_bytes(dst->format)) ||
        ((uintptr_t)dst->data % (size_t)fmt_bytes(dst->format)) || ((uintptr_t)src->data & 1)) {
        craw_log(ctx, CRAW_LOG_ERROR, "Image pitch or alignment does not fit the format\n");
        return CRAW_ERROR(EINVAL);
    }
    if (overlaps(src, dst)) {
        craw_log(ctx, CRAW_LOG_ERROR, "Source and destination must not alias\n");
        return CRAW_ERROR(EINVAL);
    }
    if (fp->profile->cfa != CRAW_CFA_RGGB)
        return CRAW_ERROR(ENOTSUP);

    /* The frame's gains go into the prep, so that the clip keeps blown
     * highlights neutral under the chosen illuminant. */
    ret = craw_profile_eval(fp->profile, fp->wb_mode, &fp->illuminant, &frame);
    if (ret < 0)
        return ret;
    for (int s = 0; s < 4; s++) {
        static const int chan[4] = { 0, 1, 1, 2 };
        float range = fp->profile->white_level[s] - fp->profile->black_level[s];

        if (!(range > 0.0f)) {
            craw_log(ctx, CRAW_LOG_ERROR, "Profile site %d has white level %f <= black level "
                     "%f\n", s, fp->profile->white_level[s], fp->profile->black_level[s]);
            return CRAW_ERROR(EINVAL);
        }
        plan->f.prep.black[s] = fp->profile->black_level[s];
        plan->f.prep.range[s] = range;
        plan->f.prep.rrange[s] = 1.0 / (double)range;
        plan->f.prep.wb[s] = frame.wb[chan[s]];
    }
    plan->f.prep.fdiv = cfg.fp16 ? 1 : 0;
    plan->f.prep.kdiv = prep_kdiv_ok(plan->f.prep) ? 1 : 0;

    /* Denoise: which filters, which state sets. */
    spatial = cfg.has_denoise && fp->denoise_strength > 0.0f;
    temporal = cfg.has_temporal && ts && fp->denoise_temporal_strength > 0.0f;
    if (temporal && !(fp->denoise_temporal_strength <= 1.0f)) {
        craw_log(ctx, CRAW_LOG_ERROR, "denoise_temporal_strength = %g; at most 1.0\n",
                 (double)fp->denoise_temporal_strength);
        return CRAW_ERROR(EINVAL);
    }
    if (temporal && (ts->width != w || ts->height != h)) {
        craw_log(ctx, CRAW_LOG_ERROR, "Temporal state is %ux%u, frame %ux%u\n", ts->width,
                 ts->height, w, h);
        return CRAW_ERROR(EINVAL);
    }
    replay = fp->denoise_temporal_replay;
    t_read = t_write = -1;
    if (ts) {
        t_read = replay ? ts->prev : ts->cur;
        if (fp->flags & CRAW_FRAME_FLAG_TEMPORAL_RESET)
            t_read = -1;
        if (temporal)
            t_write = replay && ts->cur >= 0 ? ts->cur : (t_read == 0 ? 1 : 0);
    }
    plan->t_read = t_read;
    plan->t_write = t_write;
    plan->temporal = temporal;
    plan->replay = replay;
    /* lib/pipeline.c:2308: only a recording of a pipeline (or batch) built
     * with the temporal pass moves a state; one without it leaves an
     * attached state as it was */
    plan->ts = cfg.has_temporal ? ts : NULL;

    if (spatial || temporal) {
        const uint32_t dn_min = 2 * (DN_SEARCH_R + DN_PATCH_R + 1 + (cfg.has_temporal ? 1 : 0));
        CRAWNoiseModel nm;
        float a, b, c, sc, hh;

        if (w < dn_min || h < dn_min || !buf.clean)
            return CRAW_ERROR(EINVAL);
        if (fp->noise_override.a > 0.0f) {
            nm = fp->noise_override;
        } else {
            ret = craw_profile_noise_eval(fp->profile, fp->profile->as_shot_ei, &nm);
            if (ret < 0) {
                craw_log(ctx, CRAW_LOG_ERROR, "Denoising needs a noise model: the profile `%s` "
                         "has no table and noise_override was not set\n",
                         fp->profile->name[0] ? fp->profile->name : "(unnamed)");
                return ret;
            }
        }
        a = nm.a;
        b = nm.b > 0.0f ? nm.b : 0.0f;
        if (!(a > 0.0f)) {
            craw_log(ctx, CRAW_LOG_ERROR, "Noise model a = %g is not positive\n", (double)a);
            return CRAW_ERROR(EINVAL);
        }
        /* lib/pipeline.c:1431-1467 and lib/ref/ref_denoise.c:28-37 */
        c = 3.0f * a * a / 8.0f + b;
        sc = sqrtf(c);
        hh = spatial ? fp->denoise_strength : 1.0f;
        plan->f.dn.w2 = (int)(w / 2);
        plan->f.dn.h2 = (int)(h / 2);
        plan->f.dn.gat_a = a;
        plan->f.dn.gat_c = c;
        plan->f.dn.gat_sc = sc;
        plan->f.dn.gat_fwd = 2.0f / a;
        plan->f.dn.gat_inv = a / 4.0f;
        plan->f.dn.gat_off = c / a;
        plan->f.dn.gat_t0 = plan->f.dn.gat_fwd * sc;
        plan->f.dn.inv_h2 = 1.0f / (hh * hh);
        plan->f.dn.guard_lo = DN_GUARD_LO;
        plan->f.dn.guard_inv = 1.0f / (DN_GUARD_HI - DN_GUARD_LO);
        plan->f.dn.alpha_max = temporal ? fp->denoise_temporal_strength : 0.0f;
        for (int s = 0; s < 4; s++) {
            float range = fp->profile->white_level[s] - fp->profile->black_level[s];
            plan->f.dn.black[s] = fp->profile->black_level[s];
            plan->f.dn.range[s] = range;
            plan->f.dn.inv_range[s] = 1.0f / range;
        }
        plan->f.dn.src_pitch = (int)(src->pitch / 2);
        plan->f.dn.dst_pitch = (int)w;
        plan->dn_mode = (spatial ? KDN_SPATIAL : 0) | (temporal ? KDN_TEMPORAL : 0) |
                        (temporal && t_read >= 0 ? KDN_HISTORY : 0);
        if (temporal) {
            if (t_read >= 0) {
                plan->f.st.prev_level = ts->level[t_read];
                plan->f.st.prev_var = ts->var[t_read];
            }
            plan->f.st.next_level = ts->level[t_write];
            plan->f.st.next_var = ts->var[t_write];
        }
    }

    plan->chroma = cfg.has_chroma && fp->denoise_chroma_strength > 0.0f;
    if (plan->chroma) {
        if (!(fp->denoise_chroma_strength <= 1.0f)) {
            craw_log(ctx, CRAW_LOG_ERROR, "denoise_chroma_strength = %g; at most 1.0\n",
                     (double)fp->denoise_chroma_strength);
            return CRAW_ERROR(EINVAL);
        }
        if (w < CD_MIN_SIZE || h < CD_MIN_SIZE || !buf.inter_rgb)
            return CRAW_ERROR(EINVAL);
    }
    plan->f.chroma_strength = fp->denoise_chroma_strength;
    plan->f.inter.rgb = buf.inter_rgb;
    plan->f.inter.down = buf.down;
    plan->f.inter.blur = buf.blur;
    plan->f.inter.wq = (int)((w + 3) / 4);
    plan->f.inter.hq = (int)((h + 3) / 4);
    plan->f.inter.rgb_out = buf.inter_rgb;
    plan->f.inter.means = 1;
    plan->f.inter.up = 1;

    /* lib/pipeline.c:1820-1836: zero passes leave the frame alone, with or
     * without the stage; more than the stage has are refused */
    if (cfg.has_fc && fp->false_color_passes > 0) {
        if (fp->false_color_passes > CRAW_FALSE_COLOR_MAX_PASSES) {
            craw_log(ctx, CRAW_LOG_ERROR, "false_color_passes = %u; at most %u passes are "
                     "supported\n", fp->false_color_passes,
                     (unsigned)CRAW_FALSE_COLOR_MAX_PASSES);
            return CRAW_ERROR(EINVAL);
        }
        if (!buf.inter_rgb || (plan->chroma && !buf.inter_rgb2))
            return CRAW_ERROR(EINVAL);
        plan->fc_passes = (int)fp->false_color_passes;
        /* the debayer writes its planes (no quarter means) for the false
         * colour stage, which writes the chroma stage's inputs (the planes
         * and the means) where the chroma stage reads them */
        plan->fc_to = plan->f.inter;
        if (plan->chroma)
            plan->f.inter.rgb_out = buf.inter_rgb2;
        plan->f.inter.means = 0;
    }

    /* Colour: the frame's matrix and scale, then XYZ to the output space,
     * as lib/ref/ref_color.c:10-76 applies them; gamut compression for
     * ACEScg unless vetoed. */
    out_space = fp->out_space_set ? fp->out_space : cfg.out_space;
    ret = craw_colorspace_from_xyz_d65(out_space, m_out);
    if (ret < 0) {
        craw_log(ctx, CRAW_LOG_ERROR, "Unknown output colour space %d\n", (int)out_space);
        return ret;
    }
    memcpy(plan->f.col.cam, frame.cam_to_xyz, sizeof(plan->f.col.cam));
    memcpy(plan->f.col.out, m_out, sizeof(plan->f.col.out));
    plan->f.col.scale = frame.gain * exp2f(fp->exposure_stops);
    craw_gamut_params(&gp);
    memcpy(plan->f.col.thr, gp.thr, sizeof(gp.thr));
    memcpy(plan->f.col.scl, gp.scl, sizeof(gp.scl));
    for (int i = 0; i < 3; i++)
        plan->f.col.thr_lo[i] = kgamut_screen_bound(gp.thr[i]);
    plan->f.col.pwr = gp.pwr;
    plan->f.col.inv_pwr = gp.inv_pwr;
    plan->f.col.flags = (fp->flags & CRAW_FRAME_FLAG_CLAMP_NEGATIVE) ? KCOLOR_CLAMP_NEGATIVE : 0u;
    if (out_space == CRAW_CS_ACES_AP1_LINEAR && !(fp->flags & CRAW_FRAME_FLAG_NO_GAMUT_COMPRESSION))
        plan->f.col.flags |= KCOLOR_GAMUT_COMPRESS;
    /* the precision form: no double arithmetic on the device, the colour,
     * chroma, histogram and highlight stages' transcendentals in float */
    if (cfg.fp16)
        plan->f.col.flags |= KCOLOR_FLOAT_MATH;

    /* RCD to the output with the gamut compression: the planes, then the
     * colour stage from them (the same values: the planes hold the floats
     * the fused tail would colour, and the block sums keep their order) */
    plan->split_color = cfg.debayer == CRAW_DEBAYER_RCD && !plan->chroma && !plan->fc_passes &&
                        (plan->f.col.flags & KCOLOR_GAMUT_COMPRESS) && buf.inter_rgb;
    if (plan->split_color) {
        plan->f.inter.rgb_out = buf.inter_rgb;
        plan->f.inter.means = 0;
        plan->f.inter.up = 0;
    }

    plan->f.out.dst = dst->data;
    plan->f.out.pitch = dst->pitch;
    plan->f.out.half = dst->format == CRAW_FORMAT_R16G16B16A16_SFLOAT;
    plan->f.out.hist = cfg.awb ? buf.hist : NULL;
    plan->f.out.samples = cfg.awb ? buf.awb_samples : NULL;
    if (cfg.awb && (!buf.hist || !buf.awb_samples))
        return CRAW_ERROR(EINVAL);
    plan->f.out.awb_wb0 = frame.wb[0];
    plan->f.out.awb_wb2 = frame.wb[2];
    plan->f.out.float_log = cfg.fp16 ? 1 : 0;

    plan->f.W = (int)w;
    plan->f.H = (int)h;
    plan->f.method = (int)cfg.debayer;
    if (plan->dn_mode) {
        /* the denoise reads the caller's mosaic and writes the frame's own
         * buffer, which the debayer then reads */
        plan->f.dn_src = (const uint16_t *)src->data;
        plan->f.dn_dst = buf.clean;
        plan->f.src = buf.clean;
        plan->f.src_pitch = (int)w;
    } else {
        plan->f.src = (const uint16_t *)src->data;
        plan->f.src_pitch = (int)(src->pitch / 2);
    }

    /* lib/pipeline.c:2047-2062: the rebuilt mosaic (balanced, unclipped,
     * R16F) replaces the one the debayer would read, whose prep becomes the
     * identity */
    plan->hl = cfg.has_hl && (fp->flags & CRAW_FRAME_FLAG_HIGHLIGHT_RECONSTRUCTION);
    if (plan->hl) {
        KHighlight &k = plan->hl_k;
        size_t off[KHL_LEVELS], total;

        if (!buf.hl_mosaic || !buf.hl_levels)
            return CRAW_ERROR(EINVAL);
        k.nl = craw_hl_levels(w, h, k.lw, k.lh, off, &total);
        if (k.nl < 0)
            return k.nl;
        for (int l = 0; l < k.nl; l++)
            k.level[l] = buf.hl_levels + off[l];
        k.src = plan->f.src;
        k.src_pitch = plan->f.src_pitch;
        k.mosaic = buf.hl_mosaic;
        k.W = (int)w;
        k.H = (int)h;
        for (int s = 0; s < 4; s++) {
            k.black[s] = fp->profile->black_level[s];
            k.inv_range[s] = 1.0f / (fp->profile->white_level[s] - fp->profile->black_level[s]);
        }
        for (int c = 0; c < 3; c++)
            k.wb[c] = frame.wb[c];
        k.float_math = cfg.fp16 ? 1 : 0;
        plan->f.src = buf.hl_mosaic;
        plan->f.src_pitch = (int)w;
        plan->f.prep.half = 1;
    }
    return 0;
}

void craw_frame_commit(const CRAWFramePlan *plan)
{
    CRAWTemporalState *ts = plan->ts;

    /* lib/pipeline.c:2303-2313: only a recording that succeeded moves the
     * state; a next frame without the pass leaves nothing for the frame
     * after it, a replay without it leaves the state as it was. */
    if (!ts)
        return;
    ts->prev = plan->t_read;
    if (plan->temporal)
        ts->cur = plan->t_write;
    else if (!plan->replay)
        ts->cur = -1;
}

/* ---------------------------------------------------------- temporal */

int craw_temporal_alloc(CRAWContext *ctx, uint32_t w, uint32_t h, CRAWTemporalState *ts)
{
    memset(ts, 0, sizeof(*ts));
    ts->ctx = ctx;
    ts->width = w;
    ts->height = h;
    ts->prev = ts->cur = -1;
    for (int i = 0; i < 2; i++) {
        if (hipMalloc((void **)&ts->level[i], (size_t)w * h * sizeof(float)) != hipSuccess ||
            hipMalloc((void **)&ts->var[i], (size_t)w * h) != hipSuccess) {
            craw_temporal_release(ts);
            return CRAW_ERROR(ENOMEM);
        }
    }
    return 0;
}

void craw_temporal_release(CRAWTemporalState *ts)
{
    for (int i = 0; i < 2; i++) {
        if (ts->level[i])
            CRAW_IGNORE(hipFree(ts->level[i]));
        if (ts->var[i])
            CRAW_IGNORE(hipFree(ts->var[i]));
        ts->level[i] = NULL;
        ts->var[i] = NULL;
    }
}

extern "C" {

CRAW_API int craw_temporal_state_create(CRAWContext *ctx, uint32_t width, uint32_t height,
                                        CRAWTemporalState **out)
{
    CRAWTemporalState *ts;
    int ret;

    if (!ctx || !out || !width || !height || (width & 1) || (height & 1))
        return CRAW_ERROR(EINVAL);
    if (hipSetDevice(ctx->device) != hipSuccess)
        return CRAW_ERROR(ENODEV);
    ts = (CRAWTemporalState *)calloc(1, sizeof(*ts));
    if (!ts)
        return CRAW_ERROR(ENOMEM);
    ret = craw_temporal_alloc(ctx, width, height, ts);
    if (ret < 0) {
        free(ts);
        return ret;
    }
    *out = ts;
    return 0;
}

CRAW_API void craw_temporal_state_free(CRAWTemporalState **ts)
{
    if (!ts || !*ts)
        return;
    CRAW_IGNORE(hipSetDevice((*ts)->ctx->device));
    craw_temporal_release(*ts);
    free(*ts);
    *ts = NULL;
}

CRAW_API void craw_temporal_state_reset(CRAWTemporalState *ts)
{
    if (ts)
        ts->prev = ts->cur = -1;
}

} /* extern "C" */

/* ---------------------------------------------------------- profiling */

int craw_profiler_init(CRAWProfiler *pr)
{
    memset(pr, 0, sizeof(*pr));
    for (int s = 0; s < CRAW_PIPELINE_PROFILE_SLOTS; s++)
        for (int i = 0; i < CRAW_MAX_STAGES; i++)
            for (int j = 0; j < 2; j++)
                if (hipEventCreate(&pr->ev[s][i][j]) != hipSuccess) {
                    pr->ev[s][i][j] = NULL;
                    craw_profiler_free(pr);
                    return CRAW_ERROR(ENOMEM);
                }
    pr->on = true;
    return 0;
}

void craw_profiler_free(CRAWProfiler *pr)
{
    for (int s = 0; s < CRAW_PIPELINE_PROFILE_SLOTS; s++)
        for (int i = 0; i < CRAW_MAX_STAGES; i++)
            for (int j = 0; j < 2; j++)
                if (pr->ev[s][i][j])
                    CRAW_IGNORE(hipEventDestroy(pr->ev[s][i][j]));
    memset(pr, 0, sizeof(*pr));
}

CRAWProfRun craw_profiler_start(CRAWProfiler *pr, hipStream_t s)
{
    CRAWProfRun r = { pr, s, (uint32_t)(pr->recorded % CRAW_PIPELINE_PROFILE_SLOTS), 0 };
    return r;
}

/* The open stage's events, handed to its kernel launches (KLAUNCH). */
static thread_local hipEvent_t tl_start, tl_stop;
static thread_local bool tl_used;

void craw_launch_events(hipEvent_t *start, hipEvent_t *stop)
{
    *start = tl_start;
    *stop = tl_stop;
    if (tl_stop) {
        tl_used = true;
        tl_start = NULL;
    }
}

void craw_profiler_begin(CRAWProfRun &r, const char *name, bool kernels)
{
    if (!r.pr->on || r.n >= CRAW_MAX_STAGES)
        return;
    r.pr->name[r.slot][r.n] = name;
    if (kernels) {
        tl_start = r.pr->ev[r.slot][r.n][0];
        tl_stop = r.pr->ev[r.slot][r.n][1];
        tl_used = false;
    } else {
        CRAW_IGNORE(hipEventRecord(r.pr->ev[r.slot][r.n][0], r.s));
    }
    r.kernels = kernels;
}

void craw_profiler_end(CRAWProfRun &r)
{
    if (!r.pr->on || r.n >= CRAW_MAX_STAGES)
        return;
    if (r.kernels) {
        /* a stage that launched nothing gets a zero-length interval */
        if (!tl_used) {
            CRAW_IGNORE(hipEventRecord(r.pr->ev[r.slot][r.n][0], r.s));
            CRAW_IGNORE(hipEventRecord(r.pr->ev[r.slot][r.n][1], r.s));
        }
        tl_start = tl_stop = NULL;
        tl_used = false;
    } else {
        CRAW_IGNORE(hipEventRecord(r.pr->ev[r.slot][r.n][1], r.s));
    }
    r.n++;
}

void craw_profiler_finish(CRAWProfRun &r)
{
    if (!r.pr->on)
        return;
    r.pr->nb[r.slot] = r.n;
    r.pr->recorded++;
}

int craw_profiler_times(CRAWContext *ctx, CRAWProfiler *pr, uint32_t frames_back,
                        CRAWStageTime *out, uint32_t max)
{
    uint32_t slot, nb, nb_agg = 0;

    if (max && !out)
        return CRAW_ERROR(EINVAL);
    if (!pr->on) {
        craw_log(ctx, CRAW_LOG_ERROR, "Created without CRAW_PIPELINE_FLAG_PROFILE\n");
        return CRAW_ERROR(EINVAL);
    }
    if (frames_back >= CRAW_PIPELINE_PROFILE_SLOTS || frames_back >= pr->recorded)
        return CRAW_ERROR(EINVAL);
    slot = (uint32_t)((pr->recorded - 1 - frames_back) % CRAW_PIPELINE_PROFILE_SLOTS);
    nb = pr->nb[slot];

    /* No waiting: work that has not completed is the caller's mistake. */
    for (uint32_t i = 0; i < nb; i++)
        if (hipEventQuery(pr->ev[slot][i][1]) != hipSuccess)
            return CRAW_ERROR(EAGAIN);

    for (uint32_t i = 0; i < nb; i++) {
        const char *name = pr->name[slot][i];
        float ms = 0.0f;
        uint32_t j;

        CRAW_IGNORE(hipEventElapsedTime(&ms, pr->ev[slot][i][0], pr->ev[slot][i][1]));
        for (j = 0; j < nb_agg; j++)
            if (!strcmp(pr->agg[j].name, name))
                break;
        if (j == nb_agg) {
            pr->agg[nb_agg].name = name;
            pr->agg[nb_agg].count = 0;
            pr->agg[nb_agg].ms = 0.0;
            nb_agg++;
        }
        pr->agg[j].count++;
        pr->agg[j].ms += ms;
    }
    if (max)
        memcpy(out, pr->agg, (max < nb_agg ? max : nb_agg) * sizeof(*out));
    return (int)nb_agg;
}
