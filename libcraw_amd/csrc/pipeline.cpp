/*
 * libcraw for MI355X: the pipeline, lib/pipeline.c of cyanreg/libcraw for
 * this path.
 *
 *   craw_pipeline_create          pipeline.c:714-1252
 *   craw_pipeline_record          pipeline.c:1716-2316
 *   craw_pipeline_get_awb_histogram(_n)  pipeline.c:2318-2356
 *   craw_pipeline_get_stage_times(_n)    pipeline.c:2363-2415
 *
 * A recording validates and evaluates everything on the host first (the
 * reference's "host phase", which fails without side effects), then
 * enqueues on the caller's stream: histogram clear, denoise, debayer
 * (fused with the colour stage and the histogram when there is no chroma
 * denoise), chroma bilateral and the chroma/histogram/colour pass, and the
 * histogram's copy into the next pinned readback slot.
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
    CRAW_PIPELINE_FLAG_LENS | CRAW_PIPELINE_FLAG_FALSE_COLOR |
    CRAW_PIPELINE_FLAG_HIGHLIGHT_RECONSTRUCTION | CRAW_PIPELINE_FLAG_DEFRINGE |
    CRAW_PIPELINE_FLAG_SHARPEN;

static const uint32_t known_flags =
    CRAW_PIPELINE_FLAG_AWB_HISTOGRAM | CRAW_PIPELINE_FLAG_DENOISE |
    CRAW_PIPELINE_FLAG_CHROMA_DENOISE | CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE |
    CRAW_PIPELINE_FLAG_PROFILE | CRAW_PIPELINE_FLAG_NLM_FP16 | unsupported_flags;

extern "C" {

CRAW_API void craw_pipeline_params_init(CRAWPipelineParams *p)
{
    if (!p)
        return;
    memset(p, 0, sizeof(*p));
    p->cfa = CRAW_CFA_RGGB;
    p->debayer = CRAW_DEBAYER_RCD;
    p->out_space = CRAW_CS_XYZ_D65;
    p->out_format = CRAW_FORMAT_R16G16B16A16_SFLOAT;
}

CRAW_API void craw_frame_params_init(CRAWFrameParams *fp)
{
    if (!fp)
        return;
    memset(fp, 0, sizeof(*fp));
    fp->wb_mode = CRAW_WB_AS_SHOT;
}

CRAW_API void craw_pipeline_free(CRAWPipeline **pipe)
{
    CRAWPipeline *p;

    if (!pipe || !*pipe)
        return;
    p = *pipe;
    CRAW_IGNORE(hipSetDevice(p->ctx->device));
    CRAW_IGNORE(hipFree(p->clean));
    for (int i = 0; i < 2; i++) {
        CRAW_IGNORE(hipFree(p->tn_level[i]));
        CRAW_IGNORE(hipFree(p->tn_var[i]));
    }
    CRAW_IGNORE(hipFree(p->inter_rgb));
    CRAW_IGNORE(hipFree(p->down));
    CRAW_IGNORE(hipFree(p->blur));
    CRAW_IGNORE(hipFree(p->hist));
    for (int i = 0; i < 4; i++)
        if (p->hist_host[i])
            CRAW_IGNORE(hipHostFree(p->hist_host[i]));
    if (p->profile)
        for (int s = 0; s < CRAW_PIPELINE_PROFILE_SLOTS; s++)
            for (int i = 0; i < CRAW_MAX_STAGES; i++)
                for (int j = 0; j < 2; j++)
                    if (p->ev[s][i][j])
                        CRAW_IGNORE(hipEventDestroy(p->ev[s][i][j]));
    free(p);
    *pipe = NULL;
}

static int dmalloc(void **ptr, size_t bytes)
{
    return hipMalloc(ptr, bytes) == hipSuccess ? 0 : CRAW_ERROR(ENOMEM);
}

CRAW_API int craw_pipeline_create(CRAWContext *ctx, const CRAWPipelineParams *params,
                                  CRAWPipeline **pipe)
{
    CRAWPipeline *p;
    size_t W, H;
    int ret = 0;

    if (!ctx || !params || !pipe)
        return CRAW_ERROR(EINVAL);
    if (params->cfa != CRAW_CFA_RGGB) {
        craw_log(ctx, CRAW_LOG_ERROR, "Only the RGGB CFA phase is implemented\n");
        return CRAW_ERROR(ENOTSUP);
    }
    if (params->flags & unsupported_flags) {
        craw_log(ctx, CRAW_LOG_ERROR, "Pipeline flags 0x%x name stages this library does not "
                 "have\n", params->flags & unsupported_flags);
        return CRAW_ERROR(ENOTSUP);
    }
    if (params->flags & ~known_flags)
        return CRAW_ERROR(EINVAL);
    if (!params->width || !params->height || (params->width & 1) || (params->height & 1) ||
        params->width < 8 || params->height < 8) {
        craw_log(ctx, CRAW_LOG_ERROR, "Invalid frame size %ux%u: both must be even and at "
                 "least 8\n", params->width, params->height);
        return CRAW_ERROR(EINVAL);
    }
    if (params->out_format != CRAW_FORMAT_R16G16B16A16_SFLOAT &&
        params->out_format != CRAW_FORMAT_R32G32B32A32_SFLOAT) {
        craw_log(ctx, CRAW_LOG_ERROR, "Output format must be RGBA16F or RGBA32F\n");
        return CRAW_ERROR(EINVAL);
    }
    if ((unsigned)params->debayer >= CRAW_DEBAYER_NB || (unsigned)params->out_space >= CRAW_CS_NB)
        return CRAW_ERROR(EINVAL);
    if (params->nb_awb_slots > 4)
        return CRAW_ERROR(EINVAL);

    p = (CRAWPipeline *)calloc(1, sizeof(*p));
    if (!p)
        return CRAW_ERROR(ENOMEM);
    p->ctx = ctx;
    p->params = *params;
    p->awb = !!(params->flags & CRAW_PIPELINE_FLAG_AWB_HISTOGRAM);
    p->has_denoise = !!(params->flags & CRAW_PIPELINE_FLAG_DENOISE);
    p->has_temporal = !!(params->flags & CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE);
    p->has_chroma = !!(params->flags & CRAW_PIPELINE_FLAG_CHROMA_DENOISE);
    p->profile = !!(params->flags & CRAW_PIPELINE_FLAG_PROFILE);
    p->fp16 = !!(params->flags & CRAW_PIPELINE_FLAG_NLM_FP16);
    p->tn_prev = p->tn_cur = -1;
    W = params->width;
    H = params->height;

    if (hipSetDevice(ctx->device) != hipSuccess) {
        free(p);
        return CRAW_ERROR(ENODEV);
    }

    if (p->has_denoise || p->has_temporal) {
        /* One reflection of the plane mirror must reach the stencil:
         * lib/pipeline.c:809-825. */
        const uint32_t dn_min = 2 * (DN_SEARCH_R + DN_PATCH_R + 1 + (p->has_temporal ? 1 : 0));

        if (params->width < dn_min || params->height < dn_min) {
            craw_log(ctx, CRAW_LOG_ERROR, "Denoising needs a frame of at least %ux%u, got "
                     "%ux%u\n", dn_min, dn_min, params->width, params->height);
            ret = CRAW_ERROR(EINVAL);
            goto fail;
        }
        if ((ret = dmalloc((void **)&p->clean, W * H * sizeof(uint16_t))) < 0)
            goto fail;
    }
    if (p->has_temporal) {
        for (int i = 0; i < 2; i++) {
            if ((ret = dmalloc((void **)&p->tn_level[i], W * H * sizeof(float))) < 0 ||
                (ret = dmalloc((void **)&p->tn_var[i], W * H)) < 0)
                goto fail;
        }
    }
    if (p->has_chroma) {
        if (params->width < CD_MIN_SIZE || params->height < CD_MIN_SIZE) {
            craw_log(ctx, CRAW_LOG_ERROR, "Chroma denoising needs a frame of at least %ux%u\n",
                     (unsigned)CD_MIN_SIZE, (unsigned)CD_MIN_SIZE);
            ret = CRAW_ERROR(EINVAL);
            goto fail;
        }
        p->wq = (int)((W + 3) / 4);
        p->hq = (int)((H + 3) / 4);
        if ((ret = dmalloc((void **)&p->inter_rgb, 3 * W * H * sizeof(float))) < 0 ||
            (ret = dmalloc((void **)&p->down, 3 * (size_t)p->wq * p->hq * sizeof(float))) < 0 ||
            (ret = dmalloc((void **)&p->blur, 2 * (size_t)p->wq * p->hq * sizeof(float))) < 0)
            goto fail;
    }
    if (p->awb) {
        p->nb_slots = params->nb_awb_slots ? params->nb_awb_slots : 1;
        if ((ret = dmalloc((void **)&p->hist, CRAW_HIST_WORDS * sizeof(uint32_t))) < 0)
            goto fail;
        for (uint32_t i = 0; i < p->nb_slots; i++) {
            if (hipHostMalloc((void **)&p->hist_host[i], CRAW_HIST_WORDS * sizeof(uint32_t),
                              hipHostMallocDefault) != hipSuccess) {
                p->hist_host[i] = NULL;
                ret = CRAW_ERROR(ENOMEM);
                goto fail;
            }
            memset(p->hist_host[i], 0, CRAW_HIST_WORDS * sizeof(uint32_t));
        }
    }
    if (p->profile) {
        for (int s = 0; s < CRAW_PIPELINE_PROFILE_SLOTS; s++)
            for (int i = 0; i < CRAW_MAX_STAGES; i++)
                for (int j = 0; j < 2; j++)
                    if (hipEventCreate(&p->ev[s][i][j]) != hipSuccess) {
                        p->ev[s][i][j] = NULL;
                        ret = CRAW_ERROR(ENOMEM);
                        goto fail;
                    }
    }

    craw_log(ctx, CRAW_LOG_VERBOSE, "Pipeline %ux%u, debayer %d%s%s%s%s%s\n", params->width,
             params->height, (int)params->debayer, p->awb ? ", awb histogram" : "",
             p->has_denoise ? ", denoise" : "", p->has_temporal ? ", temporal" : "",
             p->has_chroma ? ", chroma denoise" : "", p->fp16 ? ", fp16 nlm" : "");
    *pipe = p;
    return 0;

fail:
    craw_pipeline_free(&p);
    return ret;
}

/* ------------------------------------------------------------ planning */

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

int craw_pipeline_plan(CRAWPipeline *p, const CRAWImage *src, const CRAWImage *dst,
                       const CRAWFrameParams *fp, CRAWFramePlan *plan)
{
    CRAWContext *ctx;
    CRAWColorParams frame;
    CRAWGamutParams gp;
    uint32_t w, h;
    enum CRAWColorSpace out_space;
    float m_out[3][3], scale;
    bool spatial, temporal, replay;
    int t_read, t_write, ret;

    if (!p || !src || !dst || !fp || !fp->profile || !plan)
        return CRAW_ERROR(EINVAL);
    ctx = p->ctx;
    w = p->params.width;
    h = p->params.height;
    memset(plan, 0, sizeof(*plan));

    if (src->format != CRAW_FORMAT_R16_UNORM) {
        craw_log(ctx, CRAW_LOG_ERROR, "Source must be CRAW_FORMAT_R16_UNORM (got %d)\n",
                 (int)src->format);
        return CRAW_ERROR(EINVAL);
    }
    if (dst->format != p->params.out_format) {
        craw_log(ctx, CRAW_LOG_ERROR, "Destination format %d does not match the pipeline's "
                 "%d\n", (int)dst->format, (int)p->params.out_format);
        return CRAW_ERROR(EINVAL);
    }
    if (src->width != w || src->height != h || dst->width != w || dst->height != h) {
        craw_log(ctx, CRAW_LOG_ERROR, "Frame size mismatch: src %ux%u dst %ux%u, pipeline "
                 "%ux%u\n", src->width, src->height, dst->width, dst->height, w, h);
        return CRAW_ERROR(EINVAL);
    }
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

    /* The frame's gains go into the prep, so that the clip keeps blown
     * highlights neutral under the chosen illuminant. */
    ret = craw_profile_eval(fp->profile, fp->wb_mode, &fp->illuminant, &frame);
    if (ret < 0)
        return ret;
    for (int s = 0; s < 4; s++) {
        float range = fp->profile->white_level[s] - fp->profile->black_level[s];
        static const int chan[4] = { 0, 1, 1, 2 };

        if (!(range > 0.0f)) {
            craw_log(ctx, CRAW_LOG_ERROR, "Profile site %d has white level %f <= black level "
                     "%f\n", s, fp->profile->white_level[s], fp->profile->black_level[s]);
            return CRAW_ERROR(EINVAL);
        }
        plan->prep.black[s] = fp->profile->black_level[s];
        plan->prep.range[s] = range;
        plan->prep.wb[s] = frame.wb[chan[s]];
    }

    /* Denoise: which filters, which state sets (lib/pipeline.c:1784-1818). */
    spatial = p->has_denoise && fp->denoise_strength > 0.0f;
    temporal = p->has_temporal && fp->denoise_temporal_strength > 0.0f;
    if (temporal && !(fp->denoise_temporal_strength <= 1.0f)) {
        craw_log(ctx, CRAW_LOG_ERROR, "denoise_temporal_strength = %g; at most 1.0\n",
                 (double)fp->denoise_temporal_strength);
        return CRAW_ERROR(EINVAL);
    }
    replay = fp->denoise_temporal_replay;
    t_read = replay ? p->tn_prev : p->tn_cur;
    if (fp->flags & CRAW_FRAME_FLAG_TEMPORAL_RESET)
        t_read = -1;
    t_write = -1;
    if (temporal)
        t_write = replay && p->tn_cur >= 0 ? p->tn_cur : (t_read == 0 ? 1 : 0);
    plan->t_read = t_read;
    plan->t_write = t_write;
    plan->temporal = temporal;
    plan->replay = replay;

    if (spatial || temporal) {
        CRAWNoiseModel nm;
        float a, b, c, sc, hh;

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
        plan->dn.w2 = (int)(w / 2);
        plan->dn.h2 = (int)(h / 2);
        plan->dn.gat_a = a;
        plan->dn.gat_c = c;
        plan->dn.gat_sc = sc;
        plan->dn.gat_fwd = 2.0f / a;
        plan->dn.gat_inv = a / 4.0f;
        plan->dn.gat_off = c / a;
        plan->dn.gat_t0 = plan->dn.gat_fwd * sc;
        plan->dn.inv_h2 = 1.0f / (hh * hh);
        plan->dn.guard_lo = DN_GUARD_LO;
        plan->dn.guard_inv = 1.0f / (DN_GUARD_HI - DN_GUARD_LO);
        plan->dn.alpha_max = temporal ? fp->denoise_temporal_strength : 0.0f;
        for (int s = 0; s < 4; s++) {
            float range = fp->profile->white_level[s] - fp->profile->black_level[s];
            plan->dn.black[s] = fp->profile->black_level[s];
            plan->dn.range[s] = range;
            plan->dn.inv_range[s] = 1.0f / range;
        }
        plan->dn.src_pitch = (int)(src->pitch / 2);
        plan->dn.dst_pitch = (int)w;
        plan->dn_mode = (spatial ? KDN_SPATIAL : 0) | (temporal ? KDN_TEMPORAL : 0) |
                        (temporal && t_read >= 0 ? KDN_HISTORY : 0);
        if (temporal) {
            if (t_read >= 0) {
                plan->st.prev_level = p->tn_level[t_read];
                plan->st.prev_var = p->tn_var[t_read];
            }
            plan->st.next_level = p->tn_level[t_write];
            plan->st.next_var = p->tn_var[t_write];
        }
    }

    plan->chroma = p->has_chroma && fp->denoise_chroma_strength > 0.0f;
    if (plan->chroma && !(fp->denoise_chroma_strength <= 1.0f)) {
        craw_log(ctx, CRAW_LOG_ERROR, "denoise_chroma_strength = %g; at most 1.0\n",
                 (double)fp->denoise_chroma_strength);
        return CRAW_ERROR(EINVAL);
    }
    plan->chroma_strength = fp->denoise_chroma_strength;
    plan->inter.rgb = p->inter_rgb;
    plan->inter.down = p->down;
    plan->inter.blur = p->blur;
    plan->inter.wq = p->wq;
    plan->inter.hq = p->hq;

    /* Colour: the frame's matrix and scale, then XYZ to the output space,
     * as lib/ref/ref_color.c:10-76 applies them; gamut compression for
     * ACEScg unless vetoed. */
    out_space = fp->out_space_set ? fp->out_space : p->params.out_space;
    ret = craw_colorspace_from_xyz_d65(out_space, m_out);
    if (ret < 0) {
        craw_log(ctx, CRAW_LOG_ERROR, "Unknown output colour space %d\n", (int)out_space);
        return ret;
    }
    scale = frame.gain * exp2f(fp->exposure_stops);
    memcpy(plan->col.cam, frame.cam_to_xyz, sizeof(plan->col.cam));
    memcpy(plan->col.out, m_out, sizeof(plan->col.out));
    plan->col.scale = scale;
    craw_gamut_params(&gp);
    memcpy(plan->col.thr, gp.thr, sizeof(gp.thr));
    memcpy(plan->col.scl, gp.scl, sizeof(gp.scl));
    plan->col.pwr = gp.pwr;
    plan->col.inv_pwr = gp.inv_pwr;
    plan->col.flags = (fp->flags & CRAW_FRAME_FLAG_CLAMP_NEGATIVE) ? KCOLOR_CLAMP_NEGATIVE : 0u;
    if (out_space == CRAW_CS_ACES_AP1_LINEAR && !(fp->flags & CRAW_FRAME_FLAG_NO_GAMUT_COMPRESSION))
        plan->col.flags |= KCOLOR_GAMUT_COMPRESS;

    plan->out.dst = dst->data;
    plan->out.pitch = dst->pitch;
    plan->out.half = dst->format == CRAW_FORMAT_R16G16B16A16_SFLOAT;
    plan->out.hist = p->awb ? p->hist : NULL;
    plan->out.awb_wb0 = frame.wb[0];
    plan->out.awb_wb2 = frame.wb[2];

    plan->src = (const uint16_t *)src->data;
    plan->src_pitch = (int)(src->pitch / 2);
    plan->method = (int)p->params.debayer;
    return 0;
}

/* -------------------------------------------------------------- enqueue */

struct Prof {
    CRAWPipeline *p;
    hipStream_t s;
    uint32_t slot;
    uint32_t n;
};

static void prof_begin(Prof &pr, const char *name)
{
    if (!pr.p->profile || pr.n >= CRAW_MAX_STAGES)
        return;
    pr.p->ev_name[pr.slot][pr.n] = name;
    CRAW_IGNORE(hipEventRecord(pr.p->ev[pr.slot][pr.n][0], pr.s));
}

static void prof_end(Prof &pr)
{
    if (!pr.p->profile || pr.n >= CRAW_MAX_STAGES)
        return;
    CRAW_IGNORE(hipEventRecord(pr.p->ev[pr.slot][pr.n][1], pr.s));
    pr.n++;
}

int craw_pipeline_enqueue(CRAWPipeline *p, const CRAWFramePlan *plan, hipStream_t stream)
{
    const int W = (int)p->params.width, H = (int)p->params.height;
    const uint16_t *mosaic = plan->src;
    int mosaic_pitch = plan->src_pitch;
    Prof pr = { p, stream, (uint32_t)(p->prof_recorded % CRAW_PIPELINE_PROFILE_SLOTS), 0 };
    hipError_t e = hipSuccess;

    if (hipSetDevice(p->ctx->device) != hipSuccess)
        return CRAW_ERROR(ENODEV);

    if (p->awb) {
        prof_begin(pr, "awb_clear");
        e = craw_launch_zero_u32(p->hist, CRAW_HIST_WORDS, stream);
        prof_end(pr);
    }
    if (e == hipSuccess && plan->dn_mode) {
        KDenoise dn = plan->dn;

        dn.src_pitch = mosaic_pitch;
        prof_begin(pr, "denoise_nlm");
        e = craw_launch_denoise(plan->dn_mode, p->fp16, dn, mosaic, p->clean, plan->st, stream);
        prof_end(pr);
        mosaic = p->clean;
        mosaic_pitch = W;
    }
    if (e == hipSuccess) {
        static const char *names[3] = { "debayer_bilinear", "debayer_mhc", "rcd" };

        prof_begin(pr, names[plan->method]);
        e = craw_launch_debayer(plan->method, plan->chroma, mosaic, mosaic_pitch, W, H,
                                plan->prep, plan->col, plan->out, plan->inter, stream);
        prof_end(pr);
    }
    if (e == hipSuccess && plan->chroma) {
        prof_begin(pr, "chroma_blur");
        e = craw_launch_chroma_blur(plan->inter, stream);
        prof_end(pr);
        if (e == hipSuccess) {
            prof_begin(pr, "chroma_up_color");
            e = craw_launch_chroma_up_finish(plan->inter, plan->chroma_strength, W, H,
                                             plan->col, plan->out, stream);
            prof_end(pr);
        }
    }
    if (e == hipSuccess && p->awb) {
        uint32_t *slot = p->hist_host[p->nb_recorded % p->nb_slots];
        e = hipMemcpyAsync(slot, p->hist, CRAW_HIST_WORDS * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, stream);
    }
    if (e != hipSuccess) {
        craw_log(p->ctx, CRAW_LOG_ERROR, "Launch failed: %s\n", hipGetErrorString(e));
        return CRAW_ERROR(EIO);
    }
    if (p->profile)
        p->ev_nb[pr.slot] = pr.n;
    return 0;
}

void craw_pipeline_commit(CRAWPipeline *p, const CRAWFramePlan *plan)
{
    if (p->awb)
        p->nb_recorded++;
    if (p->profile)
        p->prof_recorded++;
    /* lib/pipeline.c:2303-2313: only a recording that succeeded moves the
     * state. */
    if (p->has_temporal) {
        p->tn_prev = plan->t_read;
        if (plan->temporal)
            p->tn_cur = plan->t_write;
        else if (!plan->replay)
            p->tn_cur = -1;
    }
}

CRAW_API int craw_pipeline_record(CRAWPipeline *p, CRAWStream stream, const CRAWImage *src,
                                  const CRAWImage *dst, const CRAWFrameParams *fp)
{
    CRAWFramePlan plan;
    int ret;

    ret = craw_pipeline_plan(p, src, dst, fp, &plan);
    if (ret < 0)
        return ret;
    ret = craw_pipeline_enqueue(p, &plan, (hipStream_t)stream);
    if (ret < 0)
        return ret;
    craw_pipeline_commit(p, &plan);
    return 0;
}

CRAW_API int craw_pipeline_get_awb_histogram_n(CRAWPipeline *p, uint32_t frames_back,
                                               CRAWAwbHistogram *out)
{
    const uint32_t *base;

    if (!p || !out)
        return CRAW_ERROR(EINVAL);
    if (!p->awb) {
        craw_log(p->ctx, CRAW_LOG_ERROR,
                 "Pipeline was created without CRAW_PIPELINE_FLAG_AWB_HISTOGRAM\n");
        return CRAW_ERROR(EINVAL);
    }
    if (frames_back >= p->nb_slots || frames_back >= p->nb_recorded)
        return CRAW_ERROR(EINVAL);
    base = p->hist_host[(p->nb_recorded - 1 - frames_back) % p->nb_slots];
    out->bins = CRAW_AWB_HIST_BINS;
    out->log_min = CRAW_AWB_HIST_LOG_MIN;
    out->log_max = CRAW_AWB_HIST_LOG_MAX;
    out->count = base;
    out->weight = base + CRAW_AWB_HIST_BINS * CRAW_AWB_HIST_BINS;
    out->weight_bright = base + 2 * CRAW_AWB_HIST_BINS * CRAW_AWB_HIST_BINS;
    return 0;
}

CRAW_API int craw_pipeline_get_awb_histogram(CRAWPipeline *p, CRAWAwbHistogram *out)
{
    return craw_pipeline_get_awb_histogram_n(p, 0, out);
}

CRAW_API int craw_pipeline_get_stage_times_n(CRAWPipeline *p, uint32_t frames_back,
                                             CRAWStageTime *out, uint32_t max)
{
    uint32_t slot, nb, nb_agg = 0;

    if (!p || (max && !out))
        return CRAW_ERROR(EINVAL);
    if (!p->profile) {
        craw_log(p->ctx, CRAW_LOG_ERROR, "Pipeline was created without "
                 "CRAW_PIPELINE_FLAG_PROFILE\n");
        return CRAW_ERROR(EINVAL);
    }
    if (frames_back >= CRAW_PIPELINE_PROFILE_SLOTS || frames_back >= p->prof_recorded)
        return CRAW_ERROR(EINVAL);
    slot = (uint32_t)((p->prof_recorded - 1 - frames_back) % CRAW_PIPELINE_PROFILE_SLOTS);
    nb = p->ev_nb[slot];

    /* No waiting: work that has not completed is the caller's mistake. */
    for (uint32_t i = 0; i < nb; i++)
        if (hipEventQuery(p->ev[slot][i][1]) != hipSuccess)
            return CRAW_ERROR(EAGAIN);

    for (uint32_t i = 0; i < nb; i++) {
        const char *name = p->ev_name[slot][i];
        float ms = 0.0f;
        uint32_t j;

        CRAW_IGNORE(hipEventElapsedTime(&ms, p->ev[slot][i][0], p->ev[slot][i][1]));
        for (j = 0; j < nb_agg; j++)
            if (!strcmp(p->agg[j].name, name))
                break;
        if (j == nb_agg) {
            p->agg[nb_agg].name = name;
            p->agg[nb_agg].count = 0;
            p->agg[nb_agg].ms = 0.0;
            nb_agg++;
        }
        p->agg[j].count++;
        p->agg[j].ms += ms;
    }
    if (max)
        memcpy(out, p->agg, (max < nb_agg ? max : nb_agg) * sizeof(*out));
    return (int)nb_agg;
}

CRAW_API int craw_pipeline_get_stage_times(CRAWPipeline *p, CRAWStageTime *out, uint32_t max)
{
    return craw_pipeline_get_stage_times_n(p, 0, out, max);
}

} /* extern "C" */
