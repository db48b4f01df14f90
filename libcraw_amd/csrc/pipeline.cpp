/*
 * libcraw for MI355X: the pipeline, lib/pipeline.c of cyanreg/libcraw for
 * this path.
 *
 *   craw_pipeline_create          pipeline.c:714-1252
 *   craw_pipeline_record          pipeline.c:1716-2316
 *   craw_pipeline_get_awb_histogram(_n)  pipeline.c:2318-2356
 *   craw_pipeline_get_stage_times(_n)    pipeline.c:2363-2415
 *
 * A recording plans the frame on the host (craw_frame_plan(), which fails
 * without side effects), then enqueues on the caller's stream: histogram
 * clear, denoise, debayer (fused with the colour stage and the AWB block
 * samples when there is no chroma denoise, except RCD with the gamut
 * compression, whose colour stage runs after it from its planes), the
 * chroma bilateral and the chroma/samples/colour pass, the histogram of
 * the block samples, and the histogram's copy into the next pinned
 * readback slot.
 */
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define DN_SEARCH_R 3
#define DN_PATCH_R  1
#define CD_MIN_SIZE 20

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
    CRAW_IGNORE(hipFree(p->buf.clean));
    CRAW_IGNORE(hipFree(p->buf.inter_rgb));
    CRAW_IGNORE(hipFree(p->buf.inter_rgb2));
    CRAW_IGNORE(hipFree(p->buf.hl_mosaic));
    CRAW_IGNORE(hipFree(p->buf.hl_levels));
    CRAW_IGNORE(hipFree(p->buf.down));
    CRAW_IGNORE(hipFree(p->buf.blur));
    CRAW_IGNORE(hipFree(p->buf.hist));
    CRAW_IGNORE(hipFree(p->buf.awb_samples));
    craw_temporal_release(&p->ts);
    for (int i = 0; i < 4; i++)
        if (p->hist_host[i])
            CRAW_IGNORE(hipHostFree(p->hist_host[i]));
    craw_profiler_free(&p->prof);
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
    int ret;

    if (!ctx || !params || !pipe)
        return CRAW_ERROR(EINVAL);
    if (params->cfa != CRAW_CFA_RGGB) {
        craw_log(ctx, CRAW_LOG_ERROR, "Only the RGGB CFA phase is implemented\n");
        return CRAW_ERROR(ENOTSUP);
    }
    if (!params->width || !params->height || (params->width & 1) || (params->height & 1) ||
        params->width < 8 || params->height < 8) {
        craw_log(ctx, CRAW_LOG_ERROR, "Invalid frame size %ux%u: both must be even and at "
                 "least 8\n", params->width, params->height);
        return CRAW_ERROR(EINVAL);
    }
    if (params->nb_awb_slots > 4)
        return CRAW_ERROR(EINVAL);

    p = (CRAWPipeline *)calloc(1, sizeof(*p));
    if (!p)
        return CRAW_ERROR(ENOMEM);
    p->ctx = ctx;
    p->params = *params;
    p->ts.prev = p->ts.cur = -1;
    ret = craw_stage_config(ctx, params->flags, params->debayer, params->out_space,
                            params->out_format, &p->cfg);
    if (ret < 0) {
        free(p);
        return ret;
    }
    W = params->width;
    H = params->height;

    if (hipSetDevice(ctx->device) != hipSuccess) {
        free(p);
        return CRAW_ERROR(ENODEV);
    }

    if (p->cfg.has_denoise || p->cfg.has_temporal) {
        /* One reflection of the plane mirror must reach the stencil:
         * lib/pipeline.c:809-825. */
        const uint32_t dn_min = 2 * (DN_SEARCH_R + DN_PATCH_R + 1 + (p->cfg.has_temporal ? 1 : 0));

        if (params->width < dn_min || params->height < dn_min) {
            craw_log(ctx, CRAW_LOG_ERROR, "Denoising needs a frame of at least %ux%u, got "
                     "%ux%u\n", dn_min, dn_min, params->width, params->height);
            ret = CRAW_ERROR(EINVAL);
            goto fail;
        }
        if ((ret = dmalloc((void **)&p->buf.clean, W * H * sizeof(uint16_t))) < 0)
            goto fail;
    }
    if (p->cfg.has_temporal) {
        if ((ret = craw_temporal_alloc(ctx, params->width, params->height, &p->ts)) < 0)
            goto fail;
        p->buf.ts = &p->ts;
    }
    if (p->cfg.has_chroma) {
        size_t wq = (W + 3) / 4, hq = (H + 3) / 4;

        if (params->width < CD_MIN_SIZE || params->height < CD_MIN_SIZE) {
            craw_log(ctx, CRAW_LOG_ERROR, "Chroma denoising needs a frame of at least %ux%u\n",
                     (unsigned)CD_MIN_SIZE, (unsigned)CD_MIN_SIZE);
            ret = CRAW_ERROR(EINVAL);
            goto fail;
        }
        if ((ret = dmalloc((void **)&p->buf.down, 3 * wq * hq * sizeof(float))) < 0 ||
            (ret = dmalloc((void **)&p->buf.blur, 2 * wq * hq * sizeof(float))) < 0)
            goto fail;
        if (p->cfg.has_fc &&
            (ret = dmalloc((void **)&p->buf.inter_rgb2, 3 * W * H * sizeof(float))) < 0)
            goto fail;
    }
    if (craw_needs_inter_rgb(p->cfg) &&
        (ret = dmalloc((void **)&p->buf.inter_rgb, 3 * W * H * sizeof(float))) < 0)
        goto fail;
    if (p->cfg.has_hl) {
        int lw[KHL_LEVELS], lh[KHL_LEVELS];
        size_t off[KHL_LEVELS], texels;

        if (craw_hl_levels(params->width, params->height, lw, lh, off, &texels) < 0) {
            craw_log(ctx, CRAW_LOG_ERROR, "Frame too large for the highlight field's pyramid\n");
            ret = CRAW_ERROR(EINVAL);
            goto fail;
        }
        if ((ret = dmalloc((void **)&p->buf.hl_mosaic, W * H * sizeof(uint16_t))) < 0 ||
            (ret = dmalloc((void **)&p->buf.hl_levels, texels * sizeof(float4))) < 0)
            goto fail;
    }
    if (p->cfg.awb) {
        p->nb_slots = params->nb_awb_slots ? params->nb_awb_slots : 1;
        if ((ret = dmalloc((void **)&p->buf.hist, CRAW_HIST_WORDS * sizeof(uint32_t))) < 0 ||
            (ret = dmalloc((void **)&p->buf.awb_samples,
                           craw_awb_blocks(params->width, params->height) * sizeof(uint2))) < 0)
            goto fail;
        for (uint32_t i = 0; i < p->nb_slots; i++) {
            /* coherent: the export kernel writes it directly */
            if (hipHostMalloc((void **)&p->hist_host[i], CRAW_HIST_WORDS * sizeof(uint32_t),
                              hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
                p->hist_host[i] = NULL;
                ret = CRAW_ERROR(ENOMEM);
                goto fail;
            }
            memset(p->hist_host[i], 0, CRAW_HIST_WORDS * sizeof(uint32_t));
            if (hipHostGetDevicePointer((void **)&p->hist_dev[i], p->hist_host[i], 0) !=
                hipSuccess) {
                ret = CRAW_ERROR(ENOMEM);
                goto fail;
            }
        }
        /* the device histogram is cleared by the first recording, on the
         * caller's stream (each recording's export leaves it zero again),
         * so that creating a pipeline synchronises no stream */
        p->hist_dirty = true;
    }
    if ((params->flags & CRAW_PIPELINE_FLAG_PROFILE) && (ret = craw_profiler_init(&p->prof)) < 0)
        goto fail;

    craw_log(ctx, CRAW_LOG_VERBOSE, "Pipeline %ux%u, debayer %d%s%s%s%s%s\n", params->width,
             params->height, (int)params->debayer, p->cfg.awb ? ", awb histogram" : "",
             p->cfg.has_denoise ? ", denoise" : "", p->cfg.has_temporal ? ", temporal" : "",
             p->cfg.has_chroma ? ", chroma denoise" : "",
             p->cfg.fp16 ? ", fp16 nlm" : (p->cfg.fast ? ", fast nlm" : ""));
    *pipe = p;
    return 0;

fail:
    craw_pipeline_free(&p);
    return ret;
}

/* Enqueues one planned frame (plain form); CRAW_ERROR(EIO) when a launch
 * fails. */
static int pipeline_enqueue(CRAWPipeline *p, const CRAWFramePlan *plan, hipStream_t stream)
{
    const KFrame &f = plan->f;
    CRAWProfRun pr = craw_profiler_start(&p->prof, stream);
    hipError_t e = hipSuccess;

    if (hipSetDevice(p->ctx->device) != hipSuccess)
        return CRAW_ERROR(ENODEV);

    /* after a recording that failed part way the histogram may hold
     * counts */
    if (p->cfg.awb && p->hist_dirty) {
        craw_profiler_begin(pr, "awb_clear");
        e = craw_launch_zero_u32(p->buf.hist, CRAW_HIST_WORDS, stream);
        craw_profiler_end(pr);
    }
    if (e == hipSuccess && plan->dn_mode) {
        if (p->cfg.fp16 || p->cfg.fast) {
            craw_profiler_begin(pr, p->cfg.fp16 ? "denoise_nlm_fp16" : "denoise_nlm_fast");
            e = craw_launch_denoise_fast(plan->dn_mode, p->cfg.fp16, f, stream);
        } else {
            craw_profiler_begin(pr, "denoise_nlm");
            e = craw_launch_denoise(plan->dn_mode, f, stream);
        }
        craw_profiler_end(pr);
    }
    if (e == hipSuccess && plan->hl) {
        craw_profiler_begin(pr, "highlight_field");
        e = craw_launch_hl_field(plan->hl_k, stream);
        craw_profiler_end(pr);
        if (e == hipSuccess) {
            craw_profiler_begin(pr, "highlight_push");
            e = craw_launch_hl_push(plan->hl_k, stream);
            craw_profiler_end(pr);
        }
        if (e == hipSuccess) {
            craw_profiler_begin(pr, "highlight_fill");
            e = craw_launch_hl_fill(plan->hl_k, stream);
            craw_profiler_end(pr);
        }
    }
    if (e == hipSuccess) {
        static const char *names[3] = { "debayer_bilinear", "debayer_mhc", "rcd" };

        craw_profiler_begin(pr, names[f.method]);
        e = craw_launch_debayer(plan->chroma || plan->fc_passes || plan->split_color, f, stream);
        craw_profiler_end(pr);
    }
    if (e == hipSuccess && plan->split_color) {
        craw_profiler_begin(pr, "color");
        e = craw_launch_chroma_up_finish(f, stream);
        craw_profiler_end(pr);
    }
    if (e == hipSuccess && plan->fc_passes) {
        craw_profiler_begin(pr, plan->chroma ? "false_color" : "false_color_color");
        e = craw_launch_false_color(plan->fc_passes, plan->chroma, f, plan->fc_to, stream);
        craw_profiler_end(pr);
    }
    if (e == hipSuccess && plan->chroma) {
        craw_profiler_begin(pr, "chroma_blur");
        e = craw_launch_chroma_blur(f, stream);
        craw_profiler_end(pr);
        if (e == hipSuccess) {
            craw_profiler_begin(pr, "chroma_up_color");
            e = craw_launch_chroma_up_finish(f, stream);
            craw_profiler_end(pr);
        }
    }
    if (e == hipSuccess && p->cfg.awb) {
        craw_profiler_begin(pr, "awb_hist");
        e = craw_launch_awb_hist(f, stream);
        craw_profiler_end(pr);
    }
    if (e == hipSuccess && p->cfg.awb) {
        craw_profiler_begin(pr, "awb_export");
        e = craw_launch_hist_export(p->buf.hist, p->hist_dev[p->nb_recorded % p->nb_slots],
                                    CRAW_HIST_WORDS, stream);
        craw_profiler_end(pr);
    }
    if (e != hipSuccess) {
        craw_log(p->ctx, CRAW_LOG_ERROR, "Launch failed: %s\n", hipGetErrorString(e));
        p->hist_dirty = true;
        return CRAW_ERROR(EIO);
    }
    p->hist_dirty = false;
    craw_profiler_finish(pr);
    return 0;
}

CRAW_API int craw_pipeline_record(CRAWPipeline *p, CRAWStream stream, const CRAWImage *src,
                                  const CRAWImage *dst, const CRAWFrameParams *fp)
{
    CRAWFramePlan plan;
    int ret;

    if (!p || !src || !dst || !fp || !fp->profile)
        return CRAW_ERROR(EINVAL);
    ret = craw_frame_plan(p->ctx, p->cfg, p->buf, p->params.width, p->params.height, src, dst,
                          fp, &plan);
    if (ret < 0)
        return ret;
    ret = pipeline_enqueue(p, &plan, (hipStream_t)stream);
    if (ret < 0)
        return ret;
    if (p->cfg.awb)
        p->nb_recorded++;
    craw_frame_commit(&plan);
    return 0;
}

CRAW_API int craw_pipeline_get_awb_histogram_n(CRAWPipeline *p, uint32_t frames_back,
                                               CRAWAwbHistogram *out)
{
    const uint32_t *base;

    if (!p || !out)
        return CRAW_ERROR(EINVAL);
    if (!p->cfg.awb) {
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
    if (!p)
        return CRAW_ERROR(EINVAL);
    return craw_profiler_times(p->ctx, &p->prof, frames_back, out, max);
}

CRAW_API int craw_pipeline_get_stage_times(CRAWPipeline *p, CRAWStageTime *out, uint32_t max)
{
    return craw_pipeline_get_stage_times_n(p, 0, out, max);
}

} /* extern "C" */
