/*
 * libcraw for MI355X: internal state of contexts, pipelines, batches and
 * temporal states.
 */
#ifndef CRAW_INTERNAL_H
#define CRAW_INTERNAL_H

#include <pthread.h>

#include <hip/hip_runtime.h>

#include <craw/pipeline.h>
#include <craw/batch.h>
#include "hostmath.h"
#include "kernels.h"

struct CRAWContext {
    int device;
    CRAWContextOptions opts;
    char arch[64];
    int cu_count;

    /* craw_noise_estimate_image()'s device scratch, grown to the largest
     * frame measured and kept until the context is freed; the lock makes
     * the measurements on one context take turns. */
    pthread_mutex_t noise_lock;
    void *noise_scratch;
    size_t noise_scratch_size;
};

#define CRAW_MAX_STAGES 16
#define CRAW_HIST_WORDS (3u * CRAW_AWB_HIST_BINS * CRAW_AWB_HIST_BINS)

/* Two sets of (level, variance) for one clip stream and which set holds
 * what (lib/pipeline.c:435-445, 1796-1818, 2303-2313 of the reference):
 * `cur` the set the last recording wrote, `prev` the set it read, -1
 * nothing valid. */
struct CRAWTemporalState {
    CRAWContext *ctx;
    uint32_t width, height;
    float *level[2];            /* 4 planes of (w/2) x (h/2) */
    uint8_t *var[2];            /* square root of the variance, 8 bits */
    int prev, cur;
};

/* What a pipeline or a batch runs; fixed at creation. */
struct CRAWStageConfig {
    enum CRAWDebayer debayer;
    enum CRAWColorSpace out_space;
    enum CRAWFormat out_format;
    bool awb, has_denoise, has_temporal, has_chroma, has_fc, has_hl, fp16, fast;
};

/* Whether a pipeline or batch of this configuration holds the camera RGB
 * planes: for the chroma or false-colour stage, and for the RCD debayer
 * to ACES AP1, whose colour stage with the gamut compression then runs
 * from the planes (streamed RCD holds too many registers to run the
 * compression's double-precision pows at speed). */
static inline bool craw_needs_inter_rgb(const CRAWStageConfig &cfg)
{
    return cfg.has_chroma || cfg.has_fc ||
           (cfg.debayer == CRAW_DEBAYER_RCD && cfg.out_space == CRAW_CS_ACES_AP1_LINEAR);
}

/* The device buffers one frame's recording uses. */
struct CRAWFrameBuffers {
    uint16_t *clean;            /* W x H, when denoising */
    float *inter_rgb;           /* 3 x W x H, with chroma denoise or false colour, or an
                                 * RCD debayer to ACES AP1 (craw_split_color_buffers) */
    float *inter_rgb2;          /* the debayer's planes for false colour before chroma */
    uint16_t *hl_mosaic;        /* W x H R16F, with highlight reconstruction */
    float4 *hl_levels;          /* the colour field's levels, one after another */
    float *down, *blur;         /* 3 and 2 quarter planes */
    uint32_t *hist;             /* CRAW_HIST_WORDS, with the histogram */
    uint2 *awb_samples;         /* one per 8x8 block, with the histogram */
    CRAWTemporalState *ts;      /* with the temporal pass, or NULL */
};

/* Everything one frame's recording enqueues, worked out on the host
 * before any of it is launched, so that a failing record changes nothing. */
struct CRAWFramePlan {
    KFrame f;                   /* what the kernels read */
    int dn_mode;                /* KDN_* bits, 0: no denoise */
    bool chroma;
    /* the RCD debayer writes its planes and the colour stage runs after it
     * from them (the chroma-up kernel with f.inter.up = 0): with the gamut
     * compression and no chroma or false-colour stage */
    bool split_color;
    int fc_passes;              /* false colour passes, 0: none */
    bool hl;                    /* highlight reconstruction before the debayer */
    KHighlight hl_k;
    KInter fc_to;               /* the false colour stage's output: the chroma stage's inputs */
    int t_read, t_write;        /* temporal state sets, -1: none */
    bool temporal, replay;
    CRAWTemporalState *ts;
};

struct CRAWProfiler {
    hipEvent_t ev[CRAW_PIPELINE_PROFILE_SLOTS][CRAW_MAX_STAGES][2];
    const char *name[CRAW_PIPELINE_PROFILE_SLOTS][CRAW_MAX_STAGES];
    uint32_t nb[CRAW_PIPELINE_PROFILE_SLOTS];
    uint64_t recorded;
    CRAWStageTime agg[CRAW_MAX_STAGES];
    bool on;
};

/* One recording's launches being profiled. */
struct CRAWProfRun {
    CRAWProfiler *pr;
    hipStream_t s;
    uint32_t slot, n;
    bool kernels;               /* the open stage's events ride on its launches */
};

struct CRAWPipeline {
    CRAWContext *ctx;
    CRAWPipelineParams params;
    CRAWStageConfig cfg;
    CRAWFrameBuffers buf;
    CRAWTemporalState ts;
    uint32_t *hist_host[4];     /* pinned readback slots (coherent) */
    uint32_t *hist_dev[4];      /* the same, as the device addresses them */
    bool hist_dirty;            /* a failed recording may have left counts */
    uint32_t nb_slots;
    uint64_t nb_recorded;
    CRAWProfiler prof;
};

extern "C" {
void craw_vlog_ctx(CRAWContext *ctx, enum CRAWLogLevel level, const char *fmt, va_list ap);
}

/* Checks and resolves a stage configuration's flags. */
int craw_stage_config(CRAWContext *ctx, uint32_t flags, enum CRAWDebayer debayer,
                      enum CRAWColorSpace out_space, enum CRAWFormat out_format,
                      CRAWStageConfig *cfg);

/* The highlight field's pyramid for a W x H frame: level sizes and
 * offsets (in texels) into one array; the level count, or
 * CRAW_ERROR(EINVAL) past KHL_LEVELS. */
int craw_hl_levels(uint32_t w, uint32_t h, int lw[KHL_LEVELS], int lh[KHL_LEVELS],
                   size_t off[KHL_LEVELS], size_t *total);

/* Plans one frame of W x H (validation, constants, state selection); 0 or
 * a negative errno, nothing changed on failure. */

int craw_frame_plan(CRAWContext *ctx, const CRAWStageConfig &cfg, const CRAWFrameBuffers &buf,
                    uint32_t W, uint32_t H, const CRAWImage *src, const CRAWImage *dst,
                    const CRAWFrameParams *fp, CRAWFramePlan *plan);

/* Moves a planned frame's temporal indices once it has been enqueued. */
void craw_frame_commit(const CRAWFramePlan *plan);

/* A temporal state's device memory. */
int craw_temporal_alloc(CRAWContext *ctx, uint32_t w, uint32_t h, CRAWTemporalState *ts);
void craw_temporal_release(CRAWTemporalState *ts);

/* Profiling: an event pair around every launch of a recording, the last
 * CRAW_PIPELINE_PROFILE_SLOTS recordings kept. */
int craw_profiler_init(CRAWProfiler *pr);
void craw_profiler_free(CRAWProfiler *pr);
CRAWProfRun craw_profiler_start(CRAWProfiler *pr, hipStream_t s);
/* Opens a stage. kernels: the stage is kernel launches (KLAUNCH), timed from
 * the first one's start to the last one's end; otherwise (a graph launch,
 * a copy) event markers are recorded around it on the stream. */
void craw_profiler_begin(CRAWProfRun &r, const char *name, bool kernels = true);
void craw_profiler_end(CRAWProfRun &r);
void craw_profiler_finish(CRAWProfRun &r);
int craw_profiler_times(CRAWContext *ctx, CRAWProfiler *pr, uint32_t frames_back,
                        CRAWStageTime *out, uint32_t max);

/* Looks up, once per device, the address of the allocation-free entry
 * point's scratch array on ctx->device (which must be current). */
int craw_noalloc_setup(CRAWContext *ctx);

/* Blocks of the AWB grid of a W x H frame. */
static inline size_t craw_awb_blocks(uint32_t w, uint32_t h)
{
    return (size_t)((w + 7) / 8) * ((h + 7) / 8);
}

/* Device bytes a batch of `params` allocates (no allocation). */
int craw_batch_footprint(CRAWContext *ctx, const CRAWBatchParams *params, uint64_t *bytes);

/* Plans `frames` as craw_batch_record() would, and launches nothing: 0 or
 * the error recording them would return. */
int craw_batch_check(CRAWBatch *b, const CRAWBatchFrame *frames, uint32_t nb);

/* hipFree and friends on teardown paths, whose errors change nothing. */
#define CRAW_IGNORE(x) do { hipError_t craw_ignored_ = (x); (void)craw_ignored_; } while (0)

#endif /* CRAW_INTERNAL_H */
