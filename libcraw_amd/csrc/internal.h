/*
 * libcraw for MI355X: internal state of contexts and pipelines.
 */
#ifndef CRAW_INTERNAL_H
#define CRAW_INTERNAL_H

#include <hip/hip_runtime.h>

#include <craw/pipeline.h>
#include "hostmath.h"
#include "kernels.h"

struct CRAWContext {
    int device;
    CRAWContextOptions opts;
    char arch[64];
    int cu_count;
};

#define CRAW_MAX_STAGES 8

/* Everything one recording enqueues, worked out on the host before any
 * of it is launched, so that a failing record changes nothing. */
struct CRAWFramePlan {
    KPrep prep;
    KColor col;
    KOut out;
    KInter inter;
    KDenoise dn;
    KDenoiseState st;
    int dn_mode;                /* 0: no denoise */
    bool chroma;
    float chroma_strength;
    int t_read, t_write;        /* temporal state sets, -1: none */
    bool temporal, replay;
    const uint16_t *src;
    int src_pitch;              /* uint16 elements */
    int method;
};

struct CRAWPipeline {
    CRAWContext *ctx;
    CRAWPipelineParams params;
    bool awb, has_denoise, has_temporal, has_chroma, profile, fp16;

    uint16_t *clean;            /* denoised mosaic, W x H */
    float *tn_level[2];         /* temporal state sets, 4 planes of w2 x h2 */
    uint8_t *tn_var[2];
    int tn_prev, tn_cur;
    float *inter_rgb;           /* 3 x W x H */
    float *down, *blur;         /* quarter grid, 3 and 2 planes */
    int wq, hq;
    uint32_t *hist;             /* count, weight, bright, 128^2 each */
    uint32_t *hist_host[4];     /* pinned readback slots */
    uint32_t nb_slots;
    uint64_t nb_recorded;

    /* profiling: a pair of events around every launch, four recordings */
    hipEvent_t ev[CRAW_PIPELINE_PROFILE_SLOTS][CRAW_MAX_STAGES][2];
    const char *ev_name[CRAW_PIPELINE_PROFILE_SLOTS][CRAW_MAX_STAGES];
    uint32_t ev_nb[CRAW_PIPELINE_PROFILE_SLOTS];
    uint32_t prof_slot;
    uint64_t prof_recorded;
    CRAWStageTime agg[CRAW_MAX_STAGES];
};

#define CRAW_HIST_WORDS (3u * CRAW_AWB_HIST_BINS * CRAW_AWB_HIST_BINS)

extern "C" {
void craw_vlog_ctx(CRAWContext *ctx, enum CRAWLogLevel level, const char *fmt, va_list ap);

/* Plans a frame (validation, constants, state selection); 0 or a negative
 * errno, nothing changed on failure. Shared by the plain and the batched
 * forms. */
int craw_pipeline_plan(CRAWPipeline *p, const CRAWImage *src, const CRAWImage *dst,
                       const CRAWFrameParams *fp, CRAWFramePlan *plan);

/* Enqueues a planned frame on `stream`; a launch failure is reported as
 * CRAW_ERROR(EIO). */
int craw_pipeline_enqueue(CRAWPipeline *p, const CRAWFramePlan *plan, hipStream_t stream);

/* Moves the temporal indices and the histogram slot after a successful
 * enqueue. */
void craw_pipeline_commit(CRAWPipeline *p, const CRAWFramePlan *plan);
}

/* hipFree and friends on teardown paths, whose errors change nothing. */
#define CRAW_IGNORE(x) do { hipError_t craw_ignored_ = (x); (void)craw_ignored_; } while (0)

#endif /* CRAW_INTERNAL_H */
