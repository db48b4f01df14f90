/*
 * libcraw for MI355X: the batched form and per-stream temporal state.
 *
 * This library's own addition, with no counterpart in cyanreg/libcraw,
 * whose recordings develop one frame each (include/craw/pipeline.h:450-455
 * of the reference). A batch develops many frames of possibly different sizes in
 * one recording: every stage is one persistent launch over a work-list of
 * tiles drawn from all the frames, so small frames fill the gaps large
 * ones leave and a stage's launch cost is paid once per batch. Each frame
 * is developed exactly as craw_pipeline_record() would develop it with the
 * same parameters (bit for bit in the FP32 NLM form).
 *
 * Temporal denoise in a batch uses a CRAWTemporalState per clip stream,
 * with the pipeline's rules (pipeline.h): a frame carrying a state reads
 * what the last recording of that state wrote and writes its other set,
 * replay and reset included.
 */
#ifndef CRAW_BATCH_H
#define CRAW_BATCH_H

#include <craw/pipeline.h>

CRAW_BEGIN_DECLS

typedef struct CRAWBatch CRAWBatch;
typedef struct CRAWTemporalState CRAWTemporalState;

typedef struct CRAWBatchParams {
    uint32_t max_frames;            /* frames per recording, 1 to 255 */
    uint32_t max_width, max_height; /* largest frame */
    /* CRAW_PIPELINE_FLAG_AWB_HISTOGRAM, _DENOISE, _CHROMA_DENOISE,
     * _TEMPORAL_DENOISE, _PROFILE, _NLM_FP16, _NLM_FAST */
    uint32_t flags;
    uint64_t max_pixels;            /* most pixels of all frames of one recording */
    enum CRAWDebayer debayer;       /* shared by all frames */
    enum CRAWColorSpace out_space;  /* default; frames may set their own */
    enum CRAWFormat out_format;     /* RGBA16F or RGBA32F */
    /* Nonzero: capture each distinct recording shape (the frames' sizes
     * and stages) as a hipGraph once, and replay it for every later
     * recording of the same shape. */
    uint32_t use_graphs;
    uint8_t padding[256 - 4 * 4 - 8 - 4 * 4];
} CRAWBatchParams;

/* One frame of a batch recording. `params` as for craw_pipeline_record();
 * `state` is the frame's clip stream (NULL: no temporal denoise for it). */
typedef struct CRAWBatchFrame {
    CRAWImage src, dst;
    const CRAWFrameParams *params;
    CRAWTemporalState *state;
} CRAWBatchFrame;

CRAW_API void craw_batch_params_init(CRAWBatchParams *p);

/* Allocates the intermediates for max_pixels, in one arena per kind. */
CRAW_API int craw_batch_create(CRAWContext *ctx, const CRAWBatchParams *params,
                               CRAWBatch **batch);
CRAW_API void craw_batch_free(CRAWBatch **batch);

/* Enqueues the development of `nb` frames on `stream`. Fails without side
 * effects (CRAW_ERROR(EINVAL), ENOTSUP as craw_pipeline_record(); ENOSPC
 * when the frames exceed max_frames or max_pixels). Recordings of one
 * batch go to one stream in call order; its intermediates are single
 * buffered. */
CRAW_API int craw_batch_record(CRAWBatch *batch, CRAWStream stream,
                               const CRAWBatchFrame *frames, uint32_t nb);

/* The histogram of frame `index` of the last recording; valid once that
 * recording has completed. */
CRAW_API int craw_batch_get_awb_histogram(CRAWBatch *batch, uint32_t index,
                                          CRAWAwbHistogram *out);

/* Stage times of the last recording (CRAW_PIPELINE_FLAG_PROFILE), as
 * craw_pipeline_get_stage_times(). */
CRAW_API int craw_batch_get_stage_times(CRAWBatch *batch, CRAWStageTime *out, uint32_t max);

/* Picks the launch shape of each stage kind (one workgroup per tile, or a
 * persistent 1 to 16 workgroups per CU) and of the denoise stage (its
 * CRAW_BATCH_DENOISE_* options) by timing each candidate on `frames`
 * (developing them `iterations` times per candidate on `stream`,
 * synchronously), and keeps the fastest for later recordings. Returns the
 * chosen configuration's time per recording in microseconds. Outputs are
 * the same whichever candidate runs. */
CRAW_API int craw_batch_autotune(CRAWBatch *batch, CRAWStream stream,
                                 const CRAWBatchFrame *frames, uint32_t nb,
                                 uint32_t iterations);

/* Denoise launch options: the pairwise band kernel (exact and precision
 * forms, without history) in 57-row bands instead of 48; the fast form's
 * 7 search rows unrolled. */
#define CRAW_BATCH_DENOISE_TALL   (1 << 0)
#define CRAW_BATCH_DENOISE_UNROLL (1 << 1)

/* Launch shapes of later recordings, set directly (as the autotuner
 * does): persistent workgroups per CU for the denoise, debayer, chroma
 * blur and AWB histogram stages (0: one workgroup per tile; at most 64),
 * and the CRAW_BATCH_DENOISE_* options. CRAW_ERROR(EINVAL) for a value out
 * of range. craw_batch_get_launch_shape() reads the current ones. */
CRAW_API int craw_batch_set_launch_shape(CRAWBatch *batch, const int wg_per_cu[4],
                                         int denoise_options);
CRAW_API int craw_batch_get_launch_shape(CRAWBatch *batch, int wg_per_cu[4],
                                         int *denoise_options);

/* Temporal state of one clip stream: two sets of (level, variance) for a
 * width x height mosaic, 5 bytes a sample each, on the context's device. */
CRAW_API int craw_temporal_state_create(CRAWContext *ctx, uint32_t width, uint32_t height,
                                        CRAWTemporalState **state);
CRAW_API void craw_temporal_state_free(CRAWTemporalState **state);
/* Forgets the state: the next frame is developed as a first frame. */
CRAW_API void craw_temporal_state_reset(CRAWTemporalState *state);

CRAW_END_DECLS

#endif /* CRAW_BATCH_H */
