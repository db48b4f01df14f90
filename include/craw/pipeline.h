/*
 * libcraw for MI355X: the raw development pipeline.
 *
 * Interface of include/craw/pipeline.h in cyanreg/libcraw. Every name,
 * flag value, enum value and struct layout of the reference is kept
 * (CRAWPipelineParams is 256 bytes and CRAWFrameParams 264, as
 * pipeline.c:252-253 of the reference asserts); the one change is the
 * binding of <craw/hip.h>: craw_pipeline_record() takes a hipStream_t and
 * device pointers where the reference takes a VkCommandBuffer and images.
 *
 * The stages this library develops, in order:
 *
 *   src (R16_UNORM Bayer mosaic, RGGB)
 *     -> [temporal denoise against the previous frame's state]
 *     -> [non-local means on the four CFA planes, variance stabilised]
 *     -> black/white normalisation, white balance, clip at 1.0, or
 *        [highlight reconstruction: clipped sites rebuilt from their
 *         neighbours and a multi-scale colour field, balanced, unclipped,
 *         stored as R16F]
 *     -> bilinear, MHC or RCD debayer to camera RGB
 *     -> [false colour suppression: 1 to 5 passes of a 3x3 median and a
 *         3x3 box on R - G and B - G]
 *     -> [chroma denoise of the quarter resolution chroma band]
 *     -> [auto white balance histogram]
 *     -> camera matrix, gain, exposure, XYZ to the output space
 *     -> [ACES 1.3 gamut compression, ACEScg output only]
 *   dst (RGBA16F or RGBA32F)
 *
 * The reference's lens, defringe and sharpening stages are not part of
 * this library; their pipeline flags make craw_pipeline_create() fail with
 * CRAW_ERROR(ENOTSUP) and their frame fields are ignored.
 */
#ifndef CRAW_PIPELINE_H
#define CRAW_PIPELINE_H

#include <craw/hip.h>
#include <craw/profile.h>
#include <craw/awb.h>
#include <craw/lens.h>

CRAW_BEGIN_DECLS

typedef struct CRAWPipeline CRAWPipeline;

enum CRAWDebayer {
    CRAW_DEBAYER_BILINEAR = 0,
    CRAW_DEBAYER_MHC,
    CRAW_DEBAYER_RCD,
    CRAW_DEBAYER_NB,
};

/* Pipeline flags: the reference's values (pipeline.h:47-170). */
#define CRAW_PIPELINE_FLAG_AWB_HISTOGRAM            (1 << 0)
#define CRAW_PIPELINE_FLAG_DENOISE                  (1 << 1)
#define CRAW_PIPELINE_FLAG_LENS                     (1 << 2)  /* ENOTSUP here */
#define CRAW_PIPELINE_FLAG_FALSE_COLOR              (1 << 3)
#define CRAW_PIPELINE_FLAG_CHROMA_DENOISE           (1 << 4)
#define CRAW_PIPELINE_FLAG_HIGHLIGHT_RECONSTRUCTION (1 << 5)
#define CRAW_PIPELINE_FLAG_DEFRINGE                 (1 << 6)  /* ENOTSUP here */
#define CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE         (1 << 7)
#define CRAW_PIPELINE_FLAG_SHARPEN                  (1 << 8)  /* ENOTSUP here */
#define CRAW_PIPELINE_FLAG_PROFILE                  (1 << 9)

/* This library's own addition, with no counterpart in the reference: run
 * the non-local means with its patch distances and weights in packed FP16
 * (weights and weighted sums accumulated in FP32) instead of the FP32
 * form, which reproduces the CPU reference bit for bit. The FP16 form
 * stays within 2^-11 of the clip level of an FP64 evaluation. */
#define CRAW_PIPELINE_FLAG_NLM_FP16                 (1 << 16)

/* Also this library's own: the NLM in FP32 with its 3x3 sums reordered
 * (each squared difference computed once, summed across lanes and down
 * registers) and the hardware exp2. Within a few ulps of the bit-exact
 * form in every weight; no longer bit for bit the CPU reference. */
#define CRAW_PIPELINE_FLAG_NLM_FAST                 (1 << 17)

#define CRAW_FALSE_COLOR_MAX_PASSES 5
#define CRAW_SHARPEN_MAX_RADIUS     2.0f
#define CRAW_SHARPEN_MAX_ITERATIONS 100

/* Recordings whose stage times a profiling pipeline keeps. */
#define CRAW_PIPELINE_PROFILE_SLOTS 4

typedef struct CRAWPipelineParams {
    uint32_t width, height;         /* even, at least 8 */
    enum CRAWCfa cfa;               /* CRAW_CFA_RGGB only */
    enum CRAWDebayer debayer;
    enum CRAWColorSpace out_space;  /* default for frames that set none */
    enum CRAWFormat out_format;     /* RGBA16F or RGBA32F */
    uint32_t flags;                 /* CRAW_PIPELINE_FLAG_* */
    uint32_t nb_awb_slots;          /* histogram readback slots, 1..4 (0: 1) */
    uint8_t padding[256 - 8*4];
} CRAWPipelineParams;

/* Frame flags: the reference's values (pipeline.h:172-249). */
#define CRAW_FRAME_FLAG_CLAMP_NEGATIVE           (1 << 0)
#define CRAW_FRAME_FLAG_LENS_KEEP_ALL            (1 << 1)  /* ignored */
#define CRAW_FRAME_FLAG_NO_GAMUT_COMPRESSION     (1 << 2)
#define CRAW_FRAME_FLAG_HIGHLIGHT_RECONSTRUCTION (1 << 3)  /* with the pipeline flag */
#define CRAW_FRAME_FLAG_TEMPORAL_RESET           (1 << 4)

typedef struct CRAWFrameParams {
    const CRAWProfile *profile;

    enum CRAWWbMode wb_mode;
    CRAWIlluminant  illuminant;     /* for CRAW_WB_REBALANCE */

    float exposure_stops;

    uint32_t flags;                 /* CRAW_FRAME_FLAG_* */

    /* The frame's output space when out_space_set, else the pipeline's. */
    enum CRAWColorSpace out_space;
    bool out_space_set;

    /* Spatial denoise strength in units of the modelled noise; 0 is off. */
    float denoise_strength;

    /* Used when a > 0; else the profile's table at its as_shot_ei. */
    CRAWNoiseModel noise_override;

    /* False colour passes, 0 (off) to CRAW_FALSE_COLOR_MAX_PASSES; more is
     * CRAW_ERROR(EINVAL) with CRAW_PIPELINE_FLAG_FALSE_COLOR. */
    uint32_t false_color_passes;

    /* Chroma denoise, 0 (off) to 1. */
    float denoise_chroma_strength;

    CRAWLensCorrection lens;        /* ignored */

    float defringe_strength;        /* ignored */

    /* Temporal denoise alpha_max, 0 (off) to 1, and whether this recording
     * develops the last recorded frame again. */
    float denoise_temporal_strength;
    bool denoise_temporal_replay;

    float    sharpen_radius;        /* ignored */
    uint32_t sharpen_iterations;    /* ignored */
    float    sharpen_contrast;      /* ignored */

    float rotation;                 /* ignored */

    uint8_t padding[256 - 4*4 - 8 - 12 - 64 - 8 - 4 - 4 - 4 - 4 - 4 - 12 - 4];
} CRAWFrameParams;

/* Defaults: RGGB, RCD, XYZ D65, RGBA16F, no flags; width and height 0. */
CRAW_API void craw_pipeline_params_init(CRAWPipelineParams *p);

/* Defaults: as shot, the pipeline's space, nothing optional; profile NULL. */
CRAW_API void craw_frame_params_init(CRAWFrameParams *fp);

/* Allocates the pipeline's device memory for exactly width x height.
 * CRAW_ERROR(ENOTSUP) for a non-RGGB CFA or a flag of a stage this library
 * does not have, CRAW_ERROR(EINVAL) for a bad size or format. */
CRAW_API int craw_pipeline_create(CRAWContext *ctx,
                                  const CRAWPipelineParams *params,
                                  CRAWPipeline **pipe);

/* Frees the pipeline and sets *pipe to NULL. Work recorded with it must
 * have completed. */
CRAW_API void craw_pipeline_free(CRAWPipeline **pipe);

/* Enqueues the development of `src` into `dst` on `stream`.
 *
 * src is CRAW_FORMAT_R16_UNORM, dst the pipeline's out_format, both the
 * pipeline's size, not aliasing. The pipeline's intermediates are single
 * buffered: every recording of one pipeline must go to one stream, in
 * call order. Nothing is synchronised; the caller waits on the stream.
 *
 * Temporal state (CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE) follows the
 * reference's rules (pipeline.h:457-484): nothing is valid at creation
 * or after CRAW_FRAME_FLAG_TEMPORAL_RESET; a recording reads what the last
 * one wrote and writes the other set; a replay reads what the last
 * recording read and overwrites what it wrote; a next frame with the pass
 * off leaves no state, a replay with it off leaves the state alone; a
 * failing call changes nothing.
 *
 * Errors: CRAW_ERROR(EINVAL) for bad images or parameters,
 * CRAW_ERROR(ENOTSUP) when a denoise strength is positive and no noise
 * model is available. */
CRAW_API int craw_pipeline_record(CRAWPipeline *pipe, CRAWStream stream,
                                  const CRAWImage *src, const CRAWImage *dst,
                                  const CRAWFrameParams *fp);

/* One frame developed with no pipeline and no allocation: prep,
 * debayer, colour (and gamut compression) into `dst`, and, when
 * `histogram` is not NULL, the AWB histogram into it (3 x 128^2 u32:
 * count, weight, weight_bright; device-addressable, 16-byte aligned,
 * complete when `stream` reaches the end of this call's work). The same
 * kernels and constants as craw_pipeline_record() on a pipeline built with
 * `debayer`, `out_space`, dst's format and CRAW_PIPELINE_FLAG_AWB_HISTOGRAM
 * (when `histogram` is set) and no other stage flag: the frame's denoise,
 * chroma, false-colour and highlight settings are ignored, as such a
 * pipeline ignores them.
 *
 * It has no counterpart in the reference. It is for callers that cannot
 * allocate device memory: code under HIP graph capture, serving loops and
 * harnesses that forbid hipMalloc. It calls no allocator: the histogram's
 * scratch (192 KiB plus 8 bytes an 8x8 block, at most 2^19 blocks, e.g.
 * 8192x4096 or 6000x4000; 4.2 MiB in all) is a device array each GPU the
 * library runs on holds until the process ends, shared by every call on
 * that GPU: calls that overlap on one GPU race. A frame with more blocks,
 * with a histogram, fails with CRAW_ERROR(EINVAL); the other errors are
 * craw_pipeline_record()'s. */
CRAW_API int craw_develop_noalloc(CRAWContext *ctx, CRAWStream stream, const CRAWImage *src,
                                  const CRAWImage *dst, const CRAWFrameParams *fp,
                                  enum CRAWDebayer debayer, enum CRAWColorSpace out_space,
                                  uint32_t *histogram);

/* The histogram of the last recording (or of `frames_back` recordings
 * before it, up to nb_awb_slots - 1), in pinned host memory owned by the
 * pipeline. Valid once the caller has synchronised that recording.
 * CRAW_ERROR(EINVAL) without CRAW_PIPELINE_FLAG_AWB_HISTOGRAM. */
CRAW_API int craw_pipeline_get_awb_histogram(CRAWPipeline *pipe,
                                             CRAWAwbHistogram *out);
CRAW_API int craw_pipeline_get_awb_histogram_n(CRAWPipeline *pipe, uint32_t frames_back,
                                               CRAWAwbHistogram *out);

typedef struct CRAWStageTime {
    const char *name;   /* the kernel's stage name, owned by the pipeline */
    uint32_t count;     /* launches of it in the recording */
    double   ms;        /* their GPU time, summed (hipEvent pairs) */
} CRAWStageTime;

/* Stage times of the last recording (or `frames_back` before it, up to
 * CRAW_PIPELINE_PROFILE_SLOTS - 1): returns the number of stages and
 * writes the first `max`. CRAW_ERROR(EAGAIN) before the work completed,
 * CRAW_ERROR(EINVAL) without CRAW_PIPELINE_FLAG_PROFILE. */
CRAW_API int craw_pipeline_get_stage_times(CRAWPipeline *pipe, CRAWStageTime *out,
                                           uint32_t max);
CRAW_API int craw_pipeline_get_stage_times_n(CRAWPipeline *pipe, uint32_t frames_back,
                                             CRAWStageTime *out, uint32_t max);

CRAW_END_DECLS

#endif /* CRAW_PIPELINE_H */
