/*
 * libcraw for MI355X: craw_develop_noalloc(), one frame developed without a
 * pipeline object and without allocating anything.
 *
 * The frame's only scratch is the AWB histogram's: the 3 x 128^2 u32
 * accumulator and one (bin, weight) sample per 8x8 block. Here they live in
 * `scratch`, a file-scope device array each GPU holds from the library's
 * load until the process ends; its address on a device is looked up once,
 * by craw_context_create() on that device (craw_noalloc_setup()), and kept
 * per device. The stages are the plain form's own launches (prep + debayer
 * fused with the colour stage and the block samples, the histogram, its
 * export), on the caller's stream, with the same constants planned by
 * craw_frame_plan().
 *
 * Every call on one device shares its array: calls that overlap on one
 * GPU (two streams, two threads) race on it. The accumulator is cleared at
 * the start of every call by a kernel, so nothing of an earlier call (or of
 * one that failed part way) is counted.
 */
#include <pthread.h>
#include <string.h>

#include "internal.h"

/* 8x8 AWB blocks one call can sample: 8192x4096, 7680x4320 and 6000x4000
 * frames fit. */
#define NOALLOC_MAX_BLOCKS (1u << 19)
#define NOALLOC_HIST_BYTES ((size_t)CRAW_HIST_WORDS * sizeof(uint32_t))
#define NOALLOC_BYTES (NOALLOC_HIST_BYTES + (size_t)NOALLOC_MAX_BLOCKS * sizeof(uint2))
#define NOALLOC_MAX_DEVICES 64

static_assert(NOALLOC_HIST_BYTES % 256 == 0, "the samples start 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static pthread_mutex_t scratch_lock = PTHREAD_MUTEX_INITIALIZER;
static void *scratch_ptr[NOALLOC_MAX_DEVICES];

int craw_noalloc_setup(CRAWContext *ctx)
{
    void *p = NULL;
    hipError_t e;

    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES)
        return CRAW_ERROR(ENODEV);
    pthread_mutex_lock(&scratch_lock);
    if (!scratch_ptr[ctx->device]) {
        /* the device is current: the caller set it */
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            pthread_mutex_unlock(&scratch_lock);
            craw_log(ctx, CRAW_LOG_ERROR, "Device %d: no address for the library's scratch "
                     "array: %s\n", ctx->device, hipGetErrorString(e));
            return CRAW_ERROR(EIO);
        }
        scratch_ptr[ctx->device] = p;
    }
    pthread_mutex_unlock(&scratch_lock);
    return 0;
}

extern "C" CRAW_API int craw_develop_noalloc(CRAWContext *ctx, CRAWStream stream,
                                             const CRAWImage *src, const CRAWImage *dst,
                                             const CRAWFrameParams *fp, enum CRAWDebayer debayer,
                                             enum CRAWColorSpace out_space, uint32_t *histogram)
{
    const hipStream_t s = (hipStream_t)stream;
    CRAWStageConfig cfg;
    CRAWFrameBuffers buf;
    CRAWFramePlan plan;
    unsigned char *base;
    hipError_t e;
    int ret;

    if (!ctx || !src || !dst || !fp || !fp->profile)
        return CRAW_ERROR(EINVAL);
    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES)
        return CRAW_ERROR(ENODEV);
    pthread_mutex_lock(&scratch_lock);
    base = (unsigned char *)scratch_ptr[ctx->device];
    pthread_mutex_unlock(&scratch_lock);
    if (!base)
        return CRAW_ERROR(ENODEV);
    if (histogram && ((uintptr_t)histogram & 15)) {
        craw_log(ctx, CRAW_LOG_ERROR, "craw_develop_noalloc: the histogram must be 16-byte "
                 "aligned\n");
        return CRAW_ERROR(EINVAL);
    }
    if (histogram && craw_awb_blocks(src->width, src->height) > NOALLOC_MAX_BLOCKS) {
        craw_log(ctx, CRAW_LOG_ERROR, "craw_develop_noalloc: a %ux%u frame has %zu 8x8 blocks; "
                 "the histogram's scratch holds %u\n", src->width, src->height,
                 craw_awb_blocks(src->width, src->height), NOALLOC_MAX_BLOCKS);
        return CRAW_ERROR(EINVAL);
    }
    ret = craw_stage_config(ctx, histogram ? CRAW_PIPELINE_FLAG_AWB_HISTOGRAM : 0, debayer,
                            out_space, dst->format, &cfg);
    if (ret < 0)
        return ret;
    memset(&buf, 0, sizeof(buf));
    buf.hist = (uint32_t *)base;
    buf.awb_samples = (uint2 *)(base + NOALLOC_HIST_BYTES);
    ret = craw_frame_plan(ctx, cfg, buf, src->width, src->height, src, dst, fp, &plan);
    if (ret < 0)
        return ret;
    if (hipSetDevice(ctx->device) != hipSuccess)
        return CRAW_ERROR(ENODEV);

    e = hipSuccess;
    if (histogram)
        e = craw_launch_zero_u32(buf.hist, CRAW_HIST_WORDS, s);
    if (e == hipSuccess)
        e = craw_launch_debayer(false, plan.f, s);
    if (e == hipSuccess && histogram)
        e = craw_launch_awb_hist(plan.f, s);
    if (e == hipSuccess && histogram)
        e = craw_launch_hist_export(buf.hist, histogram, CRAW_HIST_WORDS, s);
    if (e != hipSuccess) {
        craw_log(ctx, CRAW_LOG_ERROR, "craw_develop_noalloc: launch failed: %s\n",
                 hipGetErrorString(e));
        return CRAW_ERROR(EIO);
    }
    return 0;
}
