/*
 * libcraw for MI355X: the device binding.
 *
 * Replaces include/craw/vulkan.h of cyanreg/libcraw (vulkan.h:68-160). Where
 * the reference is handed Vulkan instance/device handles and images, this
 * build is handed a HIP device ordinal and device pointers:
 *
 *   reference                               here
 *   craw_context_create(ctx, VkInitParams)  craw_context_create(ctx, CRAWHipInitParams)
 *   CRAWImage{VkImage, view, VkFormat, ..}  CRAWImage{device pointer, pitch, CRAWFormat, ..}
 *   VkCommandBuffer                         hipStream_t (craw_pipeline_record)
 *
 * Nothing here needs the HIP headers: a stream is passed as an opaque
 * pointer, which is what hipStream_t is.
 */
#ifndef CRAW_HIP_H
#define CRAW_HIP_H

#include <craw/craw.h>

CRAW_BEGIN_DECLS

/* Pixel formats of the images the pipeline reads and writes. The values
 * are this library's own; they stand where the reference takes a VkFormat
 * and keep the same 4-byte size in every public struct. */
enum CRAWFormat {
    CRAW_FORMAT_UNDEFINED = 0,
    CRAW_FORMAT_R16_UNORM,            /* Bayer mosaic, one uint16 a sample */
    CRAW_FORMAT_R16G16B16A16_SFLOAT,  /* RGBA, IEEE half a component */
    CRAW_FORMAT_R32G32B32A32_SFLOAT,  /* RGBA, float a component */
};

/* An opaque hipStream_t. NULL is the device's null stream. */
typedef void *CRAWStream;

/* What craw_context_create() binds to. */
typedef struct CRAWHipInitParams {
    /* HIP device ordinal, as for hipSetDevice(). The device must be a
     * gfx950 (MI355X); the library carries no code for anything else. */
    int32_t device;

    /* Reserved, must be zero. */
    uint8_t padding[252];
} CRAWHipInitParams;

/* Creates a context on `params->device`. Fails with CRAW_ERROR(ENODEV)
 * when the ordinal names no device and CRAW_ERROR(ENOTSUP) when the device
 * is not a gfx950. Replaces craw_context_create() of vulkan.h:143-145. */
CRAW_API int craw_context_create(CRAWContext **ctx,
                                 const CRAWHipInitParams *params,
                                 const CRAWContextOptions *opts);

/* One 2D image in device memory: `data` points at the first pixel of the
 * first row, rows are `pitch` bytes apart (at least width * the format's
 * pixel size, and a multiple of it; `data` aligned to it). Replaces
 * CRAWImage of vulkan.h:150-157; there are no layouts to track. */
typedef struct CRAWImage {
    void            *data;
    size_t           pitch;
    enum CRAWFormat  format;
    uint32_t         width, height;
} CRAWImage;

/* The noise fit of craw_noise_estimate() with the frame's measurement on
 * the device: `src` is the R16_UNORM mosaic in device memory. This
 * library's addition (the reference measures on the host) for callers
 * whose frames are already on the GPU. The measurement is recorded on
 * `stream` and waited for (the host fit needs it). Its scratch, 8 bytes a
 * sample plus 8 KiB a bin, belongs to the context: allocated by the first
 * call that needs more, kept until craw_context_free(); calls on one
 * context take turns on it (a lock). The bins' counts and medians
 * are the host measurement's; their sums of means are exact here (the
 * host sums them in order in double), so the model agrees with
 * craw_noise_estimate() to a few ulp. */
struct CRAWProfile;
struct CRAWNoiseEstimateParams;
struct CRAWNoiseEstimate;
CRAW_API int craw_noise_estimate_image(CRAWContext *ctx, CRAWStream stream,
                                       const CRAWImage *src,
                                       const struct CRAWProfile *profile,
                                       const struct CRAWNoiseEstimateParams *params,
                                       struct CRAWNoiseEstimate *out);

CRAW_END_DECLS

#endif /* CRAW_HIP_H */
