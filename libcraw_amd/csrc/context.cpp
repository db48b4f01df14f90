/*
 * libcraw for MI355X: context and logging (lib/craw.c:20-43, 453-519 of
 * cyanreg/libcraw, with the Vulkan device checks replaced by a HIP device
 * query).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

extern "C" {

void craw_vlog_ctx(CRAWContext *ctx, enum CRAWLogLevel level, const char *fmt, va_list ap)
{
    enum CRAWLogLevel max = CRAW_LOG_WARN;

    if (ctx && ctx->opts.log_cb) {
        if (level <= ctx->opts.log_level)
            ctx->opts.log_cb(ctx->opts.log_opaque, level, fmt, ap);
        return;
    }
    if (ctx && ctx->opts.log_level)
        max = ctx->opts.log_level;
    if (level <= max && level != CRAW_LOG_QUIET)
        vfprintf(stderr, fmt, ap);
}

CRAW_API void craw_log(void *ctx, enum CRAWLogLevel level, const char *fmt, ...)
{
    va_list ap;

    va_start(ap, fmt);
    craw_vlog_ctx((CRAWContext *)ctx, level, fmt, ap);
    va_end(ap);
}

CRAW_API uint32_t craw_version_int(void)
{
    return (CRAW_VERSION_MAJOR << 16) | (CRAW_VERSION_MINOR << 8) | CRAW_VERSION_MICRO;
}

CRAW_API int craw_context_create(CRAWContext **out, const CRAWHipInitParams *params,
                                 const CRAWContextOptions *opts)
{
    CRAWContext *ctx;
    hipDeviceProp_t prop;
    int n = 0, ret;

    if (!out || !params)
        return CRAW_ERROR(EINVAL);
    if (hipGetDeviceCount(&n) != hipSuccess || params->device < 0 || params->device >= n)
        return CRAW_ERROR(ENODEV);
    if (hipGetDeviceProperties(&prop, params->device) != hipSuccess)
        return CRAW_ERROR(ENODEV);

    ctx = (CRAWContext *)calloc(1, sizeof(*ctx));
    if (!ctx)
        return CRAW_ERROR(ENOMEM);
    if (opts)
        ctx->opts = *opts;
    ctx->device = params->device;
    ctx->cu_count = prop.multiProcessorCount;
    snprintf(ctx->arch, sizeof(ctx->arch), "%s", prop.gcnArchName);

    /* The library carries gfx950 code objects only. */
    if (strncmp(ctx->arch, "gfx950", 6) != 0) {
        craw_log(ctx, CRAW_LOG_ERROR, "Device %d is %s; this build runs on gfx950 only\n",
                 params->device, ctx->arch);
        free(ctx);
        return CRAW_ERROR(ENOTSUP);
    }

    if (hipSetDevice(ctx->device) != hipSuccess) {
        free(ctx);
        return CRAW_ERROR(ENODEV);
    }
    ret = craw_noalloc_setup(ctx);
    if (ret < 0) {
        free(ctx);
        return ret;
    }
    pthread_mutex_init(&ctx->noise_lock, NULL);
    craw_log(ctx, CRAW_LOG_VERBOSE, "Context on device %d (%s, %d CUs)\n", ctx->device,
             ctx->arch, ctx->cu_count);
    *out = ctx;
    return 0;
}

CRAW_API void craw_context_free(CRAWContext **ctx)
{
    if (!ctx || !*ctx)
        return;
    if ((*ctx)->noise_scratch) {
        CRAW_IGNORE(hipSetDevice((*ctx)->device));
        CRAW_IGNORE(hipFree((*ctx)->noise_scratch));
    }
    pthread_mutex_destroy(&(*ctx)->noise_lock);
    free(*ctx);
    *ctx = NULL;
}

} /* extern "C" */
