/*
 * libcraw for MI355X: context and logging.
 *
 * Interface of include/craw/craw.h in cyanreg/libcraw (craw.h:69-105). The
 * ownership model is the reference's with HIP in place of Vulkan:
 *
 *   - the library never creates a stream and never submits work to one it
 *     was not handed; every GPU call is enqueued on the caller's
 *     hipStream_t (craw_pipeline_record() in <craw/pipeline.h>);
 *   - it never synchronises the caller's stream; reading a histogram or a
 *     stage time is only valid once the caller has waited for the work;
 *   - it allocates its own intermediates, temporal state and pinned
 *     histogram slots on the context's device, with hipMalloc and
 *     hipHostMalloc, at pipeline creation and nowhere else.
 *
 * A context or a pipeline is not thread safe without external locking.
 */
#ifndef CRAW_H
#define CRAW_H

#include <stdint.h>
#include <stdarg.h>

#include <craw/utils.h>

CRAW_BEGIN_DECLS

typedef struct CRAWContext CRAWContext;

/* Same values as craw.h:71-80 of the reference. */
enum CRAWLogLevel {
    CRAW_LOG_QUIET    = -(1 << 0),
    CRAW_LOG_FATAL    =  (0 << 0),
    CRAW_LOG_ERROR    = +(1 << 0),
    CRAW_LOG_WARN     = +(1 << 1),
    CRAW_LOG_INFO     = +(1 << 2),
    CRAW_LOG_VERBOSE  = +(1 << 3),
    CRAW_LOG_DEBUG    = +(1 << 4),
    CRAW_LOG_TRACE    = +(1 << 5),
};

/* Context options; zero means the default for every field. */
typedef struct CRAWContextOptions {
    void *log_opaque;

    /* Receives every message at or above log_level. NULL sends warnings
     * and errors to stderr. */
    void (*log_cb)(void *log_opaque, enum CRAWLogLevel level,
                   const char *format, va_list args);

    enum CRAWLogLevel log_level;

    /* Reserved, must be zero. Keeps sizeof at 1024 as in the reference. */
    uint8_t padding[1024 - 1*4 - 2*8];
} CRAWContextOptions;

/* Frees a context made by craw_context_create() (<craw/hip.h>) and sets
 * *ctx to NULL. Every pipeline must already be freed. */
CRAW_API void craw_context_free(CRAWContext **ctx);

/* Logs through the context's callback; `ctx` may be NULL. */
CRAW_API void craw_log(void *ctx, enum CRAWLogLevel level, const char *fmt, ...) craw_printf_format(3, 4);

/* The library's version as (major << 16) | (minor << 8) | micro. */
#define CRAW_VERSION_MAJOR 0
#define CRAW_VERSION_MINOR 1
#define CRAW_VERSION_MICRO 0
#define CRAW_GET_VER_MAJOR(v) (((v) >> 16) & 0xFF)
#define CRAW_GET_VER_MINOR(v) (((v) >>  8) & 0xFF)
#define CRAW_GET_VER_MICRO(v) (((v) >>  0) & 0xFF)
CRAW_API uint32_t craw_version_int(void);

CRAW_END_DECLS

#endif /* CRAW_H */
