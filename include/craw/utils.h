/*
 * libcraw for MI355X: common macros of the public headers.
 *
 * Interface of include/craw/utils.h in cyanreg/libcraw (utils.h:38-60): the
 * same macro names and meanings, so that callers written against it compile
 * unchanged. Errors are negative errno values throughout the library.
 */
#ifndef CRAW_UTILS_H
#define CRAW_UTILS_H

#include <stddef.h>
#include <stdint.h>
#include <stdbool.h>
#include <errno.h>

/* A negative errno, the return value of every failing call. */
#define CRAW_ERROR(err) (-(err))

#define CRAW_MIN(v1, v2) ((v1) < (v2) ? (v1) : (v2))
#define CRAW_MAX(v1, v2) ((v1) > (v2) ? (v1) : (v2))
#define CRAW_CLAMP(v, lo, hi) CRAW_MIN(CRAW_MAX((v), (lo)), (hi))
#define CRAW_ARRAY_ELEMS(arr) (sizeof(arr) / sizeof((arr)[0]))

/* Marks an exported symbol of libcraw.so. */
#ifndef CRAW_API
#  if defined(__GNUC__) || defined(__clang__)
#    define CRAW_API __attribute__((visibility("default")))
#  else
#    define CRAW_API
#  endif
#endif

#if defined(__GNUC__) || defined(__clang__)
#define craw_printf_format(fmtpos, attrpos) __attribute__((__format__(__printf__, fmtpos, attrpos)))
#else
#define craw_printf_format(fmtpos, attrpos)
#endif

#ifdef __cplusplus
#  define CRAW_BEGIN_DECLS extern "C" {
#  define CRAW_END_DECLS }
#else
#  define CRAW_BEGIN_DECLS
#  define CRAW_END_DECLS
#endif

#endif /* CRAW_UTILS_H */
