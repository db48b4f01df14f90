/*
 * libcraw for MI355X: the serving form.
 *
 * This library's own addition, with no counterpart in cyanreg/libcraw. A
 * server takes frames of many concurrent clip streams and one-off stills,
 * of any sizes, as they arrive, and develops them with continuous
 * batching: a scheduler thread packs the waiting requests into batches
 * (batch.h) on a small set of slots, each slot a batch with its own HIP
 * stream, so that batches launched while others are still on the GPU run
 * beside them. Requests that arrive while every slot is busy wait and join
 * the next batch a freed slot takes.
 *
 * Device memory is admitted against a budget fixed at creation: the slots'
 * intermediates, and each open clip stream's temporal state (10 bytes a
 * mosaic sample), which stays resident between its frames. The caller's
 * source and destination buffers are the caller's.
 *
 * A clip stream's frames are developed in submission order, each reading
 * the temporal state its predecessor wrote: a stream has at most one frame
 * on the GPU, and its next frame joins the first batch packed after that
 * one has completed (so no slot ever waits on another). A frame is
 * developed exactly as craw_pipeline_record() develops it with the same
 * parameters, whichever other frames share its batch.
 */
#ifndef CRAW_SERVE_H
#define CRAW_SERVE_H

#include <craw/batch.h>

CRAW_BEGIN_DECLS

typedef struct CRAWServer CRAWServer;

typedef struct CRAWServerParams {
    uint64_t hbm_budget;            /* bytes of device memory the server may hold */
    uint64_t max_batch_pixels;      /* pixels of all frames of one batch */
    uint32_t nb_slots;              /* batches in flight, 1 to 8 */
    uint32_t max_batch_frames;      /* frames of one batch, 1 to 255 */
    uint32_t max_width, max_height; /* largest frame */
    enum CRAWDebayer debayer;
    enum CRAWColorSpace out_space;
    enum CRAWFormat out_format;
    uint32_t flags;                 /* as CRAWBatchParams.flags */
    uint32_t use_graphs;            /* as CRAWBatchParams.use_graphs */
    uint32_t max_queue;             /* waiting requests; submit fails with EAGAIN beyond */
    uint8_t padding[256 - 8 * 2 - 4 * 10];
} CRAWServerParams;

/* One frame to develop. The frame parameters are copied at submission;
 * the profile they point to, the source and the destination must stay
 * valid until the request completes. */
typedef struct CRAWServeRequest {
    CRAWImage src, dst;
    CRAWFrameParams params;
    int32_t stream;                 /* from craw_server_open_stream(), or -1: a still */
    uint32_t reserved;
    uint64_t tag;                   /* the caller's, returned in the completion */
    /* With CRAW_PIPELINE_FLAG_AWB_HISTOGRAM: NULL, or 3 x 128 x 128 words
     * (count, weight, weight_bright) written before the completion is
     * reported. */
    uint32_t *awb_histogram;
} CRAWServeRequest;

typedef struct CRAWServeCompletion {
    uint64_t tag;
    int32_t status;                 /* 0, or the negative errno the frame failed with */
    int32_t stream;
    /* Milliseconds on the server's clock (steady, from creation): when the
     * request was submitted, when its batch was launched, and when the
     * scheduler saw the batch complete. */
    double submit_ms, launch_ms, done_ms;
    uint32_t batch_frames;          /* frames in the batch that carried it */
    uint32_t slot;
} CRAWServeCompletion;

CRAW_API void craw_server_params_init(CRAWServerParams *p);

/* Allocates the slots' intermediates; CRAW_ERROR(ENOMEM) when they do not
 * fit the budget. Starts the scheduler thread. */
CRAW_API int craw_server_create(CRAWContext *ctx, const CRAWServerParams *params,
                                CRAWServer **server);

/* Waits for every submitted request, then frees everything. */
CRAW_API void craw_server_free(CRAWServer **server);

/* A clip stream of width x height frames, its temporal state resident on
 * the device. CRAW_ERROR(ENOMEM) when the state does not fit the budget. */
CRAW_API int craw_server_open_stream(CRAWServer *server, uint32_t width, uint32_t height,
                                     int32_t *stream);

/* Waits for the stream's requests, then frees its state. */
CRAW_API int craw_server_close_stream(CRAWServer *server, int32_t stream);

/* Queues a request; returns at once. CRAW_ERROR(EINVAL) for a request that
 * can never run (sizes, unknown stream), EAGAIN when max_queue requests
 * already wait. A request that fails when its batch is planned completes
 * with that error in its status. */
CRAW_API int craw_server_submit(CRAWServer *server, const CRAWServeRequest *request);

/* Queues `nb` requests at once, in order, so that the scheduler sees them
 * together: all or none (the first error as craw_server_submit(), EAGAIN
 * when they do not all fit the queue). */
CRAW_API int craw_server_submit_n(CRAWServer *server, const CRAWServeRequest *requests,
                                  uint32_t nb);

/* Up to `max` completions, oldest first; while there are none, waits up to
 * timeout_ms (0: no wait; negative: until one is posted or nothing is left
 * in flight). Returns the number written. */
CRAW_API int craw_server_poll(CRAWServer *server, CRAWServeCompletion *out, uint32_t max,
                              int32_t timeout_ms);

/* Waits until every submitted request has completed (their completions
 * stay to be polled). */
CRAW_API int craw_server_drain(CRAWServer *server);

/* Device memory the server holds now: slots and open streams. */
CRAW_API uint64_t craw_server_get_memory_usage(CRAWServer *server);

CRAW_END_DECLS

#endif /* CRAW_SERVE_H */
