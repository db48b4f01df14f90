/*
 * libcraw for MI355X: the serving form (include/craw/serve.h).
 *
 * One scheduler thread owns the slots. Its loop:
 *   1. retires every slot whose batch has completed on the device (copies
 *      the histograms out, posts the completions, frees the slot's clip
 *      streams for their next frame);
 *   2. while a slot is free and requests wait, packs a batch from the
 *      queue, oldest first: a request joins when its frame fits the slot's
 *      remaining frame and pixel capacity and its clip stream has no frame
 *      on the GPU or already in this batch (so a stream's frames run one
 *      after another, each on the state its predecessor wrote, with no
 *      waits between slots); a frame that cannot be planned completes at
 *      once with its error;
 *   3. sleeps until a submission arrives, or polls the busy slots' events
 *      every 50 microseconds; with requests waiting and every slot busy it
 *      blocks on the oldest slot's event instead.
 * Every entry point takes the server's mutex; the scheduler holds it while
 * it plans and launches (tens of microseconds a batch) and drops it while
 * it waits.
 */
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include <craw/serve.h>
#include "internal.h"

#define MAX_SLOTS 8

namespace {

struct Req {
    CRAWServeRequest r;
    double submit_ms;
};

struct Stream {
    CRAWTemporalState *ts;      /* NULL without the temporal pass */
    uint32_t width, height;
    uint32_t pending;           /* submitted, not yet completed */
    bool on_gpu;                /* a frame of it is in a running batch */
};

struct Slot {
    CRAWBatch *batch;
    hipStream_t stream;
    hipEvent_t done;
    bool busy;
    double launch_ms;
    std::vector<Req> reqs;
};

} /* namespace */

struct CRAWServer {
    CRAWContext *ctx;
    CRAWServerParams params;
    uint64_t mem_slots, mem_streams;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req> queue;
    std::deque<CRAWServeCompletion> done;
    std::map<int32_t, Stream> streams;
    int32_t next_stream;
    uint64_t submitted, completed;
    bool stop;

    Slot slots[MAX_SLOTS];
    uint32_t nb_slots;
    std::thread thread;
    std::chrono::steady_clock::time_point t0;

    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
            .count();
    }
};

static void complete(CRAWServer *s, const Req &q, int status, double launch_ms, uint32_t nb,
                     uint32_t slot)
{
    CRAWServeCompletion c;

    memset(&c, 0, sizeof(c));
    c.tag = q.r.tag;
    c.status = status;
    c.stream = q.r.stream;
    c.submit_ms = q.submit_ms;
    c.launch_ms = launch_ms;
    c.done_ms = s->now_ms();
    c.batch_frames = nb;
    c.slot = slot;
    s->done.push_back(c);
    s->completed++;
    if (q.r.stream >= 0) {
        auto it = s->streams.find(q.r.stream);
        if (it != s->streams.end())
            it->second.pending--;
    }
}

static void retire(CRAWServer *s, uint32_t i)
{
    Slot &sl = s->slots[i];
    const uint32_t nb = (uint32_t)sl.reqs.size();

    for (uint32_t k = 0; k < nb; k++) {
        const Req &q = sl.reqs[k];

        if (q.r.awb_histogram && (s->params.flags & CRAW_PIPELINE_FLAG_AWB_HISTOGRAM)) {
            CRAWAwbHistogram h;
            const size_t n = (size_t)CRAW_AWB_HIST_BINS * CRAW_AWB_HIST_BINS;

            if (craw_batch_get_awb_histogram(sl.batch, k, &h) == 0) {
                memcpy(q.r.awb_histogram, h.count, n * 4);
                memcpy(q.r.awb_histogram + n, h.weight, n * 4);
                memcpy(q.r.awb_histogram + 2 * n, h.weight_bright, n * 4);
            }
        }
        if (q.r.stream >= 0) {
            auto it = s->streams.find(q.r.stream);
            if (it != s->streams.end())
                it->second.on_gpu = false;
        }
        complete(s, q, 0, sl.launch_ms, nb, i);
    }
    sl.reqs.clear();
    sl.busy = false;
}

static CRAWBatchFrame frame_of(CRAWServer *s, const Req &q)
{
    CRAWBatchFrame f;

    f.src = q.r.src;
    f.dst = q.r.dst;
    f.params = &q.r.params;
    f.state = NULL;
    if (q.r.stream >= 0)
        f.state = s->streams[q.r.stream].ts;
    return f;
}

/* Packs and launches one batch on free slot i; false when nothing could
 * be packed. */
static bool launch(CRAWServer *s, uint32_t i)
{
    Slot &sl = s->slots[i];
    std::vector<CRAWBatchFrame> frames;
    std::vector<int32_t> taken;
    uint64_t px = 0;
    int ret;

    sl.reqs.clear();
    for (auto it = s->queue.begin(); it != s->queue.end() &&
                                     sl.reqs.size() < s->params.max_batch_frames;) {
        const Req &q = *it;
        const uint64_t n = (uint64_t)q.r.src.width * q.r.src.height;
        bool blocked = false;

        if (q.r.stream >= 0) {
            blocked = s->streams[q.r.stream].on_gpu;
            for (int32_t t : taken)
                blocked |= t == q.r.stream;
        }
        if (blocked || px + n > s->params.max_batch_pixels) {
            /* later frames of a blocked stream stay behind it */
            if (q.r.stream >= 0)
                taken.push_back(q.r.stream);
            ++it;
            continue;
        }
        /* a frame that cannot be planned fails alone */
        {
            CRAWBatchFrame f = frame_of(s, q);
            ret = craw_batch_check(sl.batch, &f, 1);
        }
        if (ret < 0) {
            complete(s, q, ret, -1.0, 0, i);
            it = s->queue.erase(it);
            s->cv_done.notify_all();
            continue;
        }
        if (q.r.stream >= 0)
            taken.push_back(q.r.stream);
        px += n;
        sl.reqs.push_back(q);
        it = s->queue.erase(it);
    }
    if (sl.reqs.empty())
        return false;

    for (const Req &q : sl.reqs)
        frames.push_back(frame_of(s, q));
    ret = craw_batch_record(sl.batch, sl.stream, frames.data(), (uint32_t)frames.size());
    if (ret == 0 && hipEventRecord(sl.done, sl.stream) != hipSuccess)
        ret = CRAW_ERROR(EIO);
    if (ret < 0) {
        craw_log(s->ctx, CRAW_LOG_ERROR, "Serving batch of %zu frames failed: %d\n",
                 sl.reqs.size(), ret);
        for (const Req &q : sl.reqs)
            complete(s, q, ret, -1.0, (uint32_t)sl.reqs.size(), i);
        sl.reqs.clear();
        s->cv_done.notify_all();
        return true;
    }
    for (const Req &q : sl.reqs)
        if (q.r.stream >= 0)
            s->streams[q.r.stream].on_gpu = true;
    sl.busy = true;
    sl.launch_ms = s->now_ms();
    return true;
}

static void scheduler(CRAWServer *s)
{
    std::unique_lock<std::mutex> lk(s->mu);

    CRAW_IGNORE(hipSetDevice(s->ctx->device));
    for (;;) {
        bool any_busy = false, any_free = false, retired = false;
        int oldest = -1;

        for (uint32_t i = 0; i < s->nb_slots; i++) {
            Slot &sl = s->slots[i];
            if (sl.busy && hipEventQuery(sl.done) != hipErrorNotReady) {
                retire(s, i);
                retired = true;
            }
        }
        if (retired)
            s->cv_done.notify_all();

        for (uint32_t i = 0; i < s->nb_slots && !s->queue.empty(); i++)
            if (!s->slots[i].busy)
                launch(s, i);

        for (uint32_t i = 0; i < s->nb_slots; i++) {
            Slot &sl = s->slots[i];
            if (sl.busy) {
                any_busy = true;
                if (oldest < 0 || sl.launch_ms < s->slots[oldest].launch_ms)
                    oldest = (int)i;
            } else {
                any_free = true;
            }
        }
        if (!any_busy && s->queue.empty()) {
            if (s->stop)
                break;
            s->cv_work.wait(lk);
        } else if (any_busy && !s->queue.empty() && !any_free) {
            hipEvent_t ev = s->slots[oldest].done;

            lk.unlock();
            CRAW_IGNORE(hipEventSynchronize(ev));
            lk.lock();
        } else {
            /* slots busy and free with nothing to pack for the free ones:
             * poll the busy slots */
            s->cv_work.wait_for(lk, std::chrono::microseconds(50));
        }
    }
}

extern "C" {

CRAW_API void craw_server_params_init(CRAWServerParams *p)
{
    if (!p)
        return;
    memset(p, 0, sizeof(*p));
    p->hbm_budget = 64ull << 30;
    p->nb_slots = 2;
    p->max_batch_frames = 16;
    p->debayer = CRAW_DEBAYER_RCD;
    p->out_space = CRAW_CS_XYZ_D65;
    p->out_format = CRAW_FORMAT_R32G32B32A32_SFLOAT;
    p->max_queue = 1024;
}

CRAW_API void craw_server_free(CRAWServer **ps)
{
    CRAWServer *s;

    if (!ps || !*ps)
        return;
    s = *ps;
    if (s->thread.joinable()) {
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        s->thread.join();
    }
    CRAW_IGNORE(hipSetDevice(s->ctx->device));
    for (uint32_t i = 0; i < MAX_SLOTS; i++) {
        Slot &sl = s->slots[i];
        if (sl.stream)
            CRAW_IGNORE(hipStreamSynchronize(sl.stream));
        craw_batch_free(&sl.batch);
        if (sl.done)
            CRAW_IGNORE(hipEventDestroy(sl.done));
        if (sl.stream)
            CRAW_IGNORE(hipStreamDestroy(sl.stream));
    }
    for (auto &kv : s->streams)
        craw_temporal_state_free(&kv.second.ts);
    delete s;
    *ps = NULL;
}

CRAW_API int craw_server_create(CRAWContext *ctx, const CRAWServerParams *params,
                                CRAWServer **out)
{
    CRAWBatchParams bp;
    CRAWServer *s;
    uint64_t fp;
    int ret;

    if (!ctx || !params || !out)
        return CRAW_ERROR(EINVAL);
    if (!params->nb_slots || params->nb_slots > MAX_SLOTS || !params->max_queue)
        return CRAW_ERROR(EINVAL);
    craw_batch_params_init(&bp);
    bp.max_frames = params->max_batch_frames;
    bp.max_width = params->max_width;
    bp.max_height = params->max_height;
    bp.max_pixels = params->max_batch_pixels;
    bp.debayer = params->debayer;
    bp.out_space = params->out_space;
    bp.out_format = params->out_format;
    bp.flags = params->flags;
    bp.use_graphs = params->use_graphs;
    ret = craw_batch_footprint(ctx, &bp, &fp);
    if (ret < 0)
        return ret;
    if (fp * params->nb_slots > params->hbm_budget) {
        craw_log(ctx, CRAW_LOG_ERROR, "%u slots need %llu bytes, over the budget of %llu\n",
                 params->nb_slots, (unsigned long long)(fp * params->nb_slots),
                 (unsigned long long)params->hbm_budget);
        return CRAW_ERROR(ENOMEM);
    }
    if (hipSetDevice(ctx->device) != hipSuccess)
        return CRAW_ERROR(ENODEV);

    s = new CRAWServer();
    s->ctx = ctx;
    s->params = *params;
    s->nb_slots = params->nb_slots;
    s->t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 0; i < s->nb_slots; i++) {
        Slot &sl = s->slots[i];

        if ((ret = craw_batch_create(ctx, &bp, &sl.batch)) < 0)
            goto fail;
        ret = CRAW_ERROR(ENOMEM);
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess) {
            sl.stream = NULL;
            goto fail;
        }
        if (hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
            sl.done = NULL;
            goto fail;
        }
        s->mem_slots += fp;
    }
    s->thread = std::thread(scheduler, s);
    *out = s;
    return 0;

fail:
    craw_server_free(&s);
    return ret;
}

CRAW_API int craw_server_open_stream(CRAWServer *s, uint32_t width, uint32_t height,
                                     int32_t *id)
{
    Stream st;
    uint64_t bytes;
    int ret;

    if (!s || !id || !width || !height || (width & 1) || (height & 1) ||
        width > s->params.max_width || height > s->params.max_height)
        return CRAW_ERROR(EINVAL);
    memset(&st, 0, sizeof(st));
    st.width = width;
    st.height = height;
    bytes = (s->params.flags & CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE) ?
            10ull * width * height : 0;

    std::lock_guard<std::mutex> g(s->mu);
    if (s->mem_slots + s->mem_streams + bytes > s->params.hbm_budget) {
        craw_log(s->ctx, CRAW_LOG_ERROR, "A %ux%u stream's state does not fit the budget\n",
                 width, height);
        return CRAW_ERROR(ENOMEM);
    }
    if (bytes && (ret = craw_temporal_state_create(s->ctx, width, height, &st.ts)) < 0)
        return ret;
    s->mem_streams += bytes;
    *id = s->next_stream++;
    s->streams[*id] = st;
    return 0;
}

CRAW_API int craw_server_close_stream(CRAWServer *s, int32_t id)
{
    if (!s)
        return CRAW_ERROR(EINVAL);
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->streams.find(id);
    if (it == s->streams.end())
        return CRAW_ERROR(EINVAL);
    s->cv_done.wait(lk, [&] { return it->second.pending == 0 && !it->second.on_gpu; });
    if (it->second.ts) {
        s->mem_streams -= 10ull * it->second.width * it->second.height;
        craw_temporal_state_free(&it->second.ts);
    }
    s->streams.erase(it);
    return 0;
}

/* Checks what needs no lock. */
static int request_shape_ok(CRAWServer *s, const CRAWServeRequest *r)
{
    if (!r || !r->params.profile)
        return CRAW_ERROR(EINVAL);
    if (r->src.width != r->dst.width || r->src.height != r->dst.height ||
        r->src.width > s->params.max_width || r->src.height > s->params.max_height ||
        (uint64_t)r->src.width * r->src.height > s->params.max_batch_pixels)
        return CRAW_ERROR(EINVAL);
    if (r->stream < -1)
        return CRAW_ERROR(EINVAL);
    return 0;
}

CRAW_API int craw_server_submit_n(CRAWServer *s, const CRAWServeRequest *reqs, uint32_t nb)
{
    double t;
    int ret;

    if (!s || (nb && !reqs))
        return CRAW_ERROR(EINVAL);
    for (uint32_t i = 0; i < nb; i++)
        if ((ret = request_shape_ok(s, &reqs[i])) < 0)
            return ret;
    {
        std::lock_guard<std::mutex> g(s->mu);

        for (uint32_t i = 0; i < nb; i++) {
            if (reqs[i].stream >= 0) {
                auto it = s->streams.find(reqs[i].stream);
                if (it == s->streams.end() || it->second.width != reqs[i].src.width ||
                    it->second.height != reqs[i].src.height)
                    return CRAW_ERROR(EINVAL);
            }
        }
        if (s->queue.size() + nb > s->params.max_queue)
            return CRAW_ERROR(EAGAIN);
        t = s->now_ms();
        for (uint32_t i = 0; i < nb; i++) {
            Req q;

            q.r = reqs[i];
            q.submit_ms = t;
            if (q.r.stream >= 0)
                s->streams[q.r.stream].pending++;
            s->queue.push_back(q);
            s->submitted++;
        }
    }
    s->cv_work.notify_one();
    return 0;
}

CRAW_API int craw_server_submit(CRAWServer *s, const CRAWServeRequest *r)
{
    return craw_server_submit_n(s, r, 1);
}

CRAW_API int craw_server_poll(CRAWServer *s, CRAWServeCompletion *out, uint32_t max,
                              int32_t timeout_ms)
{
    uint32_t n = 0;

    if (!s || (max && !out))
        return CRAW_ERROR(EINVAL);
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && timeout_ms != 0) {
        auto ready = [&] { return !s->done.empty(); };
        if (timeout_ms < 0)
            s->cv_done.wait(lk, [&] { return ready() || s->completed == s->submitted; });
        else
            s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    }
    while (n < max && !s->done.empty()) {
        out[n++] = s->done.front();
        s->done.pop_front();
    }
    return (int)n;
}

CRAW_API int craw_server_drain(CRAWServer *s)
{
    if (!s)
        return CRAW_ERROR(EINVAL);
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->completed == s->submitted; });
    return 0;
}

CRAW_API uint64_t craw_server_get_memory_usage(CRAWServer *s)
{
    if (!s)
        return 0;
    std::lock_guard<std::mutex> g(s->mu);
    return s->mem_slots + s->mem_streams;
}

} /* extern "C" */
