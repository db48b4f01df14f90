/*
 * libcraw for MI355X: the batched form (include/craw/batch.h).
 *
 * A recording plans every frame exactly as craw_pipeline_record() plans
 * one (craw_frame_plan()), with the frame's intermediates carved out of
 * per-kind arenas, then builds one work-list of tiles per stage variant
 * (denoise mode; debayer to the chroma stage or to the output; chroma
 * blur; chroma up) across all frames and launches each list once as a
 * persistent kernel. Per-frame constants and the work-lists travel to the
 * device through a small ring of pinned staging buffers, so recordings can
 * be queued back to back. With graphs on, the launches of each distinct
 * recording shape are captured once into a hipGraph and replayed.
 */
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "internal.h"

#define RING 4
#define NB_LISTS 12     /* 5 denoise modes, 2 debayer targets, blur, up, awb, 2 RCD streams */
#define L_DN0 0
#define L_DB_OUT 5
#define L_DB_INTER 6
#define L_CB 7
#define L_CU 8
#define L_AWB 9
#define L_RS 10
#define L_RS_F 11

static const int dn_modes[5] = {
    KDN_SPATIAL, KDN_TEMPORAL, KDN_TEMPORAL | KDN_HISTORY, KDN_SPATIAL | KDN_TEMPORAL,
    KDN_SPATIAL | KDN_TEMPORAL | KDN_HISTORY,
};

struct Staging {
    void *host;             /* pinned: frames table, then the tile lists */
    hipEvent_t done;        /* the H2D copies out of it have executed */
    bool used;
};

/* A captured launch whose arguments are a frame's constants. */
enum FrameNodeKind {
    FN_UP,              /* (KFrame): chroma up + colour */
    FN_HL,              /* (KHighlight): highlight field, coarse levels, fill */
    FN_HL_PUSH,         /* (KHighlight, int level) */
    FN_FC,              /* (KFrame, KInter): false colour */
};

struct FrameNode {
    hipGraphNode_t node;
    hipKernelNodeParams params;
    uint32_t frame;
    FrameNodeKind kind;
    int level;
};

struct CRAWBatch {
    CRAWContext *ctx;
    CRAWBatchParams params;
    CRAWStageConfig cfg;
    int form;

    /* arenas */
    uint16_t *clean;
    float *inter_rgb, *inter_rgb2, *down, *blur;
    uint16_t *hl_mosaic;
    float4 *hl_levels;
    size_t hl_levels_cap;   /* texels */
    uint32_t *hist;
    uint2 *awb_samples;
    uint32_t *hist_host;    /* pinned, max_frames histograms */
    uint32_t nb_last;
    uint64_t device_bytes;

    /* device tables */
    KFrame *frames;
    uint32_t *tiles;
    size_t tiles_cap;
    Staging ring[RING];
    uint32_t ring_pos;

    /* persistent workgroups per CU for each list kind (0 denoise, 1
     * debayer, 2 chroma blur, 3 AWB histogram); 0: one workgroup per tile */
    int wg_per_cu[4];
    /* the denoise launch shape (KDN_SHAPE_*): band height of the pairwise
     * kernel, search-row unroll of the fast one */
    int dn_shape;

    /* graph of the last captured shape */
    std::vector<uint32_t> graph_key;
    hipGraph_t graph;
    hipGraphExec_t exec;
    /* the graph's launches that take a frame's constants as arguments
     * (chroma up + colour, the highlight stage, false colour): each node,
     * its captured launch shape, the frame it develops and what its
     * arguments are; they are set again before every replay */
    std::vector<FrameNode> fnode;

    CRAWProfiler prof;
};

/* Most tiles `stage` can have over the frames of one recording: a frame of
 * w x h grid samples has at most (w/tw + 1)(h/th + 1) tiles per plane, and
 * w <= W/div + 1, h <= H/div + 1; summed over at most max_frames frames of
 * at most max_pixels pixels together. */
/* Whether the frame's plan launches the chroma-up + colour kernel. */
static bool has_up(const CRAWFramePlan &pl)
{
    return pl.chroma || pl.split_color;
}

static size_t stage_tiles_bound(int stage, const CRAWBatchParams *p, int method, int dn_geom)
{
    int tw, th, div, planes;
    double F = p->max_frames, wm, hm, area, n;

    if (stage == KSTAGE_RCD_RING)       /* at most every tile of the frame */
        return stage_tiles_bound(KSTAGE_DEBAYER, p, 2, dn_geom);
    if (stage == KSTAGE_RCD_STREAM || stage == KSTAGE_RCD_STREAM_F) {
        /* strips of the interior, in groups */
        const double outw = stage == KSTAGE_RCD_STREAM ? KG_RS_OUT : KG_RS_OUT_F;
        n = (double)p->max_pixels / (outw * KG_RS_R * KG_RS_NW) +
            F * (((double)p->max_width / outw + (double)p->max_height / KG_RS_R + 1.0) /
                 KG_RS_NW + 1.0);
        return (size_t)n + 64;
    }

    craw_stage_geom(stage, method, dn_geom, &tw, &th, &div, &planes);
    wm = (double)p->max_width / div + 1.0;
    hm = (double)p->max_height / div + 1.0;
    area = (double)p->max_pixels / ((double)div * div) + F * (wm + hm);
    n = planes * (area / ((double)tw * th) + F * (wm / tw + hm / th + 1.0));
    if (stage == KSTAGE_DENOISE)
        n += 15.0 * F;          /* each frame's entries rounded to 16 */
    return (size_t)n + 64;
}

extern "C" {

CRAW_API void craw_batch_params_init(CRAWBatchParams *p)
{
    if (!p)
        return;
    memset(p, 0, sizeof(*p));
    p->max_frames = 16;
    p->debayer = CRAW_DEBAYER_RCD;
    p->out_space = CRAW_CS_XYZ_D65;
    p->out_format = CRAW_FORMAT_R32G32B32A32_SFLOAT;
}

CRAW_API void craw_batch_free(CRAWBatch **pb)
{
    CRAWBatch *b;

    if (!pb || !*pb)
        return;
    b = *pb;
    CRAW_IGNORE(hipSetDevice(b->ctx->device));
    if (b->exec)
        CRAW_IGNORE(hipGraphExecDestroy(b->exec));
    if (b->graph)
        CRAW_IGNORE(hipGraphDestroy(b->graph));
    for (int i = 0; i < RING; i++) {
        if (b->ring[i].done)
            CRAW_IGNORE(hipEventDestroy(b->ring[i].done));
        if (b->ring[i].host)
            CRAW_IGNORE(hipHostFree(b->ring[i].host));
    }
    CRAW_IGNORE(hipFree(b->clean));
    CRAW_IGNORE(hipFree(b->inter_rgb));
    CRAW_IGNORE(hipFree(b->inter_rgb2));
    CRAW_IGNORE(hipFree(b->hl_mosaic));
    CRAW_IGNORE(hipFree(b->hl_levels));
    CRAW_IGNORE(hipFree(b->down));
    CRAW_IGNORE(hipFree(b->blur));
    CRAW_IGNORE(hipFree(b->hist));
    CRAW_IGNORE(hipFree(b->awb_samples));
    CRAW_IGNORE(hipFree(b->frames));
    CRAW_IGNORE(hipFree(b->tiles));
    if (b->hist_host)
        CRAW_IGNORE(hipHostFree(b->hist_host));
    craw_profiler_free(&b->prof);
    delete b;
    *pb = NULL;
}

} /* extern "C" */

/* Bytes of each device buffer a batch of `params` allocates. */
struct BatchSizes {
    size_t clean, inter_rgb, inter_rgb2, down, blur, hl_mosaic, hl_levels, hist, awb_samples,
        frames, tiles_cap;

    size_t device() const
    {
        return clean + inter_rgb + inter_rgb2 + down + blur + hl_mosaic + hl_levels + hist +
               awb_samples + frames + tiles_cap * sizeof(uint32_t);
    }
};

static int batch_sizes(CRAWContext *ctx, const CRAWBatchParams *params, CRAWStageConfig *cfg,
                       BatchSizes *sz)
{
    size_t px, qpx;
    bool fast;
    int ret;

    if (!params->max_frames || params->max_frames > 255 || !params->max_pixels ||
        params->max_width < 8 || params->max_height < 8 ||
        (uint64_t)params->max_width * params->max_height > params->max_pixels)
        return CRAW_ERROR(EINVAL);
    ret = craw_stage_config(ctx, params->flags, params->debayer, params->out_space,
                            params->out_format, cfg);
    if (ret < 0)
        return ret;
    fast = cfg->fp16 || cfg->fast;
    px = params->max_pixels;
    /* quarter grid of every frame, with a row and a column of rounding */
    qpx = px / 16 + (size_t)params->max_frames * ((params->max_width + params->max_height) / 4 + 2);
    memset(sz, 0, sizeof(*sz));
    if (cfg->has_denoise || cfg->has_temporal)
        sz->clean = px * sizeof(uint16_t);
    if (cfg->has_chroma) {
        sz->down = 3 * qpx * sizeof(float);
        sz->blur = 2 * qpx * sizeof(float);
    }
    if (craw_needs_inter_rgb(*cfg))
        sz->inter_rgb = 3 * px * sizeof(float);
    if (cfg->has_chroma && cfg->has_fc)
        sz->inter_rgb2 = 3 * px * sizeof(float);
    if (cfg->has_hl) {
        /* a frame's field levels hold at most w h / 12 + (w + h) / 2 + 16
         * texels (each level a quarter of the one below, plus a row and a
         * column of rounding) */
        sz->hl_mosaic = px * sizeof(uint16_t);
        sz->hl_levels = (px / 12 + (size_t)params->max_frames *
                         ((params->max_width + params->max_height) / 2 + 64)) * sizeof(float4);
    }
    if (cfg->awb) {
        sz->hist = (size_t)params->max_frames * CRAW_HIST_WORDS * 4;
        /* 8x8 blocks of every frame, a row and a column of rounding each */
        sz->awb_samples = (px / 64 + (size_t)params->max_frames *
                           ((params->max_width + params->max_height) / 8 + 2)) * sizeof(uint2);
    }
    sz->frames = params->max_frames * sizeof(KFrame);
    for (int st = 0; st < KSTAGE_NB; st++)
        sz->tiles_cap += stage_tiles_bound(st, params, (int)params->debayer,
                                           fast ? KDN_GEOM_FAST : KDN_GEOM_DIRECT);
    /* frames with the pairwise-distance tiles (the exact form's spatial
     * modes, with and without history, and the precision form's spatial
     * ones) */
    sz->tiles_cap += stage_tiles_bound(KSTAGE_DENOISE, params, (int)params->debayer,
                                       KDN_GEOM_SYM);
    sz->tiles_cap += stage_tiles_bound(KSTAGE_DENOISE, params, (int)params->debayer,
                                       KDN_GEOM_SYMH);
    sz->tiles_cap += stage_tiles_bound(KSTAGE_DENOISE, params, (int)params->debayer,
                                       KDN_GEOM_SYMH16);
    sz->tiles_cap += NB_LISTS * 64;
    return 0;
}

int craw_batch_footprint(CRAWContext *ctx, const CRAWBatchParams *params, uint64_t *bytes)
{
    CRAWStageConfig cfg;
    BatchSizes sz;
    int ret = batch_sizes(ctx, params, &cfg, &sz);

    if (ret < 0)
        return ret;
    *bytes = sz.device();
    return 0;
}

extern "C" {

CRAW_API int craw_batch_create(CRAWContext *ctx, const CRAWBatchParams *params, CRAWBatch **out)
{
    CRAWBatch *b;
    BatchSizes sz;
    size_t staging;
    int ret;

    if (!ctx || !params || !out)
        return CRAW_ERROR(EINVAL);
    b = new CRAWBatch();
    b->ctx = ctx;
    b->params = *params;
    ret = batch_sizes(ctx, params, &b->cfg, &sz);
    if (ret < 0) {
        delete b;
        return ret;
    }
    b->form = b->cfg.fp16 ? KFORM_FP16 : (b->cfg.fast ? KFORM_FAST : KFORM_EXACT);
    for (int i = 0; i < 4; i++)
        b->wg_per_cu[i] = 0;
    b->dn_shape = 0;
    if (hipSetDevice(ctx->device) != hipSuccess) {
        delete b;
        return CRAW_ERROR(ENODEV);
    }

    ret = CRAW_ERROR(ENOMEM);
#define DALLOC(ptr, bytes) ((bytes) && hipMalloc((void **)&(ptr), (bytes)) != hipSuccess)
    if (DALLOC(b->clean, sz.clean) || DALLOC(b->inter_rgb, sz.inter_rgb) ||
        DALLOC(b->inter_rgb2, sz.inter_rgb2) || DALLOC(b->hl_mosaic, sz.hl_mosaic) ||
        DALLOC(b->hl_levels, sz.hl_levels) ||
        DALLOC(b->down, sz.down) || DALLOC(b->blur, sz.blur) || DALLOC(b->hist, sz.hist) ||
        DALLOC(b->awb_samples, sz.awb_samples) ||
        DALLOC(b->frames, sz.frames) || DALLOC(b->tiles, sz.tiles_cap * sizeof(uint32_t)))
        goto fail;
#undef DALLOC
    if (sz.hist && hipHostMalloc((void **)&b->hist_host, sz.hist, hipHostMallocDefault) !=
                       hipSuccess) {
        b->hist_host = NULL;
        goto fail;
    }
    b->tiles_cap = sz.tiles_cap;
    b->hl_levels_cap = sz.hl_levels / sizeof(float4);
    b->device_bytes = sz.device();
    staging = sz.frames + b->tiles_cap * sizeof(uint32_t);
    for (int i = 0; i < RING; i++) {
        if (hipHostMalloc(&b->ring[i].host, staging, hipHostMallocDefault) != hipSuccess) {
            b->ring[i].host = NULL;
            goto fail;
        }
        if (hipEventCreateWithFlags(&b->ring[i].done, hipEventDisableTiming) != hipSuccess) {
            b->ring[i].done = NULL;
            goto fail;
        }
    }
    if ((params->flags & CRAW_PIPELINE_FLAG_PROFILE) && craw_profiler_init(&b->prof) < 0)
        goto fail;

    *out = b;
    return 0;

fail:
    craw_batch_free(&b);
    return ret;
}

} /* extern "C" */

/* The recording of one batch, planned on the host. */
struct BatchPlan {
    std::vector<CRAWFramePlan> plans;
    std::vector<uint32_t> lists[NB_LISTS];
    std::vector<uint32_t> key;
};

static int batch_plan(CRAWBatch *b, const CRAWBatchFrame *frames, uint32_t nb, BatchPlan &bp)
{
    size_t off_px = 0, off_q = 0, off_awb = 0, off_lv = 0;

    if (!nb || nb > b->params.max_frames)
        return CRAW_ERROR(ENOSPC);
    bp.plans.resize(nb);
    for (uint32_t i = 0; i < nb; i++) {
        const CRAWBatchFrame &fr = frames[i];
        uint32_t w = fr.src.width, h = fr.src.height;
        size_t q = (size_t)((w + 3) / 4) * ((h + 3) / 4);
        CRAWFrameBuffers buf = {};
        int ret;

        if (!fr.params || !fr.params->profile)
            return CRAW_ERROR(EINVAL);
        if (w > b->params.max_width || h > b->params.max_height ||
            off_px + (size_t)w * h > b->params.max_pixels)
            return CRAW_ERROR(ENOSPC);
        buf.clean = b->clean ? b->clean + off_px : NULL;
        buf.inter_rgb = b->inter_rgb ? b->inter_rgb + 3 * off_px : NULL;
        buf.inter_rgb2 = b->inter_rgb2 ? b->inter_rgb2 + 3 * off_px : NULL;
        buf.hl_mosaic = b->hl_mosaic ? b->hl_mosaic + off_px : NULL;
        buf.hl_levels = b->hl_levels ? b->hl_levels + off_lv : NULL;
        buf.down = b->down ? b->down + 3 * off_q : NULL;
        buf.blur = b->blur ? b->blur + 2 * off_q : NULL;
        buf.hist = b->hist ? b->hist + (size_t)i * CRAW_HIST_WORDS : NULL;
        buf.awb_samples = b->awb_samples ? b->awb_samples + off_awb : NULL;
        buf.ts = fr.state;
        ret = craw_frame_plan(b->ctx, b->cfg, buf, w, h, &fr.src, &fr.dst, fr.params,
                              &bp.plans[i]);
        if (ret < 0)
            return ret;
        off_px += (size_t)w * h;
        off_q += q;
        if (b->hl_levels) {
            int lw[KHL_LEVELS], lh[KHL_LEVELS];
            size_t off[KHL_LEVELS], texels;

            if (craw_hl_levels(w, h, lw, lh, off, &texels) < 0)
                return CRAW_ERROR(EINVAL);
            /* the arena's bound holds for any frames within the batch's
             * limits; checked all the same before anything is written */
            if (off_lv + texels > b->hl_levels_cap)
                return CRAW_ERROR(ENOSPC);
            off_lv += texels;
        }
        off_awb += craw_awb_blocks(w, h);
    }
    /* one temporal state may appear once per recording */
    for (uint32_t i = 0; i < nb; i++)
        for (uint32_t j = i + 1; j < nb; j++)
            if (frames[i].state && frames[i].state == frames[j].state)
                return CRAW_ERROR(EINVAL);

    for (uint32_t i = 0; i < nb; i++) {
        const CRAWFramePlan &pl = bp.plans[i];
        /* the debayer writes camera RGB planes for a later stage */
        const bool inter = pl.chroma || pl.fc_passes || pl.split_color;
        int W = pl.f.W, H = pl.f.H, ntx, n;

        if (pl.dn_mode) {
            int li = -1;
            for (int m = 0; m < 5; m++)
                if (dn_modes[m] == pl.dn_mode)
                    li = L_DN0 + m;
            n = craw_stage_tiles(KSTAGE_DENOISE, W, H, pl.f.method,
                                 craw_dn_geom(b->form, pl.dn_mode, b->dn_shape), &ntx);
            for (int t = 0; t < n; t++)
                bp.lists[li].push_back(KTILE(i, t));
        }
        if (pl.f.method == 2) {
            /* RCD: ring tiles and interior strips */
            n = craw_stage_tiles(KSTAGE_RCD_RING, W, H, 2, 0, &ntx);
            for (int t = 0; t < n; t++)
                bp.lists[inter ? L_DB_INTER : L_DB_OUT].push_back(KTILE(i, t));
            n = craw_stage_tiles(inter ? KSTAGE_RCD_STREAM : KSTAGE_RCD_STREAM_F, W, H, 2, 0,
                                 &ntx);
            for (int t = 0; t < n; t++)
                bp.lists[inter ? L_RS : L_RS_F].push_back(KTILE(i, t));
        } else {
            n = craw_stage_tiles(KSTAGE_DEBAYER, W, H, pl.f.method, 0, &ntx);
            for (int t = 0; t < n; t++)
                bp.lists[inter ? L_DB_INTER : L_DB_OUT].push_back(KTILE(i, t));
        }
        if (pl.chroma) {
            n = craw_stage_tiles(KSTAGE_CHROMA_BLUR, W, H, pl.f.method, 0, &ntx);
            for (int t = 0; t < n; t++)
                bp.lists[L_CB].push_back(KTILE(i, t));
        }
        if (has_up(pl)) {
            n = craw_stage_tiles(KSTAGE_CHROMA_UP, W, H, pl.f.method, 0, &ntx);
            for (int t = 0; t < n; t++)
                bp.lists[L_CU].push_back(KTILE(i, t));
        }
        if (b->cfg.awb) {
            n = craw_stage_tiles(KSTAGE_AWB, W, H, pl.f.method, 0, &ntx);
            for (int t = 0; t < n; t++)
                bp.lists[L_AWB].push_back(KTILE(i, t));
        }
        bp.key.push_back((uint32_t)W);
        bp.key.push_back((uint32_t)H);
        bp.key.push_back((uint32_t)pl.dn_mode | (pl.chroma ? 0x100u : 0u) |
                         (pl.hl ? 0x200u : 0u) | (pl.split_color ? 0x400u : 0u) |
                         ((uint32_t)pl.fc_passes << 12));
    }
    {
        size_t total = 0;
        for (int l = 0; l < NB_LISTS; l++)
            total += bp.lists[l].size();
        if (total > b->tiles_cap)
            return CRAW_ERROR(ENOSPC);
    }
    for (int k = 0; k < 4; k++)
        bp.key.push_back((uint32_t)b->wg_per_cu[k]);
    bp.key.push_back((uint32_t)b->dn_shape);
    return 0;
}

int craw_batch_check(CRAWBatch *b, const CRAWBatchFrame *frames, uint32_t nb)
{
    BatchPlan bp;

    if (!b || !frames)
        return CRAW_ERROR(EINVAL);
    return batch_plan(b, frames, nb, bp);
}

/* Workgroups of a persistent launch over `ntiles` entries: wg_per_cu per
 * CU, or with 0 one per entry (the hardware's own scheduling). */
static int grid_for(CRAWBatch *b, int kind, size_t ntiles)
{
    size_t g = (size_t)b->wg_per_cu[kind] * (size_t)b->ctx->cu_count;

    if (!b->wg_per_cu[kind] || ntiles < g)
        return (int)ntiles;
    return (int)g;
}

/* The stage launches of a planned recording, in order: with `part` 1 those
 * that read only the recording's tables (uploaded again before each
 * replay), with 2 the rest (the chroma-up launches, which take a frame's
 * constants as arguments, and the histogram after them), with 0 all (what
 * a graph holds). */
static hipError_t batch_launch(CRAWBatch *b, const BatchPlan &bp, const size_t *list_off,
                               hipStream_t s, CRAWProfRun *pr, int part = 0)
{
    const uint32_t nb = (uint32_t)bp.plans.size();
    hipError_t e = hipSuccess;

#define PB(name) do { if (pr) craw_profiler_begin(*pr, name); } while (0)
#define PE() do { if (pr) craw_profiler_end(*pr); } while (0)
    if (part == 2)
        goto rest;
    if (b->cfg.awb) {
        PB("awb_clear");
        e = craw_launch_zero_u32(b->hist, (size_t)nb * CRAW_HIST_WORDS, s);
        PE();
    }
    for (int m = 0; m < 5 && e == hipSuccess; m++) {
        size_t n = bp.lists[L_DN0 + m].size();
        if (!n)
            continue;
        PB(b->form == KFORM_FP16 ? "denoise_nlm_fp16" : (b->form == KFORM_FAST ?
                                                            "denoise_nlm_fast" : "denoise_nlm"));
        e = craw_launch_batch_denoise(dn_modes[m], b->form, b->dn_shape, b->frames,
                                      b->tiles + list_off[L_DN0 + m], (uint32_t)n,
                                      grid_for(b, 0, n), s);
        PE();
    }
    /* highlight reconstruction and false colour: plain launches a frame
     * (in a graph, their arguments set again before each replay), each
     * step over all frames one stage */
    {
        bool any = false;

        for (uint32_t i = 0; i < nb; i++)
            any = any || bp.plans[i].hl;
        for (int step = 0; step < 3 && any && e == hipSuccess; step++) {
            static const char *names[3] = { "highlight_field", "highlight_push",
                                            "highlight_fill" };

            PB(names[step]);
            for (uint32_t i = 0; i < nb && e == hipSuccess; i++)
                if (bp.plans[i].hl)
                    e = step == 0 ? craw_launch_hl_field(bp.plans[i].hl_k, s) :
                        step == 1 ? craw_launch_hl_push(bp.plans[i].hl_k, s) :
                                    craw_launch_hl_fill(bp.plans[i].hl_k, s);
            PE();
        }
    }
    for (int l = L_DB_OUT; l <= L_DB_INTER && e == hipSuccess; l++) {
        size_t n = bp.lists[l].size();
        static const char *names[3] = { "debayer_bilinear", "debayer_mhc", "rcd" };
        if (!n)
            continue;
        PB(names[b->cfg.debayer]);
        e = craw_launch_batch_debayer((int)b->cfg.debayer, l == L_DB_INTER, b->frames,
                                      b->tiles + list_off[l], (uint32_t)n, grid_for(b, 1, n), s);
        PE();
    }
    for (int l = L_RS; l <= L_RS_F && e == hipSuccess; l++) {
        size_t n = bp.lists[l].size();
        bool gamut = false;

        if (!n)
            continue;
        /* the fused colour tail with the compression, if a frame there has it */
        for (uint32_t i = 0; i < nb && l == L_RS_F; i++) {
            const CRAWFramePlan &pl = bp.plans[i];
            gamut = gamut || (pl.f.method == 2 && !(pl.chroma || pl.fc_passes || pl.split_color) &&
                              (pl.f.col.flags & KCOLOR_GAMUT_COMPRESS));
        }
        PB("rcd");
        e = craw_launch_batch_rcd_stream(l == L_RS, gamut, b->frames, b->tiles + list_off[l],
                                         (uint32_t)n, grid_for(b, 1, n), s);
        PE();
    }
    {
        bool any = false;

        for (uint32_t i = 0; i < nb; i++)
            any = any || bp.plans[i].fc_passes;
        if (any && e == hipSuccess) {
            PB("false_color");
            for (uint32_t i = 0; i < nb && e == hipSuccess; i++) {
                const CRAWFramePlan &pl = bp.plans[i];

                if (pl.fc_passes)
                    e = craw_launch_false_color(pl.fc_passes, pl.chroma, pl.f, pl.fc_to, s);
            }
            PE();
        }
    }
    if (e == hipSuccess && bp.lists[L_CB].size()) {
        size_t n = bp.lists[L_CB].size();
        PB("chroma_blur");
        e = craw_launch_batch_chroma_blur(b->frames, b->tiles + list_off[L_CB], (uint32_t)n,
                                          grid_for(b, 2, n), s);
        PE();
    }
    if (part == 1)
        return e;
rest:
    if (e == hipSuccess && bp.lists[L_CU].size()) {
        /* one plain launch a frame, its constants in the kernel's
         * arguments: the persistent form of this stage, reading them from
         * memory, ran at 0.26 ms a 5.8K frame against 0.16. In a graph
         * these launches' arguments are set again before each replay (the
         * frames' constants change); the frames whose RCD colour stage
         * runs from its planes take the same kernel */
        PB("chroma_up_color");
        for (uint32_t i = 0; i < nb && e == hipSuccess; i++)
            if (has_up(bp.plans[i]))
                e = craw_launch_chroma_up_finish(bp.plans[i].f, s);
        PE();
    }
    if (e == hipSuccess && bp.lists[L_AWB].size()) {
        size_t n = bp.lists[L_AWB].size();
        PB("awb_hist");
        e = craw_launch_batch_awb_hist(b->frames, b->tiles + list_off[L_AWB], (uint32_t)n,
                                       grid_for(b, 3, n), s);
        PE();
    }
#undef PB
#undef PE
    return e;
}

/* The launches of a captured recording that take a frame's constants as
 * arguments: every kernel node of those kernels, matched to the frame
 * whose constants it took (its arguments copied at capture), and checked
 * against the number of such launches the recording makes, so that no
 * node keeps a previous recording's constants. */
static hipError_t graph_frame_nodes(hipGraph_t g, const BatchPlan &bp,
                                    std::vector<FrameNode> &out)
{
    const void *cu = craw_chroma_up_kernel();
    const void *hl[4] = { craw_hl_kernel(0), craw_hl_kernel(1), craw_hl_kernel(2),
                          craw_hl_kernel(3) };
    std::vector<hipGraphNode_t> nodes;
    size_t n = 0, want = 0;
    hipError_t e;

    for (const CRAWFramePlan &pl : bp.plans) {
        want += has_up(pl) ? 1 : 0;
        if (pl.hl)
            want += 2 + (size_t)craw_hl_push_launches(pl.hl_k);
        want += pl.fc_passes ? 1 : 0;
    }
    if ((e = hipGraphGetNodes(g, NULL, &n)) != hipSuccess)
        return e;
    nodes.resize(n);
    if (n && (e = hipGraphGetNodes(g, nodes.data(), &n)) != hipSuccess)
        return e;
    for (hipGraphNode_t nd : nodes) {
        hipGraphNodeType t;
        hipKernelNodeParams kp;
        FrameNode fnd;
        size_t i;

        if ((e = hipGraphNodeGetType(nd, &t)) != hipSuccess)
            return e;
        if (t != hipGraphNodeTypeKernel)
            continue;
        if ((e = hipGraphKernelNodeGetParams(nd, &kp)) != hipSuccess)
            return e;
        if (kp.func == cu)
            fnd.kind = FN_UP;
        else if (kp.func == hl[0] || kp.func == hl[1] || kp.func == hl[3])
            fnd.kind = FN_HL;
        else if (kp.func == hl[2])
            fnd.kind = FN_HL_PUSH;
        else if (craw_is_false_color_kernel(kp.func))
            fnd.kind = FN_FC;
        else
            continue;
        if (!kp.kernelParams || !kp.kernelParams[0] ||
            ((fnd.kind == FN_HL_PUSH || fnd.kind == FN_FC) && !kp.kernelParams[1]))
            return hipErrorInvalidValue;
        fnd.level = fnd.kind == FN_HL_PUSH ? *(const int *)kp.kernelParams[1] : 0;
        for (i = 0; i < bp.plans.size(); i++) {
            const CRAWFramePlan &pl = bp.plans[i];
            const void *a = kp.kernelParams[0];

            if (fnd.kind == FN_UP && has_up(pl) && !memcmp(a, &pl.f, sizeof(KFrame)))
                break;
            if ((fnd.kind == FN_HL || fnd.kind == FN_HL_PUSH) && pl.hl &&
                !memcmp(a, &pl.hl_k, sizeof(KHighlight)))
                break;
            if (fnd.kind == FN_FC && pl.fc_passes && !memcmp(a, &pl.f, sizeof(KFrame)) &&
                !memcmp(kp.kernelParams[1], &pl.fc_to, sizeof(KInter)))
                break;
        }
        if (i == bp.plans.size())
            return hipErrorInvalidValue;
        /* one node a frame for each kernel (a level for the push) */
        for (const FrameNode &o : out)
            if (o.frame == i && o.params.func == kp.func && o.level == fnd.level)
                return hipErrorInvalidValue;
        fnd.node = nd;
        fnd.params = kp;
        fnd.frame = (uint32_t)i;
        out.push_back(fnd);
    }
    return out.size() == want ? hipSuccess : hipErrorInvalidValue;
}

static int batch_record(CRAWBatch *b, hipStream_t s, const CRAWBatchFrame *frames, uint32_t nb,
                        bool graphs)
{
    BatchPlan bp;
    Staging *stg;
    size_t list_off[NB_LISTS], pos = 0, frames_bytes;
    hipError_t e;
    int ret;

    if (!b || !frames)
        return CRAW_ERROR(EINVAL);
    ret = batch_plan(b, frames, nb, bp);
    if (ret < 0)
        return ret;
    if (hipSetDevice(b->ctx->device) != hipSuccess)
        return CRAW_ERROR(ENODEV);

    /* staging: wait until the slot's previous copies have executed */
    stg = &b->ring[b->ring_pos];
    if (stg->used && hipEventSynchronize(stg->done) != hipSuccess)
        return CRAW_ERROR(EIO);
    frames_bytes = nb * sizeof(KFrame);
    for (uint32_t i = 0; i < nb; i++)
        memcpy((char *)stg->host + i * sizeof(KFrame), &bp.plans[i].f, sizeof(KFrame));
    for (int l = 0; l < NB_LISTS; l++) {
        list_off[l] = pos;
        if (bp.lists[l].size())
            memcpy((char *)stg->host + frames_bytes + pos * 4, bp.lists[l].data(),
                   bp.lists[l].size() * 4);
        pos += bp.lists[l].size();
    }
    e = hipMemcpyAsync(b->frames, stg->host, frames_bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && pos)
        e = hipMemcpyAsync(b->tiles, (char *)stg->host + frames_bytes, pos * 4,
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipEventRecord(stg->done, s);
    if (e != hipSuccess)
        return CRAW_ERROR(EIO);
    stg->used = true;
    b->ring_pos = (b->ring_pos + 1) % RING;

    if (graphs) {
        if (!b->exec || bp.key != b->graph_key) {
            hipGraph_t g = NULL;
            hipGraphExec_t x = NULL;
            std::vector<FrameNode> fnode;

            e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
            if (e == hipSuccess) {
                hipError_t le = batch_launch(b, bp, list_off, s, NULL, 0);
                e = hipStreamEndCapture(s, &g);
                if (le != hipSuccess)
                    e = le;
            }
            if (e == hipSuccess)
                e = graph_frame_nodes(g, bp, fnode);
            if (e == hipSuccess)
                e = hipGraphInstantiate(&x, g, NULL, NULL, 0);
            if (e != hipSuccess) {
                if (g)
                    CRAW_IGNORE(hipGraphDestroy(g));
                craw_log(b->ctx, CRAW_LOG_ERROR, "Graph capture failed: %s\n",
                         hipGetErrorString(e));
                return CRAW_ERROR(EIO);
            }
            if (b->exec)
                CRAW_IGNORE(hipGraphExecDestroy(b->exec));
            if (b->graph)
                CRAW_IGNORE(hipGraphDestroy(b->graph));
            b->graph = g;
            b->exec = x;
            b->graph_key = bp.key;
            b->fnode.swap(fnode);
        }
        /* this recording's constants into the launches that take them */
        for (size_t k = 0; k < b->fnode.size() && e == hipSuccess; k++) {
            const FrameNode &fnd = b->fnode[k];
            const CRAWFramePlan &pl = bp.plans[fnd.frame];
            hipKernelNodeParams kp = fnd.params;
            KFrame f = pl.f;
            KHighlight h = pl.hl_k;
            KInter to = pl.fc_to;
            int level = fnd.level;
            void *args[2];

            switch (fnd.kind) {
            case FN_UP: args[0] = &f; break;
            case FN_HL: args[0] = &h; break;
            case FN_HL_PUSH: args[0] = &h; args[1] = &level; break;
            case FN_FC: args[0] = &f; args[1] = &to; break;
            }
            kp.kernelParams = args;
            kp.extra = NULL;
            e = hipGraphExecKernelNodeSetParams(b->exec, fnd.node, &kp);
        }
        if (e != hipSuccess) {
            craw_log(b->ctx, CRAW_LOG_ERROR, "Graph update failed: %s\n", hipGetErrorString(e));
            return CRAW_ERROR(EIO);
        }
        if (b->prof.on) {
            CRAWProfRun pr = craw_profiler_start(&b->prof, s);
            craw_profiler_begin(pr, "batch_graph", false);
            e = hipGraphLaunch(b->exec, s);
            craw_profiler_end(pr);
            craw_profiler_finish(pr);
        } else {
            e = hipGraphLaunch(b->exec, s);
        }
    } else if (b->prof.on) {
        CRAWProfRun pr = craw_profiler_start(&b->prof, s);
        e = batch_launch(b, bp, list_off, s, &pr);
        craw_profiler_finish(pr);
    } else {
        e = batch_launch(b, bp, list_off, s, NULL);
    }
    if (e == hipSuccess && b->cfg.awb)
        e = hipMemcpyAsync(b->hist_host, b->hist, (size_t)nb * CRAW_HIST_WORDS * 4,
                           hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) {
        craw_log(b->ctx, CRAW_LOG_ERROR, "Batch launch failed: %s\n", hipGetErrorString(e));
        return CRAW_ERROR(EIO);
    }
    b->nb_last = nb;
    for (uint32_t i = 0; i < nb; i++)
        craw_frame_commit(&bp.plans[i]);
    return 0;
}

extern "C" {

CRAW_API int craw_batch_record(CRAWBatch *b, CRAWStream stream, const CRAWBatchFrame *frames,
                               uint32_t nb)
{
    if (!b)
        return CRAW_ERROR(EINVAL);
    /* the legacy default stream cannot be captured: launch directly there */
    return batch_record(b, (hipStream_t)stream, frames, nb,
                        b->params.use_graphs != 0 && stream != NULL);
}

CRAW_API int craw_batch_get_awb_histogram(CRAWBatch *b, uint32_t index, CRAWAwbHistogram *out)
{
    const uint32_t *base;

    if (!b || !out)
        return CRAW_ERROR(EINVAL);
    if (!b->cfg.awb || index >= b->nb_last)
        return CRAW_ERROR(EINVAL);
    base = b->hist_host + (size_t)index * CRAW_HIST_WORDS;
    out->bins = CRAW_AWB_HIST_BINS;
    out->log_min = CRAW_AWB_HIST_LOG_MIN;
    out->log_max = CRAW_AWB_HIST_LOG_MAX;
    out->count = base;
    out->weight = base + CRAW_AWB_HIST_BINS * CRAW_AWB_HIST_BINS;
    out->weight_bright = base + 2 * CRAW_AWB_HIST_BINS * CRAW_AWB_HIST_BINS;
    return 0;
}

CRAW_API int craw_batch_get_stage_times(CRAWBatch *b, CRAWStageTime *out, uint32_t max)
{
    if (!b)
        return CRAW_ERROR(EINVAL);
    return craw_profiler_times(b->ctx, &b->prof, 0, out, max);
}

static_assert(CRAW_BATCH_DENOISE_TALL == KDN_SHAPE_TALL &&
              CRAW_BATCH_DENOISE_UNROLL == KDN_SHAPE_UNROLL, "the public options are the shapes");

CRAW_API int craw_batch_set_launch_shape(CRAWBatch *b, const int wg_per_cu[4],
                                         int denoise_options)
{
    if (!b || !wg_per_cu ||
        (denoise_options & ~(CRAW_BATCH_DENOISE_TALL | CRAW_BATCH_DENOISE_UNROLL)))
        return CRAW_ERROR(EINVAL);
    for (int k = 0; k < 4; k++)
        if (wg_per_cu[k] < 0 || wg_per_cu[k] > 64)
            return CRAW_ERROR(EINVAL);
    for (int k = 0; k < 4; k++)
        b->wg_per_cu[k] = wg_per_cu[k];
    b->dn_shape = denoise_options;
    return 0;
}

CRAW_API int craw_batch_get_launch_shape(CRAWBatch *b, int wg_per_cu[4], int *denoise_options)
{
    if (!b || !wg_per_cu || !denoise_options)
        return CRAW_ERROR(EINVAL);
    for (int k = 0; k < 4; k++)
        wg_per_cu[k] = b->wg_per_cu[k];
    *denoise_options = b->dn_shape;
    return 0;
}

/* Times every stage of `iterations` recordings for each candidate launch
 * shape and keeps, per stage kind, the fastest: one workgroup per tile or
 * 1 to 16 persistent workgroups per CU for every kind, and for the denoise
 * stage also its band height (pairwise kernel) and search-row unroll (fast
 * kernel), where the batch's form has that kernel. The temporal states of
 * the frames are left as they were. */
CRAW_API int craw_batch_autotune(CRAWBatch *b, CRAWStream stream, const CRAWBatchFrame *frames,
                                 uint32_t nb, uint32_t iterations)
{
    static const int cand[] = { 0, 1, 2, 4, 8, 16 };
    const int ncand = (int)(sizeof(cand) / sizeof(cand[0]));
    double best[4] = { 1e30, 1e30, 1e30, 1e30 }, total = 0.0;
    int pick[4] = { 0, 0, 0, 0 }, pick_shape = 0;
    hipStream_t s = (hipStream_t)stream;
    CRAWProfiler saved;
    std::vector<int> prev(nb), cur(nb);
    std::vector<int> shapes;
    int ret = 0;

    if (!b || !frames || !nb)
        return CRAW_ERROR(EINVAL);
    if (!iterations)
        iterations = 3;
    /* the shapes that change a launch of this form */
    shapes.push_back(0);
    if (b->form != KFORM_FAST)
        shapes.push_back(KDN_SHAPE_TALL);
    if (b->form != KFORM_EXACT)
        shapes.push_back(KDN_SHAPE_UNROLL);
    if (b->form == KFORM_FP16)
        shapes.push_back(KDN_SHAPE_TALL | KDN_SHAPE_UNROLL);
    for (uint32_t i = 0; i < nb; i++) {
        prev[i] = frames[i].state ? frames[i].state->prev : -1;
        cur[i] = frames[i].state ? frames[i].state->cur : -1;
    }
    saved = b->prof;
    memset(&b->prof, 0, sizeof(b->prof));
    if ((ret = craw_profiler_init(&b->prof)) < 0) {
        b->prof = saved;
        return ret;
    }
    for (size_t sh = 0; sh < shapes.size() && ret >= 0; sh++)
    for (int c = 0; c < ncand && ret >= 0; c++) {
        double acc[4] = { 0, 0, 0, 0 };

        for (int k = 0; k < 4; k++)
            b->wg_per_cu[k] = cand[c];
        b->dn_shape = shapes[sh];
        for (uint32_t it = 0; it <= iterations && ret >= 0; it++) {
            CRAWStageTime st[CRAW_MAX_STAGES];
            int n;

            for (uint32_t i = 0; i < nb; i++)
                if (frames[i].state) {
                    frames[i].state->prev = prev[i];
                    frames[i].state->cur = cur[i];
                }
            ret = batch_record(b, s, frames, nb, false);
            if (ret < 0 || hipStreamSynchronize(s) != hipSuccess) {
                ret = ret < 0 ? ret : CRAW_ERROR(EIO);
                break;
            }
            if (!it)
                continue;   /* the first run of each candidate is a warm-up */
            n = craw_profiler_times(b->ctx, &b->prof, 0, st, CRAW_MAX_STAGES);
            for (int j = 0; j < n && j < CRAW_MAX_STAGES; j++) {
                /* the stages whose persistent grid wg_per_cu[kind] sets;
                 * the one-launch-a-frame stages (chroma up + colour,
                 * highlights, false colour) and the clear are not tuned */
                int kind = !strncmp(st[j].name, "denoise", 7) ? 0 :
                           (!strcmp(st[j].name, "rcd") ||
                            !strncmp(st[j].name, "debayer_", 8)) ? 1 :
                           !strcmp(st[j].name, "chroma_blur") ? 2 :
                           !strcmp(st[j].name, "awb_hist") ? 3 : -1;
                if (kind >= 0)
                    acc[kind] += st[j].ms;
            }
        }
        /* the other kinds' launches do not depend on the denoise shape:
         * they are timed at shape 0 only */
        for (int k = 0; k < 4; k++)
            if ((k == 0 || !sh) && acc[k] > 0.0 && acc[k] < best[k]) {
                best[k] = acc[k];
                pick[k] = cand[c];
                if (!k)
                    pick_shape = shapes[sh];
            }
    }
    craw_profiler_free(&b->prof);
    b->prof = saved;
    for (uint32_t i = 0; i < nb; i++)
        if (frames[i].state) {
            frames[i].state->prev = prev[i];
            frames[i].state->cur = cur[i];
        }
    if (ret < 0) {
        for (int k = 0; k < 4; k++)
            b->wg_per_cu[k] = 0;
        b->dn_shape = 0;
        return ret;
    }
    for (int k = 0; k < 4; k++) {
        b->wg_per_cu[k] = pick[k];
        if (best[k] < 1e30)
            total += best[k] / iterations;
    }
    b->dn_shape = pick_shape;
    craw_log(b->ctx, CRAW_LOG_VERBOSE, "Autotune: workgroups per CU %d/%d/%d/%d, denoise shape "
             "%d\n", pick[0], pick[1], pick[2], pick[3], pick_shape);
    return (int)(total * 1000.0);
}

} /* extern "C" */
