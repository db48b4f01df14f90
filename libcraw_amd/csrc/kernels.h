/*
 * libcraw for MI355X: the per-frame constants the kernels read, and the
 * launchers the host calls.
 *
 * Every stage is a per-tile device function. The plain form launches it
 * over one frame's tiles (tile = blockIdx); the batched form launches it
 * persistently over a work-list of tiles drawn from many frames of
 * different sizes, each workgroup taking tiles until the list is done, and
 * each tile reading its frame's constants from a device-resident table.
 */
#ifndef CRAW_KERNELS_H
#define CRAW_KERNELS_H

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>

#include "k_common.h"

/* Stage timing without markers between the launches. While a profiled
 * stage is open on this thread, craw_launch_events() hands its launches
 * the stage's events: the start event to the first launch and the stop
 * event to every launch, so the stage runs from its first kernel's start
 * to its last kernel's end, timestamped by the dispatches themselves
 * (hipExtLaunchKernel) rather than by event packets between them. Both
 * are NULL outside a profiled stage, and the launch is a plain one. */
void craw_launch_events(hipEvent_t *start, hipEvent_t *stop);

#define KLAUNCH(K, G, B, SHM, S, ...)                                                    \
    do {                                                                                \
        hipEvent_t kl_e0_, kl_e1_;                                                      \
        craw_launch_events(&kl_e0_, &kl_e1_);                                           \
        if (kl_e1_)                                                                     \
            hipExtLaunchKernelGGL(K, G, B, SHM, S, kl_e0_, kl_e1_, 0, __VA_ARGS__);     \
        else                                                                            \
            hipLaunchKernelGGL(K, G, B, SHM, S, __VA_ARGS__);                           \
    } while (0)

/* The chroma stage's buffers: camera RGB as three float planes of the
 * frame, the quarter grid (ceil(W/4) x ceil(H/4)) of block means as three
 * planes (R - G, B - G, G), and its bilateral result as two. The debayer
 * writes its planes to rgb_out, and the quarter means only with `means`:
 * with the false colour stage after it, rgb_out is that stage's input and
 * the stage writes rgb and the means (rgb_out == rgb otherwise). */
struct KInter {
    float *rgb;
    float *down;
    float *blur;
    int wq, hq;
    float *rgb_out;
    int means;
    /* the chroma-up + colour kernel: 1 the chroma stage's correction of
     * the planes, 0 the colour stage alone from them (the RCD debayer
     * without the chroma stage, when its gamut compression would otherwise
     * run inside the streamed RCD) */
    int up;
};

/* Constants of the raw-domain denoise (lib/pipeline.c:1414-1467 and
 * lib/ref/ref_denoise.c:26-46 of the reference), all formed on the host
 * in float exactly as the reference forms them. */
struct KDenoise {
    int w2, h2;                  /* plane size */
    int src_pitch;               /* uint16 elements */
    int dst_pitch;               /* uint16 elements */
    float black[4], range[4], inv_range[4];
    float gat_a, gat_c, gat_sc, gat_fwd, gat_inv, gat_off, gat_t0;
    float inv_h2;
    float guard_lo, guard_inv;
    float alpha_max;
};

/* Which filters run: spatial NLM, temporal average, and whether the
 * temporal one has a previous state to read. */
#define KDN_SPATIAL  1
#define KDN_TEMPORAL 2
#define KDN_HISTORY  4

struct KDenoiseState {
    const float *prev_level;     /* 4 planes of w2 x h2, or NULL */
    const uint8_t *prev_var;
    float *next_level;           /* or NULL without the temporal pass */
    uint8_t *next_var;
};

/* Everything one frame's kernels read. */
struct KFrame {
    int W, H;
    const uint16_t *src;         /* the mosaic the debayer reads */
    int src_pitch;               /* uint16 elements */
    int method;                  /* 0 bilinear, 1 MHC, 2 RCD */
    KPrep prep;
    KColor col;
    KOut out;
    KInter inter;
    float chroma_strength;
    /* denoise: input mosaic, constants, state, output mosaic */
    const uint16_t *dn_src;
    uint16_t *dn_dst;
    KDenoise dn;
    KDenoiseState st;
};

/* A tile of the batched form's work-list: frame index in the top 8 bits,
 * tile index within the frame's grid for the stage in the low 24. */
#define KTILE(frame, idx) (((uint32_t)(frame) << 24) | (uint32_t)(idx))
#define KTILE_FRAME(t) ((int)((t) >> 24))
#define KTILE_IDX(t) ((int)((t) & 0xffffff))

/* Tile geometry of each stage (frame pixels, or plane samples for the
 * denoise), for building work-lists on the host. */
#define KG_SIMPLE_TW 64
#define KG_SIMPLE_TH 16
#define KG_RCD_TW 32
#define KG_RCD_TH 16
#define KG_CB_T 16
#define KG_CU_TW 64
#define KG_CU_TH 16
#define KG_DN_TW 32
#define KG_DN_TH 16
#define KG_FT_TW 28
#define KG_FT_TH 64
#define KG_AWB_CHUNK 8192
#define KG_XS_TW 224
#ifndef KG_XS_TH
#define KG_XS_TH 48             /* 3 workgroups a CU in 53 KB of LDS */
#endif
#define KG_XS_TH_TALL 57        /* the taller band a batch's autotuner may pick */
#ifndef KG_XH_TH
#define KG_XH_TH 30             /* the same kernel with a temporal history */
#endif
#ifndef KG_XH_TH16
#define KG_XH_TH16 18           /* the precision form's: 3 workgroups a CU in 53 KB */
#endif
#define KG_RS_OUT 116           /* streamed RCD: output columns per strip */
#define KG_RS_OUT_F 112         /* the same without the chroma stage */
#define KG_RS_R   40            /* and rows */
#define KG_RS_NW  4             /* strips per workgroup */

enum KStage {
    KSTAGE_DENOISE = 0,
    KSTAGE_DEBAYER,
    KSTAGE_CHROMA_BLUR,
    KSTAGE_CHROMA_UP,
    KSTAGE_AWB,
    KSTAGE_RCD_RING,            /* tiled RCD tiles around the streamed interior */
    KSTAGE_RCD_STREAM,          /* groups of KG_RS_NW interior strips */
    KSTAGE_RCD_STREAM_F,        /* the same, strips to the output */
    KSTAGE_NB,
};

/* NLM forms: bit-exact FP32; FP32 reordered with the hardware exp2;
 * packed FP16 distances and weight arguments. */
#define KFORM_EXACT 0
#define KFORM_FAST  1
#define KFORM_FP16  2

/* Which denoise kernel's tiles: the exact tiled kernel of the temporal
 * filter alone (32x16), the fast forms' (28x64), the exact kernel with
 * pairwise distances for the spatial modes, without (224x48) and with
 * (224x30) a temporal history. */
enum KDnGeom {
    KDN_GEOM_DIRECT = 0,
    KDN_GEOM_FAST,
    KDN_GEOM_SYM,
    KDN_GEOM_SYMH,
    KDN_GEOM_SYMH16,            /* KDN_GEOM_SYMH of the precision form (224x18) */
    KDN_GEOM_SYM_TALL,          /* KDN_GEOM_SYM in taller bands (224x57) */
};

/* A batch's denoise launch shape, picked by its autotuner: the pairwise
 * kernel's taller band without history, and the fast form's search rows
 * unrolled (1 or all 7). */
#define KDN_SHAPE_TALL   1
#define KDN_SHAPE_UNROLL 2

static inline int craw_dn_geom(int form, int mode, int shape = 0)
{
    const int sym = (shape & KDN_SHAPE_TALL) ? KDN_GEOM_SYM_TALL : KDN_GEOM_SYM;

    /* the precision form's spatial modes share the exact form's pairwise
     * band kernel (with and without history), in packed FP16 */
    if (form == KFORM_FP16 && (mode & KDN_SPATIAL))
        return (mode & KDN_HISTORY) ? KDN_GEOM_SYMH16 : sym;
    if (form != KFORM_EXACT)
        return KDN_GEOM_FAST;
    if (!(mode & KDN_SPATIAL))
        return KDN_GEOM_DIRECT;
    return (mode & KDN_HISTORY) ? KDN_GEOM_SYMH : sym;
}

/* Tile geometry of `stage`: tw x th samples of a grid `div` times coarser
 * than the frame (2: the four planes of the mosaic; 4: the quarter grid of
 * the chroma stage), `planes` grids per frame. `method` picks the
 * debayer's tile, `dn_geom` (KDnGeom) the denoise's. */
static inline void craw_stage_geom(int stage, int method, int dn_geom, int *tw, int *th,
                                   int *div, int *planes)
{
    *tw = *th = *div = *planes = 1;
    switch (stage) {
    case KSTAGE_DENOISE:
        switch (dn_geom) {
        case KDN_GEOM_FAST:
            *tw = KG_FT_TW;
            *th = KG_FT_TH;
            break;
        case KDN_GEOM_SYM:
            *tw = KG_XS_TW;
            *th = KG_XS_TH;
            break;
        case KDN_GEOM_SYMH:
            *tw = KG_XS_TW;
            *th = KG_XH_TH;
            break;
        case KDN_GEOM_SYMH16:
            *tw = KG_XS_TW;
            *th = KG_XH_TH16;
            break;
        case KDN_GEOM_SYM_TALL:
            *tw = KG_XS_TW;
            *th = KG_XS_TH_TALL;
            break;
        default:
            *tw = KG_DN_TW;
            *th = KG_DN_TH;
        }
        *div = 2;
        *planes = 4;
        break;
    case KSTAGE_DEBAYER:
        *tw = method == 2 ? KG_RCD_TW : KG_SIMPLE_TW;
        *th = method == 2 ? KG_RCD_TH : KG_SIMPLE_TH;
        break;
    case KSTAGE_CHROMA_BLUR:
        *tw = *th = KG_CB_T;
        *div = 4;
        break;
    case KSTAGE_CHROMA_UP:
        *tw = KG_CU_TW;
        *th = KG_CU_TH;
        break;
    case KSTAGE_AWB:
        /* the 8x8-block grid in chunks of KG_AWB_CHUNK blocks, one row of
         * "tiles" per chunk */
        *div = 8;
        *tw = KG_AWB_CHUNK;
        break;
    }
}

/* The split of a frame for a streamed debayer: whole tw x th tiles whose
 * filter reach (`reach` pixels) lies inside the frame form the interior,
 * developed by the streamed kernel; the ring of tiles around it by the
 * tiled one. Tile units [itx0, itx1) x [ity0, ity1), empty for a frame too
 * small to have one. RCD splits 32x16 tiles with a reach of 5. */
struct KRcdSplit {
    int ntx, nty;
    int itx0, itx1, ity0, ity1;
    int x_lo, x_hi, y_lo, y_hi;
};

#ifdef __HIPCC__
__host__ __device__
#endif
static inline KRcdSplit craw_tile_split(int W, int H, int tw, int th, int reach)
{
    KRcdSplit s;

    s.ntx = (W + tw - 1) / tw;
    s.nty = (H + th - 1) / th;
    s.itx0 = 1;
    s.itx1 = W - tw * (s.ntx - 1) >= reach ? s.ntx - 1 : s.ntx - 2;
    s.ity0 = 1;
    s.ity1 = H - th * (s.nty - 1) >= reach ? s.nty - 1 : s.nty - 2;
    if (s.itx1 <= s.itx0 || s.ity1 <= s.ity0)
        s.itx0 = s.itx1 = s.ity0 = s.ity1 = 0;
    s.x_lo = tw * s.itx0;
    s.x_hi = tw * s.itx1;
    s.y_lo = th * s.ity0;
    s.y_hi = th * s.ity1;
    return s;
}

#ifdef __HIPCC__
__host__ __device__
#endif
static inline KRcdSplit craw_rcd_split(int W, int H)
{
    return craw_tile_split(W, H, KG_RCD_TW, KG_RCD_TH, 5);
}

/* Ring tile i: the rows above the interior, the columns either side of it,
 * the rows below; all tiles when the interior is empty. */
#ifdef __HIPCC__
__host__ __device__
#endif
static inline void craw_ring_tile(const KRcdSplit &s, int i, int *tx, int *ty)
{
    const int top = s.ity0 * s.ntx, per = s.ntx - (s.itx1 - s.itx0);
    const int mid = (s.ity1 - s.ity0) * per;

    if (i < top) {
        *ty = i / s.ntx;
        *tx = i % s.ntx;
    } else if (i < top + mid) {
        const int j = (i - top) % per;
        *ty = s.ity0 + (i - top) / per;
        *tx = j < s.itx0 ? j : s.itx1 + (j - s.itx0);
    } else {
        *ty = s.ity1 + (i - top - mid) / s.ntx;
        *tx = (i - top - mid) % s.ntx;
    }
}

/* The denoise stage's workgroups in plane-pair order. Planes 0 and 1 (and
 * 2 and 3) are the two columns of the same mosaic rows, so every cache
 * line a workgroup of one reads or writes holds its partner's samples too.
 * Entry i = 16 g + 8 m + k (k < 8) is plane 2 py + m of tile t, where
 * 2 t + py = 8 g + k: the two planes of a pair are entries i and i + 8,
 * which the dispatcher sends to the same XCD (workgroups go to the eight
 * XCDs in turn) one after the other, so the second finds the lines in
 * that XCD's L2. A frame of nt tiles a plane has 16 ceil(nt / 4) entries;
 * those past 2 nt pairs are empty. */
#ifdef __HIPCC__
__host__ __device__
#endif
static inline int craw_dn_pair_count(int nt)
{
    return 16 * ((nt + 3) / 4);
}

#ifdef __HIPCC__
__host__ __device__
#endif
static inline bool craw_dn_pair_tile(int i, int nt, int *t, int *pl)
{
    const int u = 8 * (i >> 4) + (i & 7);

    *t = u >> 1;
    *pl = ((u & 1) << 1) | ((i >> 3) & 1);
    return u < 2 * nt;
}

/* Number of tiles of `stage` for a frame of W x H (even W and H; the
 * quarter grid rounds up; the denoise stage's in plane-pair order). */
static inline int craw_stage_tiles(int stage, int W, int H, int method, int dn_geom, int *ntx)
{
    int tw, th, div, planes, w, h;

    if (stage == KSTAGE_RCD_RING || stage == KSTAGE_RCD_STREAM ||
        stage == KSTAGE_RCD_STREAM_F) {
        const KRcdSplit sp = craw_rcd_split(W, H);
        const int outw = stage == KSTAGE_RCD_STREAM_F ? KG_RS_OUT_F : KG_RS_OUT;
        const int nstrips = ((sp.x_hi - sp.x_lo + outw - 1) / outw) *
                            ((sp.y_hi - sp.y_lo + KG_RS_R - 1) / KG_RS_R);

        *ntx = 1;
        if (stage == KSTAGE_RCD_RING)
            return sp.ntx * sp.nty - (sp.itx1 - sp.itx0) * (sp.ity1 - sp.ity0);
        return (nstrips + KG_RS_NW - 1) / KG_RS_NW;
    }
    if (stage == KSTAGE_AWB) {
        int nblocks = ((W + 7) / 8) * ((H + 7) / 8);
        *ntx = (nblocks + KG_AWB_CHUNK - 1) / KG_AWB_CHUNK;
        return *ntx;
    }
    craw_stage_geom(stage, method, dn_geom, &tw, &th, &div, &planes);
    w = (W + div - 1) / div;
    h = (H + div - 1) / div;
    *ntx = (w + tw - 1) / tw;
    if (stage == KSTAGE_DENOISE)
        return craw_dn_pair_count(*ntx * ((h + th - 1) / th));
    return planes * *ntx * ((h + th - 1) / th);
}

/* The noise fit's device measurement (k_noise.hip): the frame's R16 samples,
 * per site black and range in double (the range formed in float), and the
 * scratch it fills: a key per plane sample, and per bin the count, the
 * fixed-point sum of means in KNZ_LIMBS 32-bit limbs (units of 2^-149,
 * wide enough for any finite float and 2^32 samples), a KNZ_D-digit histogram row and the selection state (prefix,
 * rank; after the last level the median's bits). */
#define KNZ_MAX_BINS 512
#define KNZ_LIMBS 9
#define KNZ_D 2048
struct KNoise {
    const uint16_t *src;
    size_t pitch;                /* uint16 elements */
    int w2, h2;
    double black[4], range[4];
    uint32_t nb_bins;
    float max_level;
    uint64_t *keys;              /* 4 planes of w2 x h2 */
    uint32_t *count;
    unsigned long long *limbs;
    uint32_t *hist;
    uint32_t *sel;
};

hipError_t craw_launch_noise_measure(const KNoise &n, hipStream_t s);

hipError_t craw_launch_zero_u32(uint32_t *p, size_t n, hipStream_t s);
/* n words (a multiple of 4, 16-byte aligned) of hist to the device-visible
 * host buffer host, zeroing hist. */
hipError_t craw_launch_hist_export(uint32_t *hist, uint32_t *host, size_t n, hipStream_t s);

/* Plain form: one frame, `f` by value. */
hipError_t craw_launch_debayer(bool to_inter, const KFrame &f, hipStream_t s);
hipError_t craw_launch_chroma_blur(const KFrame &f, hipStream_t s);
hipError_t craw_launch_chroma_up_finish(const KFrame &f, hipStream_t s);
/* The chroma-up kernel's host handle, as a graph's kernel nodes name it. */
const void *craw_chroma_up_kernel(void);
/* The same for the highlight stage's kernels (0 field, 1 coarse, 2 push,
 * 3 fill) and the false colour kernels, whose arguments are also a frame's
 * constants. */
const void *craw_hl_kernel(int which);
bool craw_is_false_color_kernel(const void *fn);
/* RCD over the frame, into the chroma stage or to the output: the
 * interior of craw_rcd_split() streamed and the ring of tiles around it,
 * in one launch. */
hipError_t craw_launch_rcd(bool to_inter, const KFrame &f, hipStream_t s);
hipError_t craw_launch_batch_rcd_stream(bool to_inter, bool gamut, const KFrame *frames,
                                        const uint32_t *tiles, uint32_t ntiles, int grid,
                                        hipStream_t s);

/* The highlight reconstruction's inputs and outputs: the mosaic it
 * rebuilds (the caller's or the denoised one), the R16F mosaic it writes,
 * the colour field's levels (level 0 a texel per 4x4 block, halved to
 * 1x1), and the per-site constants, formed on the host in float as
 * lib/ref/ref_highlight.c:120-130 forms them. */
#define KHL_LEVELS 16

struct KHighlight {
    const uint16_t *src;
    int src_pitch;               /* uint16 elements */
    uint16_t *mosaic;            /* R16F bits, W x H */
    float4 *level[KHL_LEVELS];   /* (sum w c_r, sum w c_b, sum w) / pushed (c_r, c_b, w) */
    int lw[KHL_LEVELS], lh[KHL_LEVELS];
    int nl;
    int W, H;
    float black[4], inv_range[4];
    float wb[3];
    int float_math;              /* the precision form: float log/exp (kf_log, kf_exp) */
};

/* Level 0 and its pull to levels 1-4; the rest of the pull and every push;
 * the fill into h.mosaic. */
hipError_t craw_launch_hl_field(const KHighlight &h, hipStream_t s);
hipError_t craw_launch_hl_push(const KHighlight &h, hipStream_t s);
hipError_t craw_launch_hl_fill(const KHighlight &h, hipStream_t s);
/* How many launches craw_launch_hl_push makes for `h`. */
int craw_hl_push_launches(const KHighlight &h);

/* False colour suppression, `passes` 1..5, on f.inter.rgb: into the chroma
 * stage's inputs `to` (planes and quarter means) when chroma is set, else
 * through the AWB block samples and the colour stage into f.out. */
hipError_t craw_launch_false_color(int passes, bool chroma, const KFrame &f, const KInter &to,
                                   hipStream_t s);

/* The AWB histogram from the finish stage's block samples. */
hipError_t craw_launch_awb_hist(const KFrame &f, hipStream_t s);

/* mode: KDN_* bits. The FP32 form that reproduces the CPU reference bit
 * for bit. */
hipError_t craw_launch_denoise(int mode, const KFrame &f, hipStream_t s);

/* The throughput forms: FP32 with reordered sums and the hardware exp2, or
 * with fp16 the packed-FP16 distances and weight arguments. */
hipError_t craw_launch_denoise_fast(int mode, bool fp16, const KFrame &f, hipStream_t s);

/* The exact form with pairwise weights (modes without a history). */
hipError_t craw_launch_denoise_sym(int mode, const KFrame &f, hipStream_t s);
hipError_t craw_launch_denoise_sym16(int mode, const KFrame &f, hipStream_t s);
hipError_t craw_launch_batch_denoise_sym16(int mode, const KFrame *frames, const uint32_t *tiles,
                                           uint32_t ntiles, int grid, bool tall, hipStream_t s);
hipError_t craw_launch_batch_denoise_sym(int mode, const KFrame *frames, const uint32_t *tiles,
                                         uint32_t ntiles, int grid, bool tall, hipStream_t s);

/* Batched form: `frames` is a device table, `tiles` a device work-list of
 * `ntiles` KTILE entries for the stage, `grid` the number of persistent
 * workgroups. All frames share the method, the denoise mode and whether
 * the chroma stage runs. */
hipError_t craw_launch_batch_denoise(int mode, int form, int shape, const KFrame *frames,
                                     const uint32_t *tiles, uint32_t ntiles, int grid,
                                     hipStream_t s);
hipError_t craw_launch_batch_debayer(int method, bool to_inter, const KFrame *frames,
                                     const uint32_t *tiles, uint32_t ntiles, int grid,
                                     hipStream_t s);
hipError_t craw_launch_batch_chroma_blur(const KFrame *frames, const uint32_t *tiles,
                                         uint32_t ntiles, int grid, hipStream_t s);
hipError_t craw_launch_batch_awb_hist(const KFrame *frames, const uint32_t *tiles,
                                      uint32_t ntiles, int grid, hipStream_t s);

#endif /* CRAW_KERNELS_H */
