/*
 * libcraw for MI355X: debayer, chroma denoise, AWB histogram and colour
 * kernels.
 *
 * A workgroup develops one tile of the frame. The prepped mosaic (and, for
 * RCD, each intermediate of its three passes) is staged in LDS over the
 * tile plus the apron the later stencils reach, so a sample is read from
 * HBM once and every stencil tap is an LDS read. Out-of-frame positions of
 * an apron hold the value at the mirrored position, and every stage
 * evaluates its stencil at the mirrored position's own neighbourhood:
 * exactly the reference's "fetch through the mirror" semantics
 * (lib/ref/ref_debayer.c:20-40) for every intermediate, at any border.
 *
 * The finished camera RGB of a tile stays in LDS for the AWB histogram
 * (one sample per whole 8x8 block, tiles being multiples of 8) and for the
 * colour stage, so without chroma denoise the whole chain is one kernel:
 * 2 bytes in, 8 or 16 bytes out per pixel. With chroma denoise the
 * debayer writes the camera RGB once (three float planes) and the 4x4
 * block means, the bilateral runs on the quarter grid, and a last kernel
 * puts the band back, builds the histogram and applies the colour stage.
 */
#include "k_common.h"
#include "kernels.h"
#include "k_finish.h"
#include "k_rcd_tile.h"
#include "detmath_tab.h"

/* -------------------------------------------------- bilinear and MHC */

#define BL_TW 64
#define BL_TH 16
#define BL_A  2
#define BL_NT 256

struct SmSimple {
    float vs[BL_TH + 2 * BL_A][BL_TW + 2 * BL_A];
    float rgbs[3][BL_TH][BL_TW];
};

/* DIRECT: the colour stage and the store straight from the pixel's
 * camera RGB (no AWB histogram, which sums 8x8 blocks across threads). */
template <int METHOD, bool INTERIOR, bool DIRECT = false>
KDEV void dev_debayer_simple(SmSimple &sm, const KFrame &f, int x0, int y0)
{
    constexpr int RW = BL_TW + 2 * BL_A, RH = BL_TH + 2 * BL_A;
    float (&vs)[RH][RW] = sm.vs;
    float (&rgbs)[3][BL_TH][BL_TW] = sm.rgbs;
    const uint16_t *__restrict__ src = f.src;
    const int src_pitch = f.src_pitch, W = f.W, H = f.H;
    const KPrep &prep = f.prep;
    const int tid = threadIdx.x;
    const int xe = min(x0 + BL_TW, W), ye = min(y0 + BL_TH, H);
    const int rw = xe - x0 + 2 * BL_A, rh = ye - y0 + 2 * BL_A;

    /* compile-time strides; in an interior tile no position needs the
     * mirror, and where the mosaic's rows are 4-byte aligned a thread
     * loads a pair of samples (x0 - BL_A is even) as one 32-bit word */
    static_assert(RW % 2 == 0 && BL_A % 2 == 0, "sample pairs");
    if (INTERIOR && ((uintptr_t)src & 3u) == 0 && (src_pitch & 1) == 0) {
        for (int i = tid; i < RW / 2 * RH; i += BL_NT) {
            const int lx = 2 * (i % (RW / 2)), ly = i / (RW / 2);
            const int mx = x0 - BL_A + lx, my = y0 - BL_A + ly;
            const uint32_t pair = *(const uint32_t *)(src + (size_t)my * src_pitch + mx);

            vs[ly][lx] = kprep(pair & 0xffffu, (my & 1) * 2, prep);
            vs[ly][lx + 1] = kprep(pair >> 16, (my & 1) * 2 + 1, prep);
        }
    } else {
        for (int i = tid; i < RW * RH; i += BL_NT) {
            int lx = i % RW, ly = i / RW;
            int mx = x0 - BL_A + lx, my = y0 - BL_A + ly, s;

            if (!INTERIOR) {
                if (lx >= rw || ly >= rh)
                    continue;
                mx = kmirror(mx, W);
                my = kmirror(my, H);
            }
            s = (my & 1) * 2 + (mx & 1);
            vs[ly][lx] = kprep(src[(size_t)my * src_pitch + mx], s, prep);
        }
    }
    __syncthreads();

#define V(dx, dy) vs[py + BL_A + (dy)][px + BL_A + (dx)]
#pragma unroll
    for (int k = 0; k < BL_TW * BL_TH / BL_NT; k++) {
        int p = tid + k * BL_NT, px = p % BL_TW, py = p / BL_TW;
        int x = x0 + px, y = y0 + py;
        int s = (y & 1) * 2 + (x & 1);
        float r, g, b;

        if (x >= W || y >= H)
            continue;
        if (METHOD == 0) {
            /* lib/ref/ref_debayer.c:65-96 */
            float edge = 0.25f * (V(0, -1) + V(0, 1) + V(-1, 0) + V(1, 0));
            float corner = 0.25f * (V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1));
            float horiz = 0.5f * (V(-1, 0) + V(1, 0));
            float vert = 0.5f * (V(0, -1) + V(0, 1));

            if (s == 0)      { r = V(0, 0); g = edge;    b = corner; }
            else if (s == 1) { r = horiz;   g = V(0, 0); b = vert; }
            else if (s == 2) { r = vert;    g = V(0, 0); b = horiz; }
            else             { r = corner;  g = edge;    b = V(0, 0); }
        } else {
            /* lib/ref/ref_debayer.c:98-176, the four MHC kernels over 8 */
            float c = V(0, 0);
            float green = (4.0f * c + 2.0f * (V(0, -1) + V(0, 1) + V(-1, 0) + V(1, 0))
                           - (V(0, -2) + V(0, 2) + V(-2, 0) + V(2, 0))) * 0.125f;
            float diag = (6.0f * c + 2.0f * (V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1))
                          - 1.5f * (V(0, -2) + V(0, 2) + V(-2, 0) + V(2, 0))) * 0.125f;
            float ch = (5.0f * c + 4.0f * (V(-1, 0) + V(1, 0))
                        - (V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1) + V(-2, 0) + V(2, 0))
                        + 0.5f * (V(0, -2) + V(0, 2))) * 0.125f;
            float cv = (5.0f * c + 4.0f * (V(0, -1) + V(0, 1))
                        - (V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1) + V(0, -2) + V(0, 2))
                        + 0.5f * (V(-2, 0) + V(2, 0))) * 0.125f;

            if (s == 0)      { r = c;    g = green; b = diag; }
            else if (s == 1) { r = ch;   g = c;     b = cv; }
            else if (s == 2) { r = cv;   g = c;     b = ch; }
            else             { r = diag; g = green; b = c; }
        }
        if constexpr (DIRECT) {
            float o[3];

            kcolor(f.col, r, g, b, o);
            kstore_rgba(f.out, x, y, o);
        } else {
            rgbs[0][py][px] = r;
            rgbs[1][py][px] = g;
            rgbs[2][py][px] = b;
        }
    }
#undef V
}

template <int METHOD, bool INTER>
KDEV void dev_debayer_simple_tile(SmSimple &sm, const KFrame &f, int x0, int y0)
{
    const bool interior = x0 >= BL_A && y0 >= BL_A && x0 + BL_TW + BL_A <= f.W &&
                          y0 + BL_TH + BL_A <= f.H;

    /* to the output without the histogram, a thread colours the pixels it
     * interpolated (the colour loop's pixel order is the debayer's), so
     * the tile's RGB needs no trip through LDS (MHC at 5760x3240 0.086 ->
     * 0.082 ms) */
    if (!INTER && !f.out.hist) {
        if (interior)
            dev_debayer_simple<METHOD, true, true>(sm, f, x0, y0);
        else
            dev_debayer_simple<METHOD, false, true>(sm, f, x0, y0);
        return;
    }
    if (interior)
        dev_debayer_simple<METHOD, true>(sm, f, x0, y0);
    else
        dev_debayer_simple<METHOD, false>(sm, f, x0, y0);
    __syncthreads();
    if (INTER)
        kstore_inter<BL_TW, BL_TH, BL_NT>(sm.rgbs, x0, y0, f.W, f.H, f.inter);
    else
        kfinish<BL_TW, BL_TH, BL_NT>(sm.rgbs, x0, y0, f.W, f.H, f.col, f.out);
}

/* RCD tiles: k_rcd_tile.h */

/* ------------------------------------------------------------ chroma */

#define CB_TX 16
#define CB_TY 16

/* lib/ref/ref_chroma.c:86-117: 9x9 joint bilateral on the quarter grid,
 * guided by G, one reflection at the borders, summed row by row. The
 * tile's three quarter-grid planes are staged in LDS with the 4-sample
 * reach. The exact form's tap weights go through the table-driven exp
 * (detmath.h craw_dm_expf_nonpos_t1024: the same float as the oracle's
 * craw_dm_expf_nonpos for every argument), its table staged in LDS too. */
#define CB_R 4
#define CB_LW (CB_TX + 2 * CB_R)
#define CB_LH (CB_TY + 2 * CB_R)

struct SmCb {
    float4 q[CB_LH][CB_LW];     /* a cell's R-G, B-G, G (w unused): one read a tap */
    double tab[1024];           /* 2^(j/1024), the exact form's exp table */
};

template <bool FLOAT_MATH>
KDEV void dev_chroma_blur(SmCb &sm, const KFrame &f, int qx0, int qy0)
{
    const KInter &it = f.inter;
    const int tid = threadIdx.x;
    const int lx = tid % CB_TX, ly = tid / CB_TX;
    const int qx = qx0 + lx, qy = qy0 + ly;
    const size_t qplane = (size_t)it.wq * it.hq;
    const int wq = it.wq, hq = it.hq;
    float c0, c1, c2, level, inv_range, inv_chroma;
    float acc0 = 0.0f, acc1 = 0.0f, wsum = 0.0f;

    static_assert(1024 % (CB_TX * CB_TY) == 0, "table entries a thread");
    if constexpr (!FLOAT_MATH) {
#pragma unroll
        for (int i = 0; i < 1024 / (CB_TX * CB_TY); i++)
            sm.tab[tid + i * CB_TX * CB_TY] = craw_dm_exp2_tab1024[tid + i * CB_TX * CB_TY];
    }
    /* positions up to the reach past the last valid sample, mirrored once
     * (wq, hq >= 5) */
    for (int i = tid; i < CB_LW * CB_LH; i += CB_TX * CB_TY) {
        const int sx0 = i % CB_LW, sy0 = i / CB_LW;
        int sx = qx0 - CB_R + sx0, sy = qy0 - CB_R + sy0;
        size_t si;

        if (sx >= wq + CB_R || sy >= hq + CB_R)
            continue;
        sx = sx < 0 ? -sx : (sx >= wq ? 2 * wq - 2 - sx : sx);
        sy = sy < 0 ? -sy : (sy >= hq ? 2 * hq - 2 - sy : sy);
        si = (size_t)sy * wq + sx;
        sm.q[sy0][sx0] = make_float4(it.down[si], it.down[qplane + si],
                                     it.down[2 * qplane + si], 0.0f);
    }
    __syncthreads();

    if (qx >= wq || qy >= hq)
        return;
    {
        const float4 c = sm.q[ly + CB_R][lx + CB_R];

        c0 = c.x;
        c1 = c.y;
        c2 = c.z;
    }
    level = (c2 > 0.0f ? c2 : 0.0f) + 0.02f;
    inv_range = 1.0f / (0.35f * level);
    inv_chroma = 1.0f / (0.35f * level);

    /* all 81 taps unrolled: each tap's spatial term is a constant */
#pragma unroll
    for (int dy = -CB_R; dy <= CB_R; dy++) {
#pragma unroll
        for (int dx = -CB_R; dx <= CB_R; dx++) {
            const float4 t = sm.q[ly + CB_R + dy][lx + CB_R + dx];
            const float s0 = t.x, s1 = t.y, s2 = t.z;
            const float dg = (s2 - c2) * inv_range;
            const float dd0 = (s0 - c0) * inv_chroma;
            const float dd1 = (s1 - c1) * inv_chroma;
            const float arg = -(float)(dx * dx + dy * dy) * 0.125f - dg * dg - dd0 * dd0 -
                              dd1 * dd1;
            const float wgt = FLOAT_MATH ? kf_exp(arg) : craw_dm_expf_nonpos_t1024(arg, sm.tab);

            acc0 += wgt * s0;
            acc1 += wgt * s1;
            wsum += wgt;
        }
    }
    it.blur[(size_t)qy * wq + qx] = acc0 / wsum;
    it.blur[qplane + (size_t)qy * wq + qx] = acc1 / wsum;
}

#define CU_TW 64
#define CU_TH 16
#define CU_NT 256

/* lib/ref/ref_chroma.c:119-155: the difference of the filtered and the
 * unfiltered band, bilinearly up to every pixel and blended by the
 * strength, into R and B; then the histogram and the colour stage. With
 * f.inter.up 0, the histogram and the colour stage of the planes alone. */
/* The quarter-grid cells a 64x16 tile's bilinear taps can reach: one
 * before and one after, clamped to the grid as the reference clamps. */
#define CU_QW (CU_TW / 4 + 2)
#define CU_QH (CU_TH / 4 + 2)

struct SmUp {
    float rgbs[3][CU_TH][CU_TW];
    float4 q[CU_QH][CU_QW];       /* a cell's down R-G, down B-G, blur R-G, blur B-G */
};

KDEV void dev_chroma_up_finish(SmUp &sm, const KFrame &f, int x0, int y0)
{
    float (&rgbs)[3][CU_TH][CU_TW] = sm.rgbs;
    const KInter &it = f.inter;
    const float strength = f.chroma_strength;
    const int W = f.W, H = f.H;
    const int tid = threadIdx.x;
    const size_t plane = (size_t)W * H, qplane = (size_t)it.wq * it.hq;
    const int qbx = x0 / 4 - 1, qby = y0 / 4 - 1;

    if (!it.up) {
        /* the colour stage alone from the planes */
#pragma unroll
        for (int k = 0; k < CU_TW * CU_TH / CU_NT; k++) {
            const int p = tid + k * CU_NT, px = p % CU_TW, py = p / CU_TW;
            const int x = x0 + px, y = y0 + py;

            if (x >= W || y >= H)
                continue;
            rgbs[0][py][px] = kglobal(it.rgb)[(size_t)y * W + x];
            rgbs[1][py][px] = kglobal(it.rgb)[plane + (size_t)y * W + x];
            rgbs[2][py][px] = kglobal(it.rgb)[2 * plane + (size_t)y * W + x];
        }
    } else {
        /* cell (j, i) of the stage holds grid cell clamp(qbx + i, qby + j),
         * its four planes side by side (one 16-byte read a tap) */
        for (int i = tid; i < CU_QH * CU_QW; i += CU_NT) {
            const int lj = i / CU_QW, li = i % CU_QW;
            int gx = qbx + li, gy = qby + lj;
            size_t q;

            gx = gx > 0 ? (gx < it.wq - 1 ? gx : it.wq - 1) : 0;
            gy = gy > 0 ? (gy < it.hq - 1 ? gy : it.hq - 1) : 0;
            q = (size_t)gy * it.wq + gx;
            sm.q[lj][li] = make_float4(kglobal(it.down)[q], kglobal(it.down)[qplane + q],
                                       kglobal(it.blur)[q], kglobal(it.blur)[qplane + q]);
        }
        __syncthreads();

        /* lib/ref/ref_chroma.c:128-140: pixel x's taps are the cells
         * floor(f) and floor(f) + 1 of f = (x + 0.5) / 4 - 0.5, with weight
         * frac(f) on the second. f is exact in float (x < 2^22), so floor(f)
         * = (x - 2) >> 2 and frac(f) = (2 (x & 3) + 5) / 8 mod 1: 0.625,
         * 0.875, 0.125, 0.375 for x mod 4 = 0 .. 3; likewise in y. A
         * thread's pixels (tid + 256 k of the 64x16 tile, x0 and y0 multiples
         * of 4) share x mod 4 and y mod 4, so its four weights, products of
         * those fractions (exact, the values the reference's float products
         * give), are formed once; the taps' staged cells are (px + 2) >> 2
         * and (py + 2) >> 2 (the clamp is in the staged cells). */
        static_assert(CU_NT % CU_TW == 0 && (CU_NT / CU_TW) % 4 == 0, "pixel phases per thread");
        const int px = tid % CU_TW, py0 = tid / CU_TW;
        const float tx = (float)((2 * (px & 3) + 5) & 7) * 0.125f;
        const float ty = (float)((2 * (py0 & 3) + 5) & 7) * 0.125f;
        const float w00 = (1.0f - tx) * (1.0f - ty), w10 = tx * (1.0f - ty);
        const float w01 = (1.0f - tx) * ty, w11 = tx * ty;
        const int lx0 = (px + 2) >> 2, lx1 = lx0 + 1;
        const int x = x0 + px;

#pragma unroll
        for (int k = 0; k < CU_TW * CU_TH / CU_NT; k++) {
            const int py = py0 + k * (CU_NT / CU_TW), y = y0 + py;
            const int ly0 = (py + 2) >> 2, ly1 = ly0 + 1;
            float r, g, b, d[2];

            if (x >= W || y >= H)
                continue;
            r = kglobal(it.rgb)[(size_t)y * W + x];
            g = kglobal(it.rgb)[plane + (size_t)y * W + x];
            b = kglobal(it.rgb)[2 * plane + (size_t)y * W + x];

            d[0] = r - g;
            d[1] = b - g;
            {
                const float4 t00 = sm.q[ly0][lx0], t01 = sm.q[ly0][lx1];
                const float4 t10 = sm.q[ly1][lx0], t11 = sm.q[ly1][lx1];
                const float lo0 = w00 * t00.x + w10 * t01.x + w01 * t10.x + w11 * t11.x;
                const float lf0 = w00 * t00.z + w10 * t01.z + w01 * t10.z + w11 * t11.z;
                const float lo1 = w00 * t00.y + w10 * t01.y + w01 * t10.y + w11 * t11.y;
                const float lf1 = w00 * t00.w + w10 * t01.w + w01 * t10.w + w11 * t11.w;

                d[0] += strength * (lf0 - lo0);
                d[1] += strength * (lf1 - lo1);
            }
            rgbs[0][py][px] = g + d[0];
            rgbs[1][py][px] = g;
            rgbs[2][py][px] = g + d[1];
        }
    }
    __syncthreads();

    kfinish<CU_TW, CU_TH, CU_NT>(rgbs, x0, y0, W, H, f.col, f.out);
}

/* ------------------------------------------------------ AWB histogram */

/* The histogram of one chunk of a frame's block samples, privatised in an
 * LDS hash of bins: each lane merges the runs of equal bins among its 8
 * consecutive samples (neighbouring blocks often share a bin), inserts
 * each run into the hash with LDS atomics, and the workgroup then adds
 * every occupied entry to the frame's histogram with three global
 * atomics. A run whose bin finds no slot within AWB_PROBES probes goes to
 * the global histogram directly, so the stage terminates whatever the
 * data. Integer sums: the result does not depend on the grouping.
 * A chunk of 8192 samples (a 5.8K frame has 36) is shared by 1024
 * threads: a lane's chain of hash inserts, each waiting for its
 * compare-and-swap, is what bounds a chunk's time (0.020 ms a 5.8K frame
 * with 256 threads of 32 samples, 0.014 with 1024 of 8; chunks of 4096 or
 * 16384 samples at 1024 threads: 0.016 and 0.020). */
#define AWB_NT 1024
#define AWB_PER (KG_AWB_CHUNK / AWB_NT)
#define AWB_HASH 4096
#define AWB_PROBES 32

struct SmAwb {
    uint32_t key[AWB_HASH];
    uint32_t cnt[AWB_HASH];
    uint32_t w[AWB_HASH];
    uint32_t b[AWB_HASH];
};

KDEV void awb_insert(SmAwb &sm, uint32_t *hist, uint32_t bin, uint32_t c, uint32_t w,
                     uint32_t b)
{
    uint32_t h = (bin * 0x9e3779b1u) >> 20;

    for (int probe = 0; probe < AWB_PROBES; probe++) {
        const uint32_t k = atomicCAS(&sm.key[h], KAWB_NONE, bin);

        if (k == KAWB_NONE || k == bin) {
            atomicAdd(&sm.cnt[h], c);
            atomicAdd(&sm.w[h], w);
            atomicAdd(&sm.b[h], b);
            return;
        }
        h = (h + 1) & (AWB_HASH - 1);
    }
    atomicAdd(&hist[bin], c);
    atomicAdd(&hist[16384 + bin], w);
    atomicAdd(&hist[32768 + bin], b);
}

KDEV void dev_awb_hist(SmAwb &sm, const KFrame &f, int chunk)
{
    const int tid = threadIdx.x;
    const int nblocks = ((f.W + 7) / 8) * ((f.H + 7) / 8);
    const int i0 = chunk * KG_AWB_CHUNK + tid * AWB_PER;
    const uint2 *__restrict__ smp = f.out.samples;
    uint32_t *hist = f.out.hist;
    uint32_t rb = KAWB_NONE, rc = 0, rw = 0, rbr = 0;

    for (int i = tid; i < AWB_HASH; i += AWB_NT) {
        sm.key[i] = KAWB_NONE;
        sm.cnt[i] = 0u;
        sm.w[i] = 0u;
        sm.b[i] = 0u;
    }
    __syncthreads();

    for (int k = 0; k < AWB_PER && i0 + k < nblocks; k++) {
        const uint2 v = smp[i0 + k];
        uint32_t bin;

        if (v.x == KAWB_NONE)
            continue;
        bin = v.x & 0x3fffu;
        if (bin == rb) {
            rc += 1u;
            rw += v.x >> 14;
            rbr += v.y;
        } else {
            if (rb != KAWB_NONE)
                awb_insert(sm, hist, rb, rc, rw, rbr);
            rb = bin;
            rc = 1u;
            rw = v.x >> 14;
            rbr = v.y;
        }
    }
    if (rb != KAWB_NONE)
        awb_insert(sm, hist, rb, rc, rw, rbr);
    __syncthreads();

    for (int i = tid; i < AWB_HASH; i += AWB_NT) {
        const uint32_t k = sm.key[i];

        if (k != KAWB_NONE) {
            atomicAdd(&hist[k], sm.cnt[i]);
            atomicAdd(&hist[16384 + k], sm.w[i]);
            atomicAdd(&hist[32768 + k], sm.b[i]);
        }
    }
}

__global__ __launch_bounds__(AWB_NT) void k_awb_hist(KFrame f)
{
    __shared__ SmAwb sm;
    dev_awb_hist(sm, f, blockIdx.x);
}

__global__ __launch_bounds__(AWB_NT) void kb_awb_hist(const KFrame *__restrict__ frames,
                                                      const uint32_t *__restrict__ tiles,
                                                      uint32_t ntiles)
{
    __shared__ SmAwb sm;

    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t d = tiles[t];

        __syncthreads();
        dev_awb_hist(sm, frames[KTILE_FRAME(d)], KTILE_IDX(d));
    }
}

/* ------------------------------------------------- plain and batched */

template <int METHOD, bool INTER>
__global__ __launch_bounds__(BL_NT) void k_debayer_simple(KFrame f)
{
    __shared__ SmSimple sm;
    dev_debayer_simple_tile<METHOD, INTER>(sm, f, blockIdx.x * BL_TW, blockIdx.y * BL_TH);
}

__global__ __launch_bounds__(CB_TX * CB_TY) void k_chroma_blur(KFrame f)
{
    __shared__ SmCb sm;
    if (f.col.flags & KCOLOR_FLOAT_MATH)
        dev_chroma_blur<true>(sm, f, blockIdx.x * CB_TX, blockIdx.y * CB_TY);
    else
        dev_chroma_blur<false>(sm, f, blockIdx.x * CB_TX, blockIdx.y * CB_TY);
}

/* Seven waves a SIMD: 72 VGPRs without a spill (the compiler's own choice
 * is 80, six waves); on a 5.8K frame 0.147 ms a launch against 0.161 at
 * six, and 0.178 at eight, where it spills. */
__global__ __launch_bounds__(CU_NT) __attribute__((amdgpu_waves_per_eu(7))) void
k_chroma_up_finish(KFrame f)
{
    __shared__ SmUp sm;
    dev_chroma_up_finish(sm, f, blockIdx.x * CU_TW, blockIdx.y * CU_TH);
}

/* Persistent forms: workgroup b takes tiles b, b + grid, b + 2 grid, ...
 * of the work-list; every workgroup leaves the loop after the last tile.
 * The barrier at the top of each iteration keeps a tile's LDS from being
 * overwritten while another wave still reads the previous one. */
template <int METHOD, bool INTER>
__global__ __launch_bounds__(BL_NT) void kb_debayer_simple(const KFrame *__restrict__ frames,
                                                           const uint32_t *__restrict__ tiles,
                                                           uint32_t ntiles)
{
    __shared__ SmSimple sm;

    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t d = tiles[t];
        const KFrame &f = frames[KTILE_FRAME(d)];
        const int ntx = (f.W + BL_TW - 1) / BL_TW, i = KTILE_IDX(d);

        __syncthreads();
        dev_debayer_simple_tile<METHOD, INTER>(sm, f, (i % ntx) * BL_TW, (i / ntx) * BL_TH);
    }
}

template <bool INTER>
__global__ __launch_bounds__(RC_NT) void kb_debayer_rcd(const KFrame *__restrict__ frames,
                                                        const uint32_t *__restrict__ tiles,
                                                        uint32_t ntiles)
{
    __shared__ SmRcd sm;

    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t d = tiles[t];
        const KFrame &f = frames[KTILE_FRAME(d)];
        const int ntx = (f.W + RC_TW - 1) / RC_TW, i = KTILE_IDX(d);
        int tx, ty;

        /* entries are ring tiles */
        (void)ntx;
        craw_ring_tile(craw_rcd_split(f.W, f.H), i, &tx, &ty);
        __syncthreads();
        dev_debayer_rcd_tile<INTER>(sm, f, tx * RC_TW, ty * RC_TH);
    }
}

__global__ __launch_bounds__(CB_TX * CB_TY) void kb_chroma_blur(const KFrame *__restrict__ frames,
                                                                const uint32_t *__restrict__ tiles,
                                                                uint32_t ntiles)
{
    __shared__ SmCb sm;

    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t d = tiles[t];
        const KFrame &f = frames[KTILE_FRAME(d)];
        const int ntx = (f.inter.wq + CB_TX - 1) / CB_TX, i = KTILE_IDX(d);

        __syncthreads();
        if (f.col.flags & KCOLOR_FLOAT_MATH)
            dev_chroma_blur<true>(sm, f, (i % ntx) * CB_TX, (i / ntx) * CB_TY);
        else
            dev_chroma_blur<false>(sm, f, (i % ntx) * CB_TX, (i / ntx) * CB_TY);
    }
}

__global__ void k_zero_u32(uint32_t *p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0u;
}

/* A frame's histogram to its host slot (pinned, coherent), leaving the
 * device copy zero for the next frame: one 16-byte word a thread. */
__global__ void k_hist_export(uint4 *__restrict__ hist, uint4 *__restrict__ host, uint32_t n4)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;

    if (i < n4) {
        host[i] = hist[i];
        hist[i] = make_uint4(0u, 0u, 0u, 0u);
    }
}

/* ---------------------------------------------------------- launchers */

static_assert(BL_TW == KG_SIMPLE_TW && BL_TH == KG_SIMPLE_TH, "tile geometry");
static_assert(RC_TW == KG_RCD_TW && RC_TH == KG_RCD_TH, "tile geometry");
static_assert(CB_TX == KG_CB_T && CB_TY == KG_CB_T, "tile geometry");
static_assert(CU_TW == KG_CU_TW && CU_TH == KG_CU_TH, "tile geometry");

hipError_t craw_launch_zero_u32(uint32_t *p, size_t n, hipStream_t s)
{
    KLAUNCH(k_zero_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t craw_launch_hist_export(uint32_t *hist, uint32_t *host, size_t n, hipStream_t s)
{
    const uint32_t n4 = (uint32_t)(n / 4);

    if (n % 4)
        return hipErrorInvalidValue;
    KLAUNCH(k_hist_export, dim3((n4 + 255) / 256), dim3(256), 0, s, (uint4 *)hist,
            (uint4 *)host, n4);
    return hipGetLastError();
}

hipError_t craw_launch_debayer(bool to_inter, const KFrame &f, hipStream_t s)
{
    if (f.method == 2) {
        return craw_launch_rcd(to_inter, f, s);
    } else {
        dim3 grid((f.W + BL_TW - 1) / BL_TW, (f.H + BL_TH - 1) / BL_TH);
        if (f.method == 0 && to_inter)
            KLAUNCH((k_debayer_simple<0, true>), grid, dim3(BL_NT), 0, s, f);
        else if (f.method == 0)
            KLAUNCH((k_debayer_simple<0, false>), grid, dim3(BL_NT), 0, s, f);
        else if (to_inter)
            KLAUNCH((k_debayer_simple<1, true>), grid, dim3(BL_NT), 0, s, f);
        else
            KLAUNCH((k_debayer_simple<1, false>), grid, dim3(BL_NT), 0, s, f);
    }
    return hipGetLastError();
}

hipError_t craw_launch_chroma_blur(const KFrame &f, hipStream_t s)
{
    dim3 grid((f.inter.wq + CB_TX - 1) / CB_TX, (f.inter.hq + CB_TY - 1) / CB_TY);
    KLAUNCH(k_chroma_blur, grid, dim3(CB_TX * CB_TY), 0, s, f);
    return hipGetLastError();
}

const void *craw_chroma_up_kernel(void)
{
    return (const void *)k_chroma_up_finish;
}

hipError_t craw_launch_chroma_up_finish(const KFrame &f, hipStream_t s)
{
    dim3 grid((f.W + CU_TW - 1) / CU_TW, (f.H + CU_TH - 1) / CU_TH);
    KLAUNCH(k_chroma_up_finish, grid, dim3(CU_NT), 0, s, f);
    return hipGetLastError();
}

hipError_t craw_launch_batch_debayer(int method, bool to_inter, const KFrame *frames,
                                     const uint32_t *tiles, uint32_t ntiles, int grid,
                                     hipStream_t s)
{
    if (!ntiles)
        return hipSuccess;
    if (method == 2) {
        if (to_inter)
            KLAUNCH(kb_debayer_rcd<true>, dim3(grid), dim3(RC_NT), 0, s, frames, tiles,
                               ntiles);
        else
            KLAUNCH(kb_debayer_rcd<false>, dim3(grid), dim3(RC_NT), 0, s, frames,
                               tiles, ntiles);
    } else if (method == 0) {
        if (to_inter)
            KLAUNCH((kb_debayer_simple<0, true>), dim3(grid), dim3(BL_NT), 0, s,
                               frames, tiles, ntiles);
        else
            KLAUNCH((kb_debayer_simple<0, false>), dim3(grid), dim3(BL_NT), 0, s,
                               frames, tiles, ntiles);
    } else {
        if (to_inter)
            KLAUNCH((kb_debayer_simple<1, true>), dim3(grid), dim3(BL_NT), 0, s,
                               frames, tiles, ntiles);
        else
            KLAUNCH((kb_debayer_simple<1, false>), dim3(grid), dim3(BL_NT), 0, s,
                               frames, tiles, ntiles);
    }
    return hipGetLastError();
}

hipError_t craw_launch_batch_chroma_blur(const KFrame *frames, const uint32_t *tiles,
                                         uint32_t ntiles, int grid, hipStream_t s)
{
    if (!ntiles)
        return hipSuccess;
    KLAUNCH(kb_chroma_blur, dim3(grid), dim3(CB_TX * CB_TY), 0, s, frames, tiles,
                       ntiles);
    return hipGetLastError();
}

hipError_t craw_launch_awb_hist(const KFrame &f, hipStream_t s)
{
    int ntx, n = craw_stage_tiles(KSTAGE_AWB, f.W, f.H, 0, 0, &ntx);

    KLAUNCH(k_awb_hist, dim3(n), dim3(AWB_NT), 0, s, f);
    return hipGetLastError();
}

hipError_t craw_launch_batch_awb_hist(const KFrame *frames, const uint32_t *tiles,
                                      uint32_t ntiles, int grid, hipStream_t s)
{
    if (!ntiles)
        return hipSuccess;
    KLAUNCH(kb_awb_hist, dim3(grid), dim3(AWB_NT), 0, s, frames, tiles, ntiles);
    return hipGetLastError();
}

