/*
 * libcraw for MI355X: RCD over the frame's interior, streamed down columns.
 *
 * Same values as k_develop.hip's tiled RCD (lib/ref/ref_debayer.c:192-395
 * of cyanreg/libcraw), computed without LDS: a wave owns a strip of the
 * frame's interior, each lane a pair of adjacent columns (one green and
 * one red or blue site in every row), and walks down the strip a row at a
 * time. Every stage keeps the rows its consumers reach in registers, so
 * each value is computed once; neighbouring columns come from the
 * neighbouring lanes through wave-wide DPP shifts (row_shr/shl would stop
 * at 16 lanes). Because a pair always holds one site of each kind, the
 * site-dependent stages (green, colour estimate, output) are uniform
 * branches on the row's parity.
 *
 * Per step the stages run at fixed lags behind the newest row of V:
 * low pass 1, gradients 2, direction weights and green 3, colour estimate
 * 4, output 5. A strip's V reach is 5 columns and 5 rows past its outputs,
 * all inside the frame (the interior is chosen so), so nothing here needs
 * the mirror; the tiled kernel develops the ring of tiles around it.
 *
 * Output: the chroma stage's inputs (the three camera RGB planes and the
 * 4x4 block means of R - G, B - G, G, summed in raster order), or without
 * the chroma stage the finished frame: the AWB block samples (8x8 sums in
 * raster order) and the colour stage into the destination. Strips are 116
 * columns (29 4x4 blocks) into the chroma stage, 112 (14 8x8 blocks) to
 * the output, so no block straddles two strips.
 */
#include "k_stream.h"
#include "k_rcd_tile.h"

#define RS_NW 4
#define RS_NT (64 * RS_NW)
#define RS_EPS 1e-5f
/* V rows a lane loads ahead of the step that takes them in: a wave's walk
 * down its strip would otherwise wait on each row's load (0.106 -> 0.096
 * ms at 5760x3240; 2 rows ahead 0.098, 8 rows 0.096) */
#define RS_PF 4

static_assert(KG_RS_OUT == 116 && KG_RS_OUT_F == 112 && KG_RS_R % 8 == 0, "strip geometry");

/* n / d, the float division's value: through kdiv_rn where the frame's
 * constants bound the operands to its range (a uniform branch). */
KDEV float rs_div(float n, float d, bool kdiv)
{
    return kdiv ? kdiv_rn(n, d) : n / d;
}

/* lib/ref/ref_debayer.c:195-202 */
KDEV float rs_ratio(float a, float b, bool kdiv)
{
    a = a > 0.0f ? a : 0.0f;
    b = b > 0.0f ? b : 0.0f;
    return rs_div(a - b, a + b + RS_EPS, kdiv);
}

struct RsState {
    RsRow v[7];                 /* V rows vt - 6 .. vt */
    RsRow lp[5];                /* LP rows vt - 5 .. vt - 1 */
    RsRow g0[3], g1[3], g2[3], g3[3];   /* gradient rows vt - 4 .. vt - 2 */
    RsRow wv[3];                /* rows vt - 5 .. vt - 3 */
    RsRow wp[2];                /* rows vt - 4 .. vt - 3 */
    RsRow G[4];                 /* rows vt - 6 .. vt - 3 */
    RsRow C[3];                 /* rows vt - 6 .. vt - 4 */
    float acc[3];               /* 4x4 block sums */
    uint32_t raw[RS_PF];        /* raw sample pairs of V rows vt .. vt + RS_PF - 1 */
};

/* Gradients of column C at the centre row v[4] (rows v[2] .. v[6]). */
template <int C>
KDEV void rs_gradients(RsState &s)
{
    const RsRow *v = s.v;
    const float c2 = 2.0f * v[4].c[C];

    s.g0[2].c[C] = fabsf(v[3].c[C] - v[5].c[C]) + fabsf(c2 - v[2].c[C] - v[6].c[C]);
    s.g1[2].c[C] = fabsf(rs_l1<C>(v[4]) - rs_r1<C>(v[4])) +
                   fabsf(c2 - rs_l2<C>(v[4]) - rs_r2<C>(v[4]));
    s.g2[2].c[C] = fabsf(rs_l1<C>(v[5]) - rs_r1<C>(v[3])) +
                   fabsf(c2 - rs_l2<C>(v[6]) - rs_r2<C>(v[2]));
    s.g3[2].c[C] = fabsf(rs_l1<C>(v[3]) - rs_r1<C>(v[5])) +
                   fabsf(c2 - rs_l2<C>(v[2]) - rs_r2<C>(v[6]));
}

/* 3x3 sum of gradient rows g[0..2] around column C, in raster order. */
template <int C>
KDEV float rs_box(const RsRow (&g)[3])
{
    float a = 0.0f;

#pragma unroll
    for (int dy = 0; dy < 3; dy++) {
        a += rs_l1<C>(g[dy]);
        a += g[dy].c[C];
        a += rs_r1<C>(g[dy]);
    }
    return a;
}

template <int C>
KDEV void rs_weights(RsState &s, bool kdiv)
{
    const float a0 = rs_box<C>(s.g0), a1 = rs_box<C>(s.g1);
    const float b0 = rs_box<C>(s.g2), b1 = rs_box<C>(s.g3);
    const float vsum = a0 * (1.0f / 9.0f), hsum = a1 * (1.0f / 9.0f);
    const float ps = b0 * (1.0f / 9.0f), qs = b1 * (1.0f / 9.0f);

    s.wv[2].c[C] = rs_div(hsum, hsum + vsum + RS_EPS, kdiv);
    s.wp[1].c[C] = rs_div(qs, ps + qs + RS_EPS, kdiv);
}

/* Green of the red/blue column N of the centre row v[3]. */
template <int N>
KDEV float rs_green(const RsState &s, bool kdiv)
{
    const float l00 = s.lp[2].c[N], lu = s.lp[0].c[N], ld = s.lp[4].c[N];
    const float ll = rs_l2<N>(s.lp[2]), lr = rs_r2<N>(s.lp[2]);
    const float gv = 0.5f * (s.v[2].c[N] * (1.0f + rs_ratio(l00, lu, kdiv)) +
                             s.v[4].c[N] * (1.0f + rs_ratio(l00, ld, kdiv)));
    const float gh = 0.5f * (rs_l1<N>(s.v[3]) * (1.0f + rs_ratio(l00, ll, kdiv)) +
                             rs_r1<N>(s.v[3]) * (1.0f + rs_ratio(l00, lr, kdiv)));
    const float w = s.wv[2].c[N];

    return w * gv + (1.0f - w) * gh;
}

/* Colour estimate of the red/blue column N of the centre row v[2]. */
template <int N>
KDEV float rs_colour(const RsState &s)
{
    const float g00 = s.G[2].c[N];
    const float dsw = rs_l1<N>(s.v[3]) - rs_l1<N>(s.G[3]);
    const float dne = rs_r1<N>(s.v[1]) - rs_r1<N>(s.G[1]);
    const float dnw = rs_l1<N>(s.v[1]) - rs_l1<N>(s.G[1]);
    const float dse = rs_r1<N>(s.v[3]) - rs_r1<N>(s.G[3]);
    const float w = s.wp[0].c[N];

    return w * (g00 + 0.5f * (dsw + dne)) + (1.0f - w) * (g00 + 0.5f * (dnw + dse));
}

/* One step: P is the parity of the step (the newest V row is odd for
 * P = 0, the strip starting on an even row). L: the stages that step
 * computes, from its first ten steps (whose rows of the later stages no
 * output reaches): 0 the V row, 1 and the low pass, 2 and the gradients,
 * 3 and the weights and green, 4 and the colour estimate, 5 and the
 * output. */
/* The lane's two samples of row y (ld0 even) as one 32-bit word: one load
 * where the mosaic's rows are 4-byte aligned (a uniform branch). */
KDEV uint32_t rs_load_pair(const KFrame &f, int y, int ld0)
{
    const uint16_t *row = f.src + (size_t)y * f.src_pitch;

    if (((uintptr_t)f.src & 3u) == 0 && (f.src_pitch & 1) == 0)
        return *(const uint32_t *)(row + ld0);
    return (uint32_t)row[ld0] | (uint32_t)row[ld0 + 1] << 16;
}

template <bool INTER, bool GAMUT, int P, int L = 5>
KDEV void rs_step(RsState &s, const KFrame &f, int vt, int ld0, bool out_row, bool out_cols,
                  int X0)
{
    /* V row vt from the pair loaded RS_PF steps before; row vt + RS_PF
     * loaded now (clamped to the frame: past its last row the values are
     * never used) */
    rs_shift(s.v);
    {
        const uint32_t pair = s.raw[0];
        const int sb = (vt & 1) * 2;

#pragma unroll
        for (int k = 0; k < RS_PF - 1; k++)
            s.raw[k] = s.raw[k + 1];
        s.raw[RS_PF - 1] = rs_load_pair(f, vt + RS_PF < f.H ? vt + RS_PF : f.H - 1, ld0);
        s.v[6].c[0] = kprep(pair & 0xffffu, sb, f.prep);
        s.v[6].c[1] = kprep(pair >> 16, sb + 1, f.prep);
    }

    /* low pass, row vt - 1 */
    rs_shift(s.lp);
#pragma unroll
    for (int c = 0; c < 2 && L >= 1; c++) {
        const RsRow *v = s.v;
        const float w_ = c == 0 ? rs_shr(v[5].c[1]) : v[5].c[0];
        const float e_ = c == 0 ? v[5].c[1] : rs_shl(v[5].c[0]);
        const float nw = c == 0 ? rs_shr(v[4].c[1]) : v[4].c[0];
        const float ne = c == 0 ? v[4].c[1] : rs_shl(v[4].c[0]);
        const float sw = c == 0 ? rs_shr(v[6].c[1]) : v[6].c[0];
        const float se = c == 0 ? v[6].c[1] : rs_shl(v[6].c[0]);

        s.lp[4].c[c] = (4.0f * v[5].c[c] + 2.0f * (v[4].c[c] + v[6].c[c] + w_ + e_) + nw + ne +
                        sw + se) * (1.0f / 16.0f);
    }

    /* gradients, row vt - 2 */
    rs_shift(s.g0);
    rs_shift(s.g1);
    rs_shift(s.g2);
    rs_shift(s.g3);
    if constexpr (L >= 2) {
        rs_gradients<0>(s);
        rs_gradients<1>(s);
    }

    /* direction weights, row vt - 3 */
    rs_shift(s.wv);
    rs_shift(s.wp);
    if constexpr (L >= 3) {
        rs_weights<0>(s, f.prep.kdiv != 0);
        rs_weights<1>(s, f.prep.kdiv != 0);
    }

    /* green, row vt - 3: the red/blue site is column P there */
    rs_shift(s.G);
    if constexpr (L >= 3) {
        s.G[3].c[P] = rs_green<P>(s, f.prep.kdiv != 0);
        s.G[3].c[1 - P] = s.v[3].c[1 - P];
    }

    /* colour estimate, row vt - 4: red/blue column 1 - P */
    rs_shift(s.C);
    if constexpr (L >= 4) {
        s.C[2].c[1 - P] = rs_colour<1 - P>(s);
        s.C[2].c[P] = 0.0f;
    }

    /* output, row vt - 5: green column 1 - P, red/blue column P */
    if (L >= 5 && out_row) {
        constexpr int CG = 1 - P, CN = P;
        const int oy = vt - 5;
        float r[2], g[2], b[2];

        {
            const float c = s.C[1].c[CN], gg = s.G[1].c[CN], vv = s.v[1].c[CN];
            g[CN] = gg;
            r[CN] = P == 0 ? vv : c;
            b[CN] = P == 0 ? c : vv;
        }
        {
            const float g00 = s.G[1].c[CG], w = s.wv[0].c[CG];
            const float dh = 0.5f * ((rs_l1<CG>(s.v[1]) - rs_l1<CG>(s.G[1])) +
                                     (rs_r1<CG>(s.v[1]) - rs_r1<CG>(s.G[1])));
            const float ev = 0.5f * ((s.C[0].c[CG] - s.G[0].c[CG]) + (s.C[2].c[CG] - s.G[2].c[CG]));
            const float ch = w * (g00 + ev) + (1.0f - w) * (g00 + dh);
            const float dv = 0.5f * ((s.v[0].c[CG] - s.G[0].c[CG]) + (s.v[2].c[CG] - s.G[2].c[CG]));
            const float eh = 0.5f * ((rs_l1<CG>(s.C[1]) - rs_l1<CG>(s.G[1])) +
                                     (rs_r1<CG>(s.C[1]) - rs_r1<CG>(s.G[1])));
            const float cv = w * (g00 + dv) + (1.0f - w) * (g00 + eh);

            g[CG] = s.v[1].c[CG];
            r[CG] = P == 0 ? ch : cv;
            b[CG] = P == 0 ? cv : ch;
        }

        if (INTER)
            rs_store_inter(s.acc, f, oy, out_cols, X0, r, g, b);
        else
            rs_finish<GAMUT>(s.acc, f, oy, out_cols, X0, r, g, b);
    }
}

/* One strip: output columns [xs, xe), rows [ys, ye) of the interior.
 * GAMUT (to the output only): the colour stage may have the gamut
 * compression. Its code holds registers the streamed kernel needs
 * elsewhere, and the planner sends frames with the compression through
 * the planes instead wherever it has them, so the kernel comes both
 * ways. */
template <bool INTER, bool GAMUT>
KDEV void dev_rcd_stream(const KFrame &f, int strip)
{
    constexpr int OUTW = INTER ? KG_RS_OUT : KG_RS_OUT_F;
    const KRcdSplit sp = craw_rcd_split(f.W, f.H);
    const int nsx = (sp.x_hi - sp.x_lo + OUTW - 1) / OUTW;
    const int nsy = (sp.y_hi - sp.y_lo + KG_RS_R - 1) / KG_RS_R;
    const int lane = threadIdx.x & 63;

    if (strip >= nsx * nsy)
        return;
    {
        const int sx = strip % nsx, sy = strip / nsx;
        const int xs = sp.x_lo + sx * OUTW, ys = sp.y_lo + sy * KG_RS_R;
        const int xe = min(xs + OUTW, sp.x_hi), ye = min(ys + KG_RS_R, sp.y_hi);
        const int X0 = xs - 6 + 2 * lane;
        const bool out_cols = X0 >= xs && X0 < xe;
        /* the pair's load column, kept inside the frame for lanes past the
         * strip's reach (their values are never used) */
        const int ld0 = X0 < f.W - 2 ? X0 : f.W - 2;
        const int nsteps = (ye - ys) + 10;
        RsState s;

#pragma unroll
        for (int k = 0; k < 7; k++)
            s.v[k].c[0] = s.v[k].c[1] = 0.0f;
#pragma unroll
        for (int k = 0; k < 5; k++)
            s.lp[k].c[0] = s.lp[k].c[1] = 0.0f;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            s.g0[k].c[0] = s.g0[k].c[1] = s.g1[k].c[0] = s.g1[k].c[1] = 0.0f;
            s.g2[k].c[0] = s.g2[k].c[1] = s.g3[k].c[0] = s.g3[k].c[1] = 0.0f;
            s.wv[k].c[0] = s.wv[k].c[1] = s.C[k].c[0] = s.C[k].c[1] = 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 4; k++)
            s.G[k].c[0] = s.G[k].c[1] = 0.0f;
        s.wp[0].c[0] = s.wp[0].c[1] = s.wp[1].c[0] = s.wp[1].c[1] = 0.0f;
        s.acc[0] = s.acc[1] = s.acc[2] = 0.0f;
#pragma unroll
        for (int k = 0; k < RS_PF; k++)
            s.raw[k] = rs_load_pair(f, ys - 5 + k, ld0);

        /* ys is even and nsteps is even: steps alternate parity 0, 1. The
         * first ten steps fill the stages: V rows from ys - 5, the low pass
         * from row ys - 4 (step 2), the gradients from ys - 3 (step 4), the
         * weights and green from ys - 2 (step 6), the colour estimate from
         * ys - 1 (step 8): the first rows output row ys reaches through
         * them. A stage's rows before those are left as they were. */
        rs_step<INTER, GAMUT, 0, 0>(s, f, ys - 5, ld0, false, out_cols, X0);
        rs_step<INTER, GAMUT, 1, 0>(s, f, ys - 4, ld0, false, out_cols, X0);
        rs_step<INTER, GAMUT, 0, 1>(s, f, ys - 3, ld0, false, out_cols, X0);
        rs_step<INTER, GAMUT, 1, 1>(s, f, ys - 2, ld0, false, out_cols, X0);
        rs_step<INTER, GAMUT, 0, 2>(s, f, ys - 1, ld0, false, out_cols, X0);
        rs_step<INTER, GAMUT, 1, 2>(s, f, ys, ld0, false, out_cols, X0);
        rs_step<INTER, GAMUT, 0, 3>(s, f, ys + 1, ld0, false, out_cols, X0);
        rs_step<INTER, GAMUT, 1, 3>(s, f, ys + 2, ld0, false, out_cols, X0);
        rs_step<INTER, GAMUT, 0, 4>(s, f, ys + 3, ld0, false, out_cols, X0);
        rs_step<INTER, GAMUT, 1, 4>(s, f, ys + 4, ld0, false, out_cols, X0);
#pragma unroll 1
        for (int t = 10; t < nsteps; t += 2) {
            rs_step<INTER, GAMUT, 0>(s, f, ys - 5 + t, ld0, true, out_cols, X0);
            rs_step<INTER, GAMUT, 1>(s, f, ys - 4 + t, ld0, true, out_cols, X0);
        }
    }
}

/* The frame's RCD in one launch: workgroups [0, nstream) stream the
 * interior's strips, the rest develop the ring of tiles around it
 * (k_rcd_tile.h). On a 5.8K frame the 1,013 groups of strips hold about
 * four fifths of the GPU's workgroup slots for the whole launch, one round
 * of them, and are dispatched first; the 762 ring tiles then run in the
 * slots beside them instead of in a launch of their own before them (0.120
 * -> 0.105 ms). The ring's LDS is allocated to every workgroup: 32 KB,
 * which fits five a CU, more than the strips' registers allow. Four waves a SIMD for
 * the fused colour tail without the compression (128 VGPRs; the
 * compiler's own choice is 130, three waves). */
static_assert(RS_NT == RC_NT, "one workgroup size for strips and ring tiles");

template <bool INTER, bool GAMUT>
__global__ __launch_bounds__(RS_NT) __attribute__((amdgpu_waves_per_eu(GAMUT ? 1 : 4)))
void k_rcd(KFrame f, int nstream)
{
    __shared__ SmRcd sm;

    if ((int)blockIdx.x < nstream) {
        dev_rcd_stream<INTER, GAMUT>(f, blockIdx.x * RS_NW + (threadIdx.x >> 6));
    } else {
        int tx, ty;

        craw_ring_tile(craw_rcd_split(f.W, f.H), (int)blockIdx.x - nstream, &tx, &ty);
        dev_debayer_rcd_tile<INTER>(sm, f, tx * RC_TW, ty * RC_TH);
    }
}

/* Persistent form: work-list entries are groups of RS_NW strips. */
template <bool INTER, bool GAMUT>
__global__ __launch_bounds__(RS_NT) void kb_rcd_stream(const KFrame *__restrict__ frames,
                                                       const uint32_t *__restrict__ tiles,
                                                       uint32_t ntiles)
{
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t d = tiles[t];
        dev_rcd_stream<INTER, GAMUT>(frames[KTILE_FRAME(d)],
                                     KTILE_IDX(d) * RS_NW + (threadIdx.x >> 6));
    }
}

hipError_t craw_launch_rcd(bool to_inter, const KFrame &f, hipStream_t s)
{
    int ntx, nring = craw_stage_tiles(KSTAGE_RCD_RING, f.W, f.H, 2, 0, &ntx);
    int nstream = craw_stage_tiles(to_inter ? KSTAGE_RCD_STREAM : KSTAGE_RCD_STREAM_F, f.W,
                                   f.H, 2, 0, &ntx);
    const dim3 grid((unsigned)(nstream + nring));

    if (!nstream && !nring)
        return hipSuccess;
    if (to_inter)
        KLAUNCH((k_rcd<true, false>), grid, dim3(RS_NT), 0, s, f, nstream);
    else if (f.col.flags & KCOLOR_GAMUT_COMPRESS)
        KLAUNCH((k_rcd<false, true>), grid, dim3(RS_NT), 0, s, f, nstream);
    else
        KLAUNCH((k_rcd<false, false>), grid, dim3(RS_NT), 0, s, f, nstream);
    return hipGetLastError();
}

hipError_t craw_launch_batch_rcd_stream(bool to_inter, bool gamut, const KFrame *frames,
                                        const uint32_t *tiles, uint32_t ntiles, int grid,
                                        hipStream_t s)
{
    if (!ntiles)
        return hipSuccess;
    if (to_inter)
        KLAUNCH((kb_rcd_stream<true, false>), dim3(grid), dim3(RS_NT), 0, s, frames, tiles,
                ntiles);
    else if (gamut)
        KLAUNCH((kb_rcd_stream<false, true>), dim3(grid), dim3(RS_NT), 0, s, frames, tiles,
                ntiles);
    else
        KLAUNCH((kb_rcd_stream<false, false>), dim3(grid), dim3(RS_NT), 0, s, frames, tiles,
                ntiles);
    return hipGetLastError();
}
