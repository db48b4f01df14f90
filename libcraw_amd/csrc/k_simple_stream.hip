/*
 * libcraw for MI355X: bilinear and MHC over the frame's interior, streamed
 * down columns.
 *
 * Same values as k_develop.hip's tiled kernel (lib/ref/ref_debayer.c:65-176
 * of cyanreg/libcraw), computed without LDS or barriers: a wave owns a
 * strip of 120 columns of the frame's interior, each lane a pair of
 * adjacent columns, and walks down it a row at a time, keeping the five
 * rows of prepped samples the filters reach in registers; the columns
 * either side come from the neighbouring lanes through wave-wide DPP
 * shifts. The output row trails the newest row by two. The interior is the
 * whole 64x16 tiles whose 2-pixel reach lies inside the frame, so nothing
 * here needs the mirror; the tiled kernel develops the ring of tiles
 * around it. The tails are the streamed RCD's: the chroma stage's inputs,
 * or the AWB block samples and the colour stage into the destination
 * (strips of 120 columns hold whole 8x8 blocks).
 */
#include "k_stream.h"

#define SS_NW 4
#define SS_NT (64 * SS_NW)
#define SS_PF 4                 /* rows loaded ahead */

static_assert(KG_SS_OUT % 8 == 0 && KG_SS_R % 8 == 0 && KG_SS_OUT / 2 + 2 <= 64,
              "strip geometry");

struct SsState {
    RsRow v[5];                 /* prepped rows vt - 4 .. vt */
    float acc[3];               /* block sums */
};

/* Sample (dx, dy) from column C of the output row v[2]. */
template <int C>
KDEV float ss_at(const SsState &s, int dx, int dy)
{
    const RsRow &r = s.v[2 + dy];

    switch (dx) {
    case -2: return rs_l2<C>(r);
    case -1: return rs_l1<C>(r);
    case 1: return rs_r1<C>(r);
    case 2: return rs_r2<C>(r);
    default: return r.c[C];
    }
}

/* Camera RGB of column C at site `site` (RGGB: 0 R, 1 and 2 G, 3 B) of the
 * output row, as the tiled kernel forms it. */
template <int METHOD, int C>
KDEV void ss_pixel(const SsState &s, int site, float &r, float &g, float &b)
{
#define V(dx, dy) ss_at<C>(s, dx, dy)
    if (METHOD == 0) {
        /* lib/ref/ref_debayer.c:65-96 */
        const float edge = 0.25f * (V(0, -1) + V(0, 1) + V(-1, 0) + V(1, 0));
        const float corner = 0.25f * (V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1));
        const float horiz = 0.5f * (V(-1, 0) + V(1, 0));
        const float vert = 0.5f * (V(0, -1) + V(0, 1));

        if (site == 0)      { r = V(0, 0); g = edge;    b = corner; }
        else if (site == 1) { r = horiz;   g = V(0, 0); b = vert; }
        else if (site == 2) { r = vert;    g = V(0, 0); b = horiz; }
        else                { r = corner;  g = edge;    b = V(0, 0); }
    } else {
        /* lib/ref/ref_debayer.c:98-176, the four MHC kernels over 8; only
         * the two a site uses are formed */
        const float c = V(0, 0);

        if (site == 0 || site == 3) {
            const float green = (4.0f * c + 2.0f * (V(0, -1) + V(0, 1) + V(-1, 0) + V(1, 0))
                                 - (V(0, -2) + V(0, 2) + V(-2, 0) + V(2, 0))) * 0.125f;
            const float diag = (6.0f * c + 2.0f * (V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1))
                                - 1.5f * (V(0, -2) + V(0, 2) + V(-2, 0) + V(2, 0))) * 0.125f;

            g = green;
            r = site == 0 ? c : diag;
            b = site == 0 ? diag : c;
        } else {
            const float ch = (5.0f * c + 4.0f * (V(-1, 0) + V(1, 0))
                              - (V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1) + V(-2, 0) + V(2, 0))
                              + 0.5f * (V(0, -2) + V(0, 2))) * 0.125f;
            const float cv = (5.0f * c + 4.0f * (V(0, -1) + V(0, 1))
                              - (V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1) + V(0, -2) + V(0, 2))
                              + 0.5f * (V(-2, 0) + V(2, 0))) * 0.125f;

            g = c;
            r = site == 1 ? ch : cv;
            b = site == 1 ? cv : ch;
        }
    }
#undef V
}

/* One step: the newest row vt (its two raw samples loaded ahead) enters;
 * row vt - 2 is developed. P is the parity of both (the strip starts on an
 * even row). */
template <int METHOD, bool INTER, int P>
KDEV void ss_step(SsState &s, const KFrame &f, int vt, const uint32_t raw[2], bool out_row,
                  bool out_cols, int X0)
{
    rs_shift(s.v);
    s.v[4].c[0] = kprep(raw[0], 2 * P, f.prep);
    s.v[4].c[1] = kprep(raw[1], 2 * P + 1, f.prep);
    if (out_row) {
        const int oy = vt - 2;
        float r[2], g[2], b[2];

        ss_pixel<METHOD, 0>(s, 2 * P, r[0], g[0], b[0]);
        ss_pixel<METHOD, 1>(s, 2 * P + 1, r[1], g[1], b[1]);
        if (INTER)
            rs_store_inter(s.acc, f, oy, out_cols, X0, r, g, b);
        else
            rs_finish(s.acc, f, oy, out_cols, X0, r, g, b);
    }
}

/* One strip: output columns [xs, xe), rows [ys, ye) of the interior. */
template <int METHOD, bool INTER>
KDEV void dev_simple_stream(const KFrame &f, int strip)
{
    const KRcdSplit sp = craw_simple_split(f.W, f.H);
    const int nsx = (sp.x_hi - sp.x_lo + KG_SS_OUT - 1) / KG_SS_OUT;
    const int nsy = (sp.y_hi - sp.y_lo + KG_SS_R - 1) / KG_SS_R;
    const int lane = threadIdx.x & 63;

    if (strip >= nsx * nsy)
        return;
    {
        const int sx = strip % nsx, sy = strip / nsx;
        const int xs = sp.x_lo + sx * KG_SS_OUT, ys = sp.y_lo + sy * KG_SS_R;
        const int xe = min(xs + KG_SS_OUT, sp.x_hi), ye = min(ys + KG_SS_R, sp.y_hi);
        const int X0 = xs - 2 + 2 * lane;
        const bool out_cols = X0 >= xs && X0 < xe;
        /* the pair's load column, kept inside the frame for lanes past the
         * strip's reach (their values are never used) */
        const int ld0 = X0 < f.W - 2 ? X0 : f.W - 2;
        const int nsteps = (ye - ys) + 4;
        SsState s;

#pragma unroll
        for (int k = 0; k < 5; k++)
            s.v[k].c[0] = s.v[k].c[1] = 0.0f;
        s.acc[0] = s.acc[1] = s.acc[2] = 0.0f;

        /* rows ys - 2 .. ye + 1 enter, SS_PF at a time: their loads are
         * issued together, then the steps run on them (a step's load would
         * otherwise wait out its own latency). ys is even, so step t's
         * rows have the parity of t; steps past the strip's end (the last
         * group's padding) load a row inside the frame and develop
         * nothing. */
#pragma unroll 1
        for (int t0 = 0; t0 < nsteps; t0 += SS_PF) {
            uint32_t raw[SS_PF][2];

#pragma unroll
            for (int k = 0; k < SS_PF; k++) {
                const int vt = min(ys - 2 + t0 + k, f.H - 1);
                const uint16_t *row = f.src + (size_t)vt * f.src_pitch;

                raw[k][0] = row[ld0];
                raw[k][1] = row[ld0 + 1];
            }
#pragma unroll
            for (int k = 0; k < SS_PF; k += 2) {
                const int t = t0 + k;

                ss_step<METHOD, INTER, 0>(s, f, ys - 2 + t, raw[k], t >= 4 && t < nsteps,
                                          out_cols, X0);
                ss_step<METHOD, INTER, 1>(s, f, ys - 1 + t, raw[k + 1],
                                          t + 1 >= 4 && t + 1 < nsteps, out_cols, X0);
            }
        }
    }
}

template <int METHOD, bool INTER>
__global__ __launch_bounds__(SS_NT) void k_simple_stream(KFrame f)
{
    dev_simple_stream<METHOD, INTER>(f, blockIdx.x * SS_NW + (threadIdx.x >> 6));
}

/* Strips of the frame's interior, in groups of SS_NW (one per wave). */
int craw_simple_stream_groups(int W, int H)
{
    const KRcdSplit sp = craw_simple_split(W, H);
    const int n = ((sp.x_hi - sp.x_lo + KG_SS_OUT - 1) / KG_SS_OUT) *
                  ((sp.y_hi - sp.y_lo + KG_SS_R - 1) / KG_SS_R);

    return (n + SS_NW - 1) / SS_NW;
}

hipError_t craw_launch_simple_stream(bool to_inter, const KFrame &f, hipStream_t s)
{
    const int n = craw_simple_stream_groups(f.W, f.H);

    if (!n)
        return hipSuccess;
    if (f.method == 0 && to_inter)
        KLAUNCH((k_simple_stream<0, true>), dim3(n), dim3(SS_NT), 0, s, f);
    else if (f.method == 0)
        KLAUNCH((k_simple_stream<0, false>), dim3(n), dim3(SS_NT), 0, s, f);
    else if (to_inter)
        KLAUNCH((k_simple_stream<1, true>), dim3(n), dim3(SS_NT), 0, s, f);
    else
        KLAUNCH((k_simple_stream<1, false>), dim3(n), dim3(SS_NT), 0, s, f);
    return hipGetLastError();
}
