/*
 * libcraw for MI355X: what the streamed debayer kernels share.
 *
 * A wave owns a strip of the frame, each lane a pair of adjacent columns
 * (one green and one red or blue site in every row), and walks down the
 * strip a row at a time; neighbouring columns come from the neighbouring
 * lanes through wave-wide DPP shifts (row_shr/shl would stop at 16
 * lanes). The tails take a finished row of camera RGB: into the chroma
 * stage's inputs (rs_store_inter), or through the AWB block samples and
 * the colour stage to the destination (rs_finish).
 */
#ifndef CRAW_K_STREAM_H
#define CRAW_K_STREAM_H

#include "k_common.h"
#include "kernels.h"

#define DPP_WAVE_SHR1 0x138
#define DPP_WAVE_SHL1 0x130

/* The value of the lane to the left / right (0 past the wave's ends). */
KDEV float rs_shr(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 DPP_WAVE_SHR1, 0xf, 0xf, true));
}

KDEV float rs_shl(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 DPP_WAVE_SHL1, 0xf, 0xf, true));
}

/* A row of a stage: the lane's two columns. */
struct RsRow {
    float c[2];
};

/* Shifts a window of rows by one: row k + 1 becomes row k. */
template <int N>
KDEV void rs_shift(RsRow (&a)[N])
{
#pragma unroll
    for (int k = 0; k + 1 < N; k++)
        a[k] = a[k + 1];
}

/* Column C's left / right neighbour of row r, for the lane's column C. */
template <int C> KDEV float rs_l1(const RsRow &r) { return C == 0 ? rs_shr(r.c[1]) : r.c[0]; }
template <int C> KDEV float rs_r1(const RsRow &r) { return C == 0 ? r.c[1] : rs_shl(r.c[0]); }
template <int C> KDEV float rs_l2(const RsRow &r) { return rs_shr(r.c[C]); }
template <int C> KDEV float rs_r2(const RsRow &r) { return rs_shl(r.c[C]); }

/* The output of row oy without the chroma stage: each 8x8 block's sums in
 * raster order (the block's first lane, column X0 % 8 == 0, adds its two
 * columns and then its three right neighbours'; lib/awb.c:249-256), its
 * sample on the block's last row, and the colour stage into the
 * destination (lib/ref/ref_color.c:10-76). */
template <bool GAMUT>
KDEV void rs_finish(float (&acc)[3], const KFrame &f, int oy, bool out_cols, int X0,
                    const float r[2], const float g[2], const float b[2])
{
    if (f.out.hist) {
        const bool first_row = (oy & 7) == 0;
        const float *chan[3] = { r, g, b };

#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float n10 = rs_shl(chan[c][0]), n11 = rs_shl(chan[c][1]);
            const float n20 = rs_shl(n10), n21 = rs_shl(n11);
            const float n30 = rs_shl(n20), n31 = rs_shl(n21);
            float a = (first_row ? 0.0f : acc[c]) + chan[c][0];

            a = a + chan[c][1];
            a = a + n10;
            a = a + n11;
            a = a + n20;
            a = a + n21;
            a = a + n30;
            a = a + n31;
            acc[c] = a;
        }
        if ((oy & 7) == 7 && out_cols && (X0 & 7) == 0) {
            CRAWAwbSample smp;
            const bool valid = craw_awb_sample_log(acc, 64.0f, f.out.awb_wb0, f.out.awb_wb2,
                                                   f.out.float_log, &smp);

            f.out.samples[(size_t)(oy / 8) * ((f.W + 7) / 8) + X0 / 8] =
                valid ? make_uint2(smp.bin | (smp.weight << 14), smp.bright)
                      : make_uint2(KAWB_NONE, 0u);
        }
    }
    if (out_cols) {
#pragma unroll 1
        for (int c = 0; c < 2; c++) {
            float o[3];

            kcolor<GAMUT>(f.col, r[c], g[c], b[c], o);
            kstore_rgba(f.out, X0 + c, oy, o);
        }
    }
}

/* The chroma stage's inputs from row oy: the three camera RGB planes and
 * the 4x4 block sums of R - G, B - G, G in raster order (the block's first
 * lane, column X0 % 4 == 0, adds its two columns and then its right
 * neighbour's; lib/ref/ref_chroma.c:59-82), stored as means on the
 * block's last row with f.inter.means (without it the false colour
 * stage, not the chroma stage, reads the planes next). */
KDEV void rs_store_inter(float (&acc)[3], const KFrame &f, int oy, bool out_cols, int X0,
                         const float r[2], const float g[2], const float b[2])
{
    const KInter &it = f.inter;
    const size_t plane = (size_t)f.W * f.H;

    if (out_cols) {
        const size_t o = (size_t)oy * f.W + X0;
        *(float2 *)&it.rgb_out[o] = make_float2(r[0], r[1]);
        *(float2 *)&it.rgb_out[plane + o] = make_float2(g[0], g[1]);
        *(float2 *)&it.rgb_out[2 * plane + o] = make_float2(b[0], b[1]);
    }
    {
        const float d0[2] = { r[0] - g[0], r[1] - g[1] };
        const float d1[2] = { b[0] - g[0], b[1] - g[1] };
        const float n00 = rs_shl(d0[0]), n01 = rs_shl(d0[1]);
        const float n10 = rs_shl(d1[0]), n11 = rs_shl(d1[1]);
        const float n20 = rs_shl(g[0]), n21 = rs_shl(g[1]);
        const bool first_row = (oy & 3) == 0;

        acc[0] = (first_row ? 0.0f : acc[0]) + d0[0];
        acc[0] = acc[0] + d0[1];
        acc[0] = acc[0] + n00;
        acc[0] = acc[0] + n01;
        acc[1] = (first_row ? 0.0f : acc[1]) + d1[0];
        acc[1] = acc[1] + d1[1];
        acc[1] = acc[1] + n10;
        acc[1] = acc[1] + n11;
        acc[2] = (first_row ? 0.0f : acc[2]) + g[0];
        acc[2] = acc[2] + g[1];
        acc[2] = acc[2] + n20;
        acc[2] = acc[2] + n21;
        if ((oy & 3) == 3 && out_cols && (X0 & 3) == 0 && it.means) {
            const size_t q = (size_t)(oy / 4) * it.wq + X0 / 4;
            const size_t qplane = (size_t)it.wq * it.hq;

            it.down[q] = acc[0] / 16.0f;
            it.down[qplane + q] = acc[1] / 16.0f;
            it.down[2 * qplane + q] = acc[2] / 16.0f;
        }
    }
}

#endif
