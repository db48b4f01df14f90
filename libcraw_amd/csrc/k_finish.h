/*
 * libcraw for MI355X: the end of a tile's development, shared by the
 * kernels that finish a tile of camera RGB held in LDS (structure of
 * arrays, rgbs[channel][row][column]): the AWB block samples and the
 * colour stage into the destination (kfinish), or the chroma stage's
 * inputs (kstore_inter).
 */
#ifndef CRAW_K_FINISH_H
#define CRAW_K_FINISH_H

#include "k_common.h"
#include "kernels.h"

/* ------------------------------------------------------------ finish */

/* One channel of an AWB block summed row by row from 0 (lib/awb.c:249-256).
 * A whole 8x8 block is unrolled, so its 64 LDS reads are issued ahead of
 * the adds rather than one behind each; the adds keep their order. */
template <int TW, int TH>
KDEV float kblock_sum(const float (&ch)[TH][TW], int bx, int by, int nx, int ny)
{
    float sum = 0.0f;

    if (nx == 8 && ny == 8) {
        float v[8][8];

#pragma unroll
        for (int y = 0; y < 8; y++)
#pragma unroll
            for (int x = 0; x < 8; x++)
                v[y][x] = ch[by + y][bx + x];
#pragma unroll
        for (int y = 0; y < 8; y++)
#pragma unroll
            for (int x = 0; x < 8; x++)
                sum += v[y][x];
    } else {
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++)
                sum += ch[by + y][bx + x];
    }
    return sum;
}

/* AWB samples of the tile's camera RGB (SoA in LDS) and the colour stage
 * into the destination. Expects the LDS tile complete (after a barrier).
 * Blocks are summed row by row as lib/awb.c:249-256 sums them, one lane of
 * the first wave per block and channel (a lane per block when a tile has
 * more than 21 blocks); each block's packed sample goes to the frame's
 * sample array for the histogram stage. */
template <int TW, int TH, int NT>
KDEV void kfinish(const float (*rgbs)[TH][TW], int x0, int y0, int W, int H,
                  const KColor &col, const KOut &out)
{
    constexpr int NBX = TW / 8, NBY = TH / 8, NB = NBX * NBY;
    const int tid = threadIdx.x;

    if constexpr (3 * NB > 64) {
        /* more blocks than a wave has lanes for three each: a lane per
         * block sums its three channels, each in the same order */
        static_assert(NB <= NT, "a lane for every block");
        if (out.hist && tid < NB) {
            const int gx0 = x0 + (tid % NBX) * 8, gy0 = y0 + (tid / NBX) * 8;
            const int gx1 = gx0 + 8 < W ? gx0 + 8 : W, gy1 = gy0 + 8 < H ? gy0 + 8 : H;
            float acc[3] = { 0.0f, 0.0f, 0.0f };
            CRAWAwbSample smp = { 0u, 0u, 0u };

            if (gx0 < W && gy0 < H) {
                for (int c = 0; c < 3; c++)
                    acc[c] = kblock_sum<TW, TH>(rgbs[c], gx0 - x0, gy0 - y0, gx1 - gx0, gy1 - gy0);
                const bool valid = craw_awb_sample_log(acc, (float)((gy1 - gy0) * (gx1 - gx0)),
                                                       out.awb_wb0, out.awb_wb2, out.float_log,
                                                       &smp);
                out.samples[(size_t)(gy0 / 8) * ((W + 7) / 8) + gx0 / 8] =
                    valid ? make_uint2(smp.bin | (smp.weight << 14), smp.bright)
                          : make_uint2(KAWB_NONE, 0u);
            }
        }
    } else if (out.hist && tid < 64) {
        /* lane c * NB + b sums channel c of block b (64 sequential adds);
         * the channel-0 lane gathers the other two and bins the block */
        const int c = tid / NB, b = tid % NB;
        const int gx0 = x0 + (b % NBX) * 8, gy0 = y0 + (b / NBX) * 8;
        const int gx1 = gx0 + 8 < W ? gx0 + 8 : W, gy1 = gy0 + 8 < H ? gy0 + 8 : H;
        const bool inside = c < 3 && gx0 < W && gy0 < H;
        float sum = 0.0f, acc[3];
        CRAWAwbSample smp = { 0u, 0u, 0u };
        bool valid = false;

        if (inside)
            sum = kblock_sum<TW, TH>(rgbs[c], gx0 - x0, gy0 - y0, gx1 - gx0, gy1 - gy0);
        acc[0] = sum;
        acc[1] = __shfl(sum, tid + NB, 64);
        acc[2] = __shfl(sum, tid + 2 * NB, 64);
        if (inside && c == 0) {
            valid = craw_awb_sample_log(acc, (float)((gy1 - gy0) * (gx1 - gx0)), out.awb_wb0,
                                        out.awb_wb2, out.float_log, &smp);
            out.samples[(size_t)(gy0 / 8) * ((W + 7) / 8) + gx0 / 8] =
                valid ? make_uint2(smp.bin | (smp.weight << 14), smp.bright)
                      : make_uint2(KAWB_NONE, 0u);
        }
    }

    /* With the histogram the first wave's block sums are a serial chain,
     * so the other waves take its share of the colour stage: the tile
     * ends when the slower of the two is done rather than after both. */
    const int first = out.hist ? 64 : 0;

    if (tid < first)
        return;
#pragma unroll 1
    for (int p = tid - first; p < TW * TH; p += NT - first) {
        const int px = p % TW, py = p / TW, x = x0 + px, y = y0 + py;
        float o[3];

        if (x >= W || y >= H)
            continue;
        kcolor(col, rgbs[0][py][px], rgbs[1][py][px], rgbs[2][py][px], o);
        kstore_rgba(out, x, y, o);
    }
}

/* Camera RGB of the tile to the chroma stage's inputs: the three planes
 * (to it.rgb_out) and, with it.means, the 4x4 block means of (R - G,
 * B - G, G), summed row by row over the pixels inside the frame
 * (lib/ref/ref_chroma.c:59-82). */
template <int TW, int TH, int NT>
KDEV void kstore_inter(const float (*rgbs)[TH][TW], int x0, int y0, int W, int H,
                       const KInter &it)
{
    const int tid = threadIdx.x;
    const size_t plane = (size_t)W * H;

#pragma unroll
    for (int k = 0; k < TW * TH / NT; k++) {
        int p = tid + k * NT, px = p % TW, py = p / TW;
        int x = x0 + px, y = y0 + py;

        if (x >= W || y >= H)
            continue;
        it.rgb_out[(size_t)y * W + x] = rgbs[0][py][px];
        it.rgb_out[plane + (size_t)y * W + x] = rgbs[1][py][px];
        it.rgb_out[2 * plane + (size_t)y * W + x] = rgbs[2][py][px];
    }

    if (!it.means)
        return;
    for (int b = tid; b < (TW / 4) * (TH / 4); b += NT) {
        int qx = x0 / 4 + b % (TW / 4), qy = y0 / 4 + b / (TW / 4);
        float sum[3] = { 0.0f, 0.0f, 0.0f }, n = 0.0f;

        if (4 * qx >= W || 4 * qy >= H)
            continue;
        for (int j = 0; j < 4; j++) {
            for (int i = 0; i < 4; i++) {
                int x = 4 * qx + i, y = 4 * qy + j;
                float r, g, bl;

                if (x >= W || y >= H)
                    continue;
                r = rgbs[0][y - y0][x - x0];
                g = rgbs[1][y - y0][x - x0];
                bl = rgbs[2][y - y0][x - x0];
                sum[0] += r - g;
                sum[1] += bl - g;
                sum[2] += g;
                n += 1.0f;
            }
        }
        {
            size_t q = (size_t)qy * it.wq + qx, qplane = (size_t)it.wq * it.hq;
            it.down[q] = sum[0] / n;
            it.down[qplane + q] = sum[1] / n;
            it.down[2 * qplane + q] = sum[2] / n;
        }
    }
}

#endif
