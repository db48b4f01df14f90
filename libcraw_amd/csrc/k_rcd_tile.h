/*
 * libcraw for MI355X: the tiled RCD (lib/ref/ref_debayer.c:192-395 of
 * cyanreg/libcraw), shared by the kernels that develop RCD tiles: the
 * ring of tiles around the streamed interior (k_rcd_stream.hip) and the
 * batched form's ring launch (k_develop.hip).
 *
 * A workgroup develops one 32x16 tile. The prepped mosaic and each
 * intermediate of RCD's passes are staged in LDS over the tile plus the
 * apron the later stencils reach; out-of-frame positions of an apron hold
 * the value at the mirrored position, and every stage evaluates its
 * stencil at the mirrored position's own neighbourhood (the reference's
 * fetch through the mirror, lib/ref/ref_debayer.c:20-40).
 */
#ifndef CRAW_K_RCD_TILE_H
#define CRAW_K_RCD_TILE_H

#include "k_common.h"
#include "kernels.h"
#include "k_finish.h"

#define RC_TW 32
#define RC_TH 16
#define RC_NT 256
#define RCD_EPS 1e-5f

/* One LDS array of a tile stage with apron A: indexed by frame
 * coordinates relative to the tile origin. */
template <int A>
struct RcStage {
    static constexpr int W = RC_TW + 2 * A, H = RC_TH + 2 * A, N = W * H;
};

/* n / d, the float division's value: through kdiv_rn where the frame's
 * constants bound the operands to its range (a uniform branch). */
KDEV float rcd_div(float n, float d, bool kdiv)
{
    return kdiv ? kdiv_rn(n, d) : n / d;
}

/* lib/ref/ref_debayer.c:195-202 */
KDEV float rcd_ratio(float a, float b, bool kdiv)
{
    a = a > 0.0f ? a : 0.0f;
    b = b > 0.0f ? b : 0.0f;
    return rcd_div(a - b, a + b + RCD_EPS, kdiv);
}

/* lib/ref/ref_debayer.c:204-395. Stages and the apron each
 * is needed over: V 5, low pass 4, gradients 3, wv 2, wp 1, green 2,
 * colour estimate 1, output 0. Every stage computes at the mirrored
 * position of each apron position, from the previous stage's values at
 * that position's own (virtual) neighbourhood. */
struct SmRcd {
    float sv[RcStage<5>::N];
    float slp[RcStage<4>::N];
    /* gradients: vertical, horizontal, NE-SW, NW-SE; reused as the RGB
     * tile once the weights are formed */
    float sg[4][RcStage<3>::N];
    float swv[RcStage<2>::N];
    float swp[RcStage<1>::N];
    float sgp[RcStage<2>::N];
    float scp[RcStage<1>::N];
};

/* A stage loop over the full (RC_TW + 2A) x (RC_TH + 2A) region of a
 * stage with apron A, with compile-time strides. In an interior tile
 * (every stage's region inside the frame) a position is its own mirrored
 * position; elsewhere positions past the partial tile's own reach are
 * skipped and the rest are mirrored. Sets lx, ly (region coordinates) and
 * mx, my (the mirrored frame position). */
#define RC_STAGE_BEGIN(A)                                                                  \
    for (int i_ = tid; i_ < RcStage<A>::N; i_ += RC_NT) {                                  \
        const int lx = i_ % RcStage<A>::W, ly = i_ / RcStage<A>::W;                        \
        int mx = x0 - (A) + lx, my = y0 - (A) + ly;                                        \
        if (!INTERIOR) {                                                                   \
            if (lx >= tw + 2 * (A) || ly >= th + 2 * (A))                                  \
                continue;                                                                  \
            mx = kmirror(mx, W);                                                           \
            my = kmirror(my, H);                                                           \
        }
#define RC_STAGE_END }

template <bool INTERIOR>
KDEV void dev_debayer_rcd(SmRcd &sm, const KFrame &f, int x0, int y0)
{
    typedef RcStage<5> S5;
    typedef RcStage<4> S4;
    typedef RcStage<3> S3;
    typedef RcStage<2> S2;
    typedef RcStage<1> S1;
    float *sv = sm.sv, *slp = sm.slp, *swv = sm.swv, *swp = sm.swp, *sgp = sm.sgp, *scp = sm.scp;
    float (&sg)[4][S3::N] = sm.sg;
    const uint16_t *__restrict__ src = f.src;
    const int src_pitch = f.src_pitch, W = f.W, H = f.H;
    const KPrep &prep = f.prep;
    const bool kdiv = prep.kdiv != 0;
    const int tid = threadIdx.x;
    const int xe = min(x0 + RC_TW, W), ye = min(y0 + RC_TH, H);
    const int tw = xe - x0, th = ye - y0;

    /* Position (X, Y) in frame coordinates of a stage with apron A. */
#define AT(arr, A, X, Y) arr[((Y) - y0 + (A)) * RcStage<A>::W + ((X) - x0 + (A))]
#define V(dx, dy) AT(sv, 5, mx + (dx), my + (dy))
#define L(dx, dy) AT(slp, 4, mx + (dx), my + (dy))
#define G(dx, dy) AT(sgp, 2, mx + (dx), my + (dy))
#define C(dx, dy) AT(scp, 1, mx + (dx), my + (dy))

    /* V: prep, apron 5 */
    RC_STAGE_BEGIN(5)
        int s = (my & 1) * 2 + (mx & 1);
        sv[ly * S5::W + lx] = kprep(src[(size_t)my * src_pitch + mx], s, prep);
    RC_STAGE_END
    __syncthreads();

    /* Pass A: low pass (apron 4) and the four gradients (apron 3). */
    RC_STAGE_BEGIN(4)
        slp[ly * S4::W + lx] =
            (4.0f * V(0, 0)
             + 2.0f * (V(0, -1) + V(0, 1) + V(-1, 0) + V(1, 0))
             + V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1)) * (1.0f / 16.0f);
        if (lx >= 1 && ly >= 1 && lx < S4::W - 1 && ly < S4::H - 1 &&
            (INTERIOR || (lx < tw + 7 && ly < th + 7))) {
            int j = (ly - 1) * S3::W + (lx - 1);
            sg[0][j] = fabsf(V(0, -1) - V(0, 1))
                     + fabsf(2.0f * V(0, 0) - V(0, -2) - V(0, 2));
            sg[1][j] = fabsf(V(-1, 0) - V(1, 0))
                     + fabsf(2.0f * V(0, 0) - V(-2, 0) - V(2, 0));
            sg[2][j] = fabsf(V(-1, 1) - V(1, -1))
                     + fabsf(2.0f * V(0, 0) - V(-2, 2) - V(2, -2));
            sg[3][j] = fabsf(V(-1, -1) - V(1, 1))
                     + fabsf(2.0f * V(0, 0) - V(-2, -2) - V(2, 2));
        }
    RC_STAGE_END
    __syncthreads();

    /* Pass A: direction weights, wv over apron 2 and wp over apron 1, from
     * 3x3 means of the gradients summed row by row
     * (lib/ref/ref_debayer.c:179-190). */
    RC_STAGE_BEGIN(2)
        float a0 = 0.0f, a1 = 0.0f;

#pragma unroll
        for (int dy = -1; dy <= 1; dy++) {
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                a0 += AT(sg[0], 3, mx + dx, my + dy);
                a1 += AT(sg[1], 3, mx + dx, my + dy);
            }
        }
        {
            float vsum = a0 * (1.0f / 9.0f), hsum = a1 * (1.0f / 9.0f);
            swv[ly * S2::W + lx] = rcd_div(hsum, hsum + vsum + RCD_EPS, kdiv);
        }
        if (lx >= 1 && ly >= 1 && lx < S2::W - 1 && ly < S2::H - 1 &&
            (INTERIOR || (lx < tw + 3 && ly < th + 3))) {
            float b0 = 0.0f, b1 = 0.0f;

#pragma unroll
            for (int dy = -1; dy <= 1; dy++) {
#pragma unroll
                for (int dx = -1; dx <= 1; dx++) {
                    b0 += AT(sg[2], 3, mx + dx, my + dy);
                    b1 += AT(sg[3], 3, mx + dx, my + dy);
                }
            }
            {
                float ps = b0 * (1.0f / 9.0f), qs = b1 * (1.0f / 9.0f);
                swp[(ly - 1) * S1::W + (lx - 1)] = rcd_div(qs, ps + qs + RCD_EPS, kdiv);
            }
        }
    RC_STAGE_END
    __syncthreads();

    /* Pass B: green, apron 2. */
    RC_STAGE_BEGIN(2)
        int s = (my & 1) * 2 + (mx & 1);
        float gval;

        if (s == 1 || s == 2) {
            gval = V(0, 0);
        } else {
            float l00 = L(0, 0), lu = L(0, -2), ld = L(0, 2), ll = L(-2, 0), lr = L(2, 0);
            float gv = 0.5f * (V(0, -1) * (1.0f + rcd_ratio(l00, lu, kdiv))
                             + V(0, 1) * (1.0f + rcd_ratio(l00, ld, kdiv)));
            float gh = 0.5f * (V(-1, 0) * (1.0f + rcd_ratio(l00, ll, kdiv))
                             + V(1, 0) * (1.0f + rcd_ratio(l00, lr, kdiv)));
            float w = AT(swv, 2, mx, my);
            gval = w * gv + (1.0f - w) * gh;
        }
        sgp[ly * S2::W + lx] = gval;
    RC_STAGE_END
    __syncthreads();

    /* Pass C, first half: colour estimate at the red and blue sites,
     * apron 1. */
    RC_STAGE_BEGIN(1)
        int s = (my & 1) * 2 + (mx & 1);
        float c = 0.0f;

        if (s == 0 || s == 3) {
            float g00 = G(0, 0);
            float dsw = V(-1, 1) - G(-1, 1);
            float dne = V(1, -1) - G(1, -1);
            float dnw = V(-1, -1) - G(-1, -1);
            float dse = V(1, 1) - G(1, 1);
            float w = AT(swp, 1, mx, my);

            c = w * (g00 + 0.5f * (dsw + dne)) + (1.0f - w) * (g00 + 0.5f * (dnw + dse));
        }
        scp[ly * S1::W + lx] = c;
    RC_STAGE_END
    __syncthreads();

    /* Output: RGB of the tile's own pixels into the gradient space. */
    {
        float (*rgbs)[RC_TH][RC_TW] = (float (*)[RC_TH][RC_TW])&sg[0][0];

#pragma unroll
        for (int k = 0; k < RC_TW * RC_TH / RC_NT; k++) {
            int p = tid + k * RC_NT, px = p % RC_TW, py = p / RC_TW;
            int mx = x0 + px, my = y0 + py;
            int s = (my & 1) * 2 + (mx & 1);
            float r, g, b;

            if (mx >= W || my >= H)
                continue;
            if (s == 0 || s == 3) {
                float c = C(0, 0);
                g = G(0, 0);
                r = s == 0 ? V(0, 0) : c;
                b = s == 0 ? c : V(0, 0);
            } else {
                /* Pass C, second half (lib/ref/ref_debayer.c:350-380) */
                float g00 = G(0, 0), w = AT(swv, 2, mx, my);
                float dh = 0.5f * ((V(-1, 0) - G(-1, 0)) + (V(1, 0) - G(1, 0)));
                float ev = 0.5f * ((C(0, -1) - G(0, -1)) + (C(0, 1) - G(0, 1)));
                float ch = w * (g00 + ev) + (1.0f - w) * (g00 + dh);
                float dv = 0.5f * ((V(0, -1) - G(0, -1)) + (V(0, 1) - G(0, 1)));
                float eh = 0.5f * ((C(-1, 0) - G(-1, 0)) + (C(1, 0) - G(1, 0)));
                float cv = w * (g00 + dv) + (1.0f - w) * (g00 + eh);

                r = s == 1 ? ch : cv;
                g = V(0, 0);
                b = s == 1 ? cv : ch;
            }
            rgbs[0][py][px] = r;
            rgbs[1][py][px] = g;
            rgbs[2][py][px] = b;
        }
    }
#undef AT
#undef V
#undef L
#undef G
#undef C
}
#undef RC_STAGE_BEGIN
#undef RC_STAGE_END

/* Whether every stage region of the tile at (x0, y0) lies in the frame. */
KDEV bool rcd_interior(const KFrame &f, int x0, int y0)
{
    return x0 >= 5 && y0 >= 5 && x0 + RC_TW + 5 <= f.W && y0 + RC_TH + 5 <= f.H;
}

/* One RCD tile: the stages (with or without the mirror) leave the tile's
 * camera RGB in LDS, then one copy of the tail stores it. */
template <bool INTER>
KDEV void dev_debayer_rcd_tile(SmRcd &sm, const KFrame &f, int x0, int y0)
{
    const float (*rgbs)[RC_TH][RC_TW] = (const float (*)[RC_TH][RC_TW])&sm.sg[0][0];

    if (rcd_interior(f, x0, y0))
        dev_debayer_rcd<true>(sm, f, x0, y0);
    else
        dev_debayer_rcd<false>(sm, f, x0, y0);
    __syncthreads();
    if (INTER)
        kstore_inter<RC_TW, RC_TH, RC_NT>(rgbs, x0, y0, f.W, f.H, f.inter);
    else
        kfinish<RC_TW, RC_TH, RC_NT>(rgbs, x0, y0, f.W, f.H, f.col, f.out);
}

#endif /* CRAW_K_RCD_TILE_H */
