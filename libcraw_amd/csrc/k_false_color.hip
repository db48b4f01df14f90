/*
 * libcraw for MI355X: false colour suppression (lib/ref/ref_false_color.c:
 * 56-134 of cyanreg/libcraw) on the camera RGB the debayer left in three
 * float planes, fused with what reads it next: the chroma stage's inputs
 * (planes and 4x4 means) or the AWB block samples and the colour stage
 * into the destination.
 *
 * One workgroup per 64x32 pixels, the pass count P a template parameter.
 * The colour differences (R - G, B - G) of the tile and an apron of 2P
 * pixels are staged in LDS as float2, the apron through the frame's mirror
 * (one reflection, repeated for a frame narrower than the apron). Sweep s
 * of the 2P (a 3x3 median when s is odd, a 3x3 box mean when even) is
 * computed on the pixels within 2P - s of the tile, ping-ponging between
 * two LDS buffers: exactly what the next sweep and the tile's own pixels
 * read. A position outside the frame takes the sweep at the frame pixel it
 * mirrors, evaluated on that pixel's own neighbourhood in its own order, so
 * the box sums are the sequential definition's bit for bit at the borders
 * too; that pixel and its neighbourhood always lie in the part the previous
 * sweep computed.
 *
 * The median is a selection by min3 / med3 / max3: each row of three
 * sorted, then the median of (the largest minimum, the median of the
 * medians, the smallest maximum), which is the fifth of the nine exactly
 * (the reference sorts the nine; both return one of the inputs).
 */
#include "k_common.h"
#include "kernels.h"
#include "k_finish.h"

#define FC_TW 64
#define FC_TH 32
#define FC_NT 256

template <int P>
struct FcGeom {
    static constexpr int R = 2 * P;
    static constexpr int LW = FC_TW + 2 * R, LH = FC_TH + 2 * R;
    static constexpr int BUF = LW * LH;                      /* float2 a buffer */
    static constexpr int RGB_BYTES = 3 * FC_TH * FC_TW * 4;  /* kfinish's tile */
    static constexpr int BYTES = BUF * 8 + (BUF * 8 > RGB_BYTES ? BUF * 8 : RGB_BYTES);
};

/* lib/ref/ref_false_color.c:16-23, repeated until inside. */
KDEV int fc_mirror(int v, int n)
{
    while (v < 0 || v >= n) {
        if (v < 0)
            v = -v;
        if (v >= n)
            v = 2 * n - 2 - v;
    }
    return v;
}

KDEV float fc_med9(float a0, float a1, float a2, float a3, float a4, float a5, float a6,
                   float a7, float a8)
{
    const float l0 = fminf(fminf(a0, a1), a2), m0 = __builtin_amdgcn_fmed3f(a0, a1, a2);
    const float h0 = fmaxf(fmaxf(a0, a1), a2);
    const float l1 = fminf(fminf(a3, a4), a5), m1 = __builtin_amdgcn_fmed3f(a3, a4, a5);
    const float h1 = fmaxf(fmaxf(a3, a4), a5);
    const float l2 = fminf(fminf(a6, a7), a8), m2 = __builtin_amdgcn_fmed3f(a6, a7, a8);
    const float h2 = fmaxf(fmaxf(a6, a7), a8);

    return __builtin_amdgcn_fmed3f(fmaxf(fmaxf(l0, l1), l2), __builtin_amdgcn_fmed3f(m0, m1, m2),
                                   fminf(fminf(h0, h1), h2));
}

/* lib/ref/ref_false_color.c:46-54: stencil order, then the nearest 1/9. */
KDEV float fc_mean9(float a0, float a1, float a2, float a3, float a4, float a5, float a6,
                    float a7, float a8)
{
    return ((((((((a0 + a1) + a2) + a3) + a4) + a5) + a6) + a7) + a8) * (1.0f / 9.0f);
}

/* Sweep S of a tile whose staged region lies wholly inside the frame (no
 * mirror): a work item is a column and a run of FC_K rows, which load
 * FC_K + 2 rows of three once; the median sorts each row's three once
 * (min3 / med3 / max3) for the up to three outputs that use it, the box
 * sum starts from the top row's (a + b) + c, the prefix of the stencil
 * order, and adds the other six in that order. */
#define FC_K 4

template <int P, int S>
KDEV void fc_sweep_inside(const float2 *cur, float2 *nxt)
{
    constexpr int R = 2 * P, LW = FC_TW + 2 * R;
    constexpr int CW = FC_TW + 2 * (R - S), CH = FC_TH + 2 * (R - S);
    constexpr int NCH = (CH + FC_K - 1) / FC_K;

    for (int it = threadIdx.x; it < CW * NCH; it += FC_NT) {
        const int lx = S + it % CW, ly0 = S + (it / CW) * FC_K;
        float2 v[FC_K + 2][3];

#pragma unroll
        for (int j = 0; j < FC_K + 2; j++) {
            /* a short last run reads no further than the sweep's last row
             * below the region (which is inside the staged tile) */
            const int r = min(ly0 - 1 + j, S + CH);
            const float2 *row = cur + r * LW + lx;

            v[j][0] = row[-1];
            v[j][1] = row[0];
            v[j][2] = row[1];
        }
        if constexpr (S & 1) {
            float lo[FC_K + 2][2], md[FC_K + 2][2], hi[FC_K + 2][2];

#pragma unroll
            for (int j = 0; j < FC_K + 2; j++) {
                lo[j][0] = fminf(fminf(v[j][0].x, v[j][1].x), v[j][2].x);
                md[j][0] = __builtin_amdgcn_fmed3f(v[j][0].x, v[j][1].x, v[j][2].x);
                hi[j][0] = fmaxf(fmaxf(v[j][0].x, v[j][1].x), v[j][2].x);
                lo[j][1] = fminf(fminf(v[j][0].y, v[j][1].y), v[j][2].y);
                md[j][1] = __builtin_amdgcn_fmed3f(v[j][0].y, v[j][1].y, v[j][2].y);
                hi[j][1] = fmaxf(fmaxf(v[j][0].y, v[j][1].y), v[j][2].y);
            }
#pragma unroll
            for (int j = 0; j < FC_K; j++) {
                float o[2];

                if (ly0 + j >= S + CH)
                    break;
#pragma unroll
                for (int c = 0; c < 2; c++)
                    o[c] = __builtin_amdgcn_fmed3f(
                        fmaxf(fmaxf(lo[j][c], lo[j + 1][c]), lo[j + 2][c]),
                        __builtin_amdgcn_fmed3f(md[j][c], md[j + 1][c], md[j + 2][c]),
                        fminf(fminf(hi[j][c], hi[j + 1][c]), hi[j + 2][c]));
                nxt[(ly0 + j) * LW + lx] = make_float2(o[0], o[1]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < FC_K; j++) {
                float2 t;

                if (ly0 + j >= S + CH)
                    break;
                t.x = (v[j][0].x + v[j][1].x) + v[j][2].x;
                t.y = (v[j][0].y + v[j][1].y) + v[j][2].y;
#pragma unroll
                for (int k = 1; k < 3; k++)
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        t.x = t.x + v[j + k][i].x;
                        t.y = t.y + v[j + k][i].y;
                    }
                nxt[(ly0 + j) * LW + lx] = make_float2(t.x * (1.0f / 9.0f), t.y * (1.0f / 9.0f));
            }
        }
    }
}

template <int P, int S>
struct FcSweeps {
    static KDEV void run(float2 *&cur, float2 *&nxt)
    {
        fc_sweep_inside<P, S>(cur, nxt);
        __syncthreads();
        float2 *t = cur;
        cur = nxt;
        nxt = t;
        if constexpr (S < 2 * P)
            FcSweeps<P, S + 1>::run(cur, nxt);
    }
};

template <int P, bool CHROMA>
__global__ __launch_bounds__(FC_NT) void k_false_color(KFrame f, KInter to)
{
    using G = FcGeom<P>;
    constexpr int R = G::R, LW = G::LW;
    __shared__ __align__(16) char smem[G::BYTES];
    float2 *cur = (float2 *)smem, *nxt = cur + G::BUF;
    const int tid = threadIdx.x;
    const int W = f.W, H = f.H;
    const int x0 = blockIdx.x * FC_TW, y0 = blockIdx.y * FC_TH;
    const int xe = min(x0 + FC_TW, W) - x0, ye = min(y0 + FC_TH, H) - y0;
    const size_t plane = (size_t)W * H;
    const float *rp = f.inter.rgb_out, *gp = rp + plane, *bp = rp + 2 * plane;

    /* the staged region (tile and apron) wholly inside the frame */
    const bool inside = x0 >= R && y0 >= R && x0 + FC_TW + R <= W && y0 + FC_TH + R <= H;

    /* local (lx, ly) stands for frame (x0 - R + lx, y0 - R + ly) */
    if (inside) {
        for (int i = tid; i < G::BUF; i += FC_NT) {
            const int lx = i % LW, ly = i / LW;
            const size_t o = (size_t)(y0 - R + ly) * W + x0 - R + lx;
            const float g = gp[o];

            cur[i] = make_float2(rp[o] - g, bp[o] - g);
        }
    } else {
        for (int i = tid; i < G::BUF; i += FC_NT) {
            const int lx = i % LW, ly = i / LW;
            const size_t o = (size_t)fc_mirror(y0 - R + ly, H) * W + fc_mirror(x0 - R + lx, W);
            const float g = gp[o];

            cur[i] = make_float2(rp[o] - g, bp[o] - g);
        }
    }
    __syncthreads();

    if (inside) {
        FcSweeps<P, 1>::run(cur, nxt);
    } else {
        /* near the frame's edges: a position outside the frame takes the
         * sweep at the pixel it mirrors, on that pixel's neighbourhood */
#pragma unroll
        for (int s = 1; s <= 2 * P; s++) {
            const int cw = xe + 2 * (R - s), ch = ye + 2 * (R - s);

            for (int i = tid; i < cw * ch; i += FC_NT) {
                const int lx = s + i % cw, ly = s + i / cw;
                const int cx = fc_mirror(x0 - R + lx, W) - x0 + R;
                const int cy = fc_mirror(y0 - R + ly, H) - y0 + R;
                const float2 *c = cur + cy * LW + cx;
                const float2 v0 = c[-LW - 1], v1 = c[-LW], v2 = c[-LW + 1];
                const float2 v3 = c[-1], v4 = c[0], v5 = c[1];
                const float2 v6 = c[LW - 1], v7 = c[LW], v8 = c[LW + 1];
                float2 r;

                if (s & 1) {
                    r.x = fc_med9(v0.x, v1.x, v2.x, v3.x, v4.x, v5.x, v6.x, v7.x, v8.x);
                    r.y = fc_med9(v0.y, v1.y, v2.y, v3.y, v4.y, v5.y, v6.y, v7.y, v8.y);
                } else {
                    r.x = fc_mean9(v0.x, v1.x, v2.x, v3.x, v4.x, v5.x, v6.x, v7.x, v8.x);
                    r.y = fc_mean9(v0.y, v1.y, v2.y, v3.y, v4.y, v5.y, v6.y, v7.y, v8.y);
                }
                nxt[ly * LW + lx] = r;
            }
            __syncthreads();
            float2 *t = cur;
            cur = nxt;
            nxt = t;
        }
    }

    /* 2P sweeps leave the result in the first buffer; the tile's camera
     * RGB goes after it */
    {
        float (*rgbs)[FC_TH][FC_TW] = (float (*)[FC_TH][FC_TW])(smem + G::BUF * 8);

        for (int i = tid; i < FC_TW * FC_TH; i += FC_NT) {
            const int px = i % FC_TW, py = i / FC_TW;
            float g;
            float2 d;

            if (px >= xe || py >= ye)
                continue;
            g = gp[(size_t)(y0 + py) * W + x0 + px];
            d = cur[(R + py) * LW + R + px];
            rgbs[0][py][px] = g + d.x;
            rgbs[1][py][px] = g;
            rgbs[2][py][px] = g + d.y;
        }
        __syncthreads();
        if constexpr (CHROMA)
            kstore_inter<FC_TW, FC_TH, FC_NT>(rgbs, x0, y0, W, H, to);
        else
            kfinish<FC_TW, FC_TH, FC_NT>(rgbs, x0, y0, W, H, f.col, f.out);
    }
}

template <int P>
static void fc_launch(bool chroma, const KFrame &f, const KInter &to, hipStream_t s)
{
    const dim3 grid((f.W + FC_TW - 1) / FC_TW, (f.H + FC_TH - 1) / FC_TH);

    if (chroma)
        KLAUNCH((k_false_color<P, true>), grid, dim3(FC_NT), 0, s, f, to);
    else
        KLAUNCH((k_false_color<P, false>), grid, dim3(FC_NT), 0, s, f, to);
}

bool craw_is_false_color_kernel(const void *fn)
{
    static const void *const k[10] = {
        (const void *)k_false_color<1, false>, (const void *)k_false_color<1, true>,
        (const void *)k_false_color<2, false>, (const void *)k_false_color<2, true>,
        (const void *)k_false_color<3, false>, (const void *)k_false_color<3, true>,
        (const void *)k_false_color<4, false>, (const void *)k_false_color<4, true>,
        (const void *)k_false_color<5, false>, (const void *)k_false_color<5, true>,
    };

    for (int i = 0; i < 10; i++)
        if (fn == k[i])
            return true;
    return false;
}

hipError_t craw_launch_false_color(int passes, bool chroma, const KFrame &f, const KInter &to,
                                   hipStream_t s)
{
    switch (passes) {
    case 1: fc_launch<1>(chroma, f, to, s); break;
    case 2: fc_launch<2>(chroma, f, to, s); break;
    case 3: fc_launch<3>(chroma, f, to, s); break;
    case 4: fc_launch<4>(chroma, f, to, s); break;
    case 5: fc_launch<5>(chroma, f, to, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
