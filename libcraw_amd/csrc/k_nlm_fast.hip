/*
 * libcraw for MI355X: the fast forms of the raw-domain denoise.
 *
 * Same filter as k_nlm_sym.hip (the definition of
 * lib/ref/ref_denoise.c:114-154 of cyanreg/libcraw), evaluated for throughput
 * rather than bit-exactness: the patch distance of every offset is built
 * from each position's squared difference computed once, summed in three
 * across the lanes of a DPP row (row_shr:1 / row_shl:1) and in three down
 * the lane's own registers, and the weight uses the hardware exp2. With
 * FP16 the squared differences, the 3x3 sums and the weight's argument are
 * packed half2 arithmetic over pairs of rows; weights and weighted sums
 * are always accumulated in FP32. The FP32 form differs from the CPU
 * reference by summation order and exp rounding only; the FP16 form's
 * error is bounded in the tests against an FP64 evaluation (within 2^-11
 * of the clip level).
 *
 * Layout: a workgroup of 256 lanes is 16 DPP rows of 16 lanes. DPP row g
 * covers plane columns x0 + 14 (g & 1) - 1 ... + 14 (lane 0 and lane 15
 * only feed their neighbours' row sums) and plane rows y0 + 8 (g >> 1) ...
 * + 7, so a workgroup's tile is 28 x 64 samples; the stabilised plane is
 * staged in LDS over the tile plus the 4-sample reach.
 */
#include "k_common.h"
#include "kernels.h"

#define FT_CW 14                  /* output columns per DPP row */
#define FT_TW (2 * FT_CW)         /* 28 */
#define FT_R  8                   /* rows per lane */
#define FT_TH (8 * FT_R)          /* 64 */
#define FT_NT 256
#define FT_A  4                   /* search 3 + patch 1 */
#define FT_LW (FT_TW + 2 * FT_A)  /* 36 */
#define FT_LH (FT_TH + 2 * FT_A)  /* 72 */

#define DPP_ROW_SHR1 0x111
#define DPP_ROW_SHL1 0x101

KDEV float dpp_f(float v, int ctrl)
{
    int r;
    if (ctrl == DPP_ROW_SHR1)
        r = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), DPP_ROW_SHR1, 0xf, 0xf,
                                        true);
    else
        r = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), DPP_ROW_SHL1, 0xf, 0xf,
                                        true);
    return __builtin_bit_cast(float, r);
}

typedef _Float16 h2v __attribute__((ext_vector_type(2)));

KDEV h2v dpp_h2(h2v v, int ctrl)
{
    int r;
    if (ctrl == DPP_ROW_SHR1)
        r = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), DPP_ROW_SHR1, 0xf, 0xf,
                                        true);
    else
        r = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), DPP_ROW_SHL1, 0xf, 0xf,
                                        true);
    return __builtin_bit_cast(h2v, r);
}

KDEV float fgat(const KDenoise &p, float u)
{
    if (u >= 0.0f)
        return p.gat_fwd * sqrtf(p.gat_a * u + p.gat_c);
    return p.gat_t0 + u / p.gat_sc;
}

KDEV float figat(const KDenoise &p, float t)
{
    if (t >= p.gat_t0)
        return p.gat_inv * t * t - p.gat_off;
    return (t - p.gat_t0) * p.gat_sc;
}

template <int MODE>
struct SmFt {
    static constexpr bool HISTORY = MODE & KDN_HISTORY;
    static constexpr int CW = FT_LW + 2, CH = FT_LH + 2;
    float splane[FT_LH * FT_LW];
    float svar[HISTORY ? FT_LH * FT_LW : 1];
    float scur[HISTORY ? CH * CW : 1];
    float stp[HISTORY ? CH * CW : 1];
};

template <int MODE, bool FP16, int UNROLL>
KDEV void dev_denoise_fast(SmFt<MODE> &sm, const KFrame &f, int x0, int y0, int pl)
{
    constexpr bool SPATIAL = MODE & KDN_SPATIAL;
    constexpr bool TEMPORAL = MODE & KDN_TEMPORAL;
    constexpr bool HISTORY = MODE & KDN_HISTORY;
    constexpr int CW = SmFt<MODE>::CW;
    float *splane = sm.splane, *svar = sm.svar, *scur = sm.scur, *stp = sm.stp;
    const KDenoise &p = f.dn;
    const KDenoiseState &st = f.st;
    const uint16_t *__restrict__ src = f.dn_src;
    uint16_t *__restrict__ dst = f.dn_dst;
    const int tid = threadIdx.x;
    const int px = pl & 1, py = pl >> 1;
    const int w2 = p.w2, h2 = p.h2;
    const int xe = min(x0 + FT_TW, w2), ye = min(y0 + FT_TH, h2);
    const int tw = xe - x0, th = ye - y0;
    const float black = ksel4(pl, p.black[0], p.black[1], p.black[2], p.black[3]);
    const float range = ksel4(pl, p.range[0], p.range[1], p.range[2], p.range[3]);
    const float inv_range = ksel4(pl, p.inv_range[0], p.inv_range[1], p.inv_range[2],
                                  p.inv_range[3]);
    const size_t psize = (size_t)w2 * h2;

#define RAW(X, Y) src[(size_t)(2 * (Y) + py) * p.src_pitch + 2 * (X) + px]

    /* Stage the stabilised plane (and with history the average and its
     * variance) over the tile plus the reach, through the plane mirror. */
    if (HISTORY) {
        const int rw = tw + 2 * (FT_A + 1), rh = th + 2 * (FT_A + 1);

        for (int i = tid; i < rw * rh; i += FT_NT) {
            int lx = i % rw, ly = i / rw;
            int mx = kmirror(x0 - FT_A - 1 + lx, w2), my = kmirror(y0 - FT_A - 1 + ly, h2);
            float raw = (float)RAW(mx, my) * (1.0f / 65535.0f);

            scur[ly * CW + lx] = fgat(p, (raw - black) * inv_range);
            stp[ly * CW + lx] = fgat(p, st.prev_level[pl * psize + (size_t)my * w2 + mx]);
        }
        __syncthreads();
        {
            const int qw = tw + 2 * FT_A, qh = th + 2 * FT_A;

            for (int i = tid; i < qw * qh; i += FT_NT) {
                int lx = i % qw, ly = i / qw;
                int vx = x0 - FT_A + lx, vy = y0 - FT_A + ly;
                int mx = kmirror(vx, w2), my = kmirror(vy, h2);
                int cx = mx - (x0 - FT_A - 1), cy = my - (y0 - FT_A - 1);
                float d2 = 0.0f, a, t, pv, v;

                for (int ky = -1; ky <= 1; ky++)
                    for (int kx = -1; kx <= 1; kx++) {
                        int j = (cy + ky) * CW + cx + kx;
                        float e = stp[j] - scur[j];
                        d2 = fmaf(e, e, d2);
                    }
                a = p.alpha_max * exp2f(fminf(fmaf(d2, -(1.0f / 9.0f) * p.inv_h2 * 1.44269504f,
                                                   2.0f * 1.44269504f), 0.0f));
                t = scur[cy * CW + cx] * (1.0f - a) + stp[cy * CW + cx] * a;
                {
                    float q = (float)st.prev_var[pl * psize + (size_t)my * w2 + mx] *
                              (1.0f / 255.0f);
                    pv = q * q;
                }
                v = fmaxf((1.0f - a) * (1.0f - a) + a * a * pv, 1.0e-4f);
                splane[ly * FT_LW + lx] = t;
                svar[ly * FT_LW + lx] = v;
                if (vx == mx && vy == my && vx >= x0 && vx < xe && vy >= y0 && vy < ye) {
                    size_t o = pl * psize + (size_t)vy * w2 + vx;
                    float vc = fminf(fmaxf(v, 0.0f), 1.0f);
                    st.next_level[o] = figat(p, t);
                    st.next_var[o] = (uint8_t)rintf(sqrtf(vc) * 255.0f);
                }
            }
        }
        __syncthreads();
    } else {
        const int rw = tw + 2 * FT_A, rh = th + 2 * FT_A;

        for (int i = tid; i < rw * rh; i += FT_NT) {
            int lx = i % rw, ly = i / rw;
            int vx = x0 - FT_A + lx, vy = y0 - FT_A + ly;
            int mx = kmirror(vx, w2), my = kmirror(vy, h2);
            float raw = (float)RAW(mx, my) * (1.0f / 65535.0f);
            float t = fgat(p, (raw - black) * inv_range);

            splane[ly * FT_LW + lx] = t;
            if (TEMPORAL && vx == mx && vy == my && vx >= x0 && vx < xe && vy >= y0 && vy < ye) {
                size_t o = pl * psize + (size_t)vy * w2 + vx;
                st.next_level[o] = figat(p, t);
                st.next_var[o] = 255;
            }
        }
        __syncthreads();
    }

    /* This lane's column and rows. */
    const int lane = tid & 63, g = (tid >> 6) * 4 + (lane >> 4), li = lane & 15;
    const int x = x0 + (g & 1) * FT_CW + li - 1;
    const int ys = y0 + (g >> 1) * FT_R;
    const int lx = x - (x0 - FT_A);                    /* column in the LDS tile */
    const int ly0 = ys - 1 - (y0 - FT_A);              /* LDS row of strip row -1 */
    float out_t[FT_R];

    if (SPATIAL) {
        const float l2e = 1.44269504088896f;
        const float kd = -(1.0f / 9.0f) * p.inv_h2 * l2e;
        float c[FT_R + 2], wsum[FT_R], acc[FT_R], vcs[FT_R], ivc[FT_R];

#pragma unroll
        for (int j = 0; j < FT_R + 2; j++)
            c[j] = splane[(ly0 + j) * FT_LW + lx];
#pragma unroll
        for (int i = 0; i < FT_R; i++) {
            wsum[i] = 0.0f;
            acc[i] = 0.0f;
            if (HISTORY) {
                vcs[i] = svar[(ly0 + 1 + i) * FT_LW + lx];
                ivc[i] = 1.0f / vcs[i];
            }
        }

        /* the search rows: a loop, or (UNROLL 7) unrolled */
#pragma unroll UNROLL
        for (int dy = -3; dy <= 3; dy++) {
#pragma unroll
            for (int dx = -3; dx <= 3; dx++) {
                const float *nrow = &splane[(ly0 + dy) * FT_LW + lx + dx];
                float nb[FT_R + 2], rs[FT_R + 2];

#pragma unroll
                for (int j = 0; j < FT_R + 2; j++)
                    nb[j] = nrow[j * FT_LW];

                if (FP16) {
                    /* pairs of rows (2k, 2k + 1), packed half2 */
                    h2v rsh[(FT_R + 2) / 2];
#pragma unroll
                    for (int k = 0; k < (FT_R + 2) / 2; k++) {
                        h2v e = __builtin_bit_cast(h2v, __builtin_amdgcn_cvt_pkrtz(
                            c[2 * k] - nb[2 * k], c[2 * k + 1] - nb[2 * k + 1]));
                        h2v sq = e * e;
                        rsh[k] = sq + dpp_h2(sq, DPP_ROW_SHR1) + dpp_h2(sq, DPP_ROW_SHL1);
                    }
                    /* 3-row sums for output rows (2k, 2k + 1) = rows
                     * (2k..2k+2, 2k+1..2k+3) of the strip-with-halo; then
                     * the weight's argument d2 kd + 2 log2 e, clamped at 0,
                     * in half2 too (with a history the argument needs the
                     * two variances and is formed in FP32). */
#pragma unroll
                    for (int k = 0; k < FT_R / 2; k++) {
                        h2v mid = { rsh[k].y, rsh[k + 1].x };
                        h2v d2 = rsh[k] + rsh[k + 1] + mid;

                        if (!HISTORY) {
                            const h2v kk = { (_Float16)kd, (_Float16)kd };
                            const h2v c2 = { (_Float16)(2.0f * l2e), (_Float16)(2.0f * l2e) };
                            const h2v zero = { (_Float16)0.0f, (_Float16)0.0f };
                            d2 = __builtin_elementwise_min(__builtin_elementwise_fma(d2, kk, c2),
                                                           zero);
                        }
                        rs[2 * k] = (float)d2.x;
                        rs[2 * k + 1] = (float)d2.y;
                    }
                } else {
                    float sq[FT_R + 2];
#pragma unroll
                    for (int j = 0; j < FT_R + 2; j++) {
                        float e = c[j] - nb[j];
                        sq[j] = e * e;
                    }
#pragma unroll
                    for (int j = 0; j < FT_R + 2; j++)
                        rs[j] = sq[j] + dpp_f(sq[j], DPP_ROW_SHR1) + dpp_f(sq[j], DPP_ROW_SHL1);
                }

#pragma unroll
                for (int i = 0; i < FT_R; i++) {
                    float w;

                    if (FP16) {
                        /* rs[i]: the clamped argument, or with a history d2 */
                        if (HISTORY) {
                            float v2 = vcs[i] + svar[(ly0 + 1 + i + dy) * FT_LW + lx + dx];
                            w = exp2f(fminf(fmaf(rs[i], kd, v2 * l2e), 0.0f) * ivc[i]);
                        } else {
                            w = exp2f(rs[i]);
                        }
                    } else {
                        float d2 = rs[i] + rs[i + 1] + rs[i + 2];

                        if (HISTORY) {
                            float v2 = vcs[i] + svar[(ly0 + 1 + i + dy) * FT_LW + lx + dx];
                            w = exp2f(fminf(fmaf(d2, kd, v2 * l2e), 0.0f) * ivc[i]);
                        } else {
                            w = exp2f(fminf(fmaf(d2, kd, 2.0f * l2e), 0.0f));
                        }
                    }
                    wsum[i] += w;
                    acc[i] = fmaf(w, nb[i + 1], acc[i]);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < FT_R; i++)
            out_t[i] = acc[i] / wsum[i];
    } else {
#pragma unroll
        for (int i = 0; i < FT_R; i++)
            out_t[i] = splane[(ly0 + 1 + i) * FT_LW + lx];
    }

    if (li < 1 || li > FT_CW || x >= xe)
        return;
#pragma unroll
    for (int i = 0; i < FT_R; i++) {
        int y = ys + i;
        float u, raw0, u0, gr, raw;

        if (y >= ye)
            break;
        u = figat(p, out_t[i]);
        raw0 = (float)RAW(x, y) * (1.0f / 65535.0f);
        u0 = (raw0 - black) * inv_range;
        gr = fminf(fmaxf((u0 - p.guard_lo) * p.guard_inv, 0.0f), 1.0f);
        u = u + (u0 - u) * gr;
        raw = fminf(fmaxf(black + u * range, 0.0f), 1.0f);
        dst[(size_t)(2 * y + py) * p.dst_pitch + 2 * x + px] = (uint16_t)rintf(raw * 65535.0f);
    }
#undef RAW
}

template <int MODE, bool FP16>
__global__ __launch_bounds__(FT_NT) void k_denoise_fast(KFrame f)
{
    __shared__ SmFt<MODE> sm;
    const int ntx = (f.dn.w2 + FT_TW - 1) / FT_TW, nty = (f.dn.h2 + FT_TH - 1) / FT_TH;
    int t, pl;

    if (!craw_dn_pair_tile(blockIdx.x, ntx * nty, &t, &pl))
        return;
    dev_denoise_fast<MODE, FP16, 1>(sm, f, (t % ntx) * FT_TW, (t / ntx) * FT_TH, pl);
}

template <int MODE, bool FP16, int UNROLL>
__global__ __launch_bounds__(FT_NT) void kb_denoise_fast(const KFrame *__restrict__ frames,
                                                         const uint32_t *__restrict__ tiles,
                                                         uint32_t ntiles)
{
    __shared__ SmFt<MODE> sm;

    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t d = tiles[t];
        const KFrame &f = frames[KTILE_FRAME(d)];
        const int ntx = (f.dn.w2 + FT_TW - 1) / FT_TW, nty = (f.dn.h2 + FT_TH - 1) / FT_TH;
        int r, pl;

        if (!craw_dn_pair_tile(KTILE_IDX(d), ntx * nty, &r, &pl))
            continue;
        __syncthreads();
        dev_denoise_fast<MODE, FP16, UNROLL>(sm, f, (r % ntx) * FT_TW, (r / ntx) * FT_TH, pl);
    }
}

static_assert(FT_TW == KG_FT_TW && FT_TH == KG_FT_TH, "tile geometry");

#define FT_MODES(X)                                                                        \
    X(KDN_SPATIAL) X(KDN_TEMPORAL) X(KDN_TEMPORAL | KDN_HISTORY)                           \
    X(KDN_SPATIAL | KDN_TEMPORAL) X(KDN_SPATIAL | KDN_TEMPORAL | KDN_HISTORY)

template <bool FP16>
static hipError_t launch_fast(int mode, const KFrame &f, hipStream_t s)
{
    dim3 grid(craw_dn_pair_count(((f.dn.w2 + FT_TW - 1) / FT_TW) *
                                 ((f.dn.h2 + FT_TH - 1) / FT_TH)));

    switch (mode) {
#define X(M)                                                                               \
    case M:                                                                                \
        KLAUNCH((k_denoise_fast<M, FP16>), grid, dim3(FT_NT), 0, s, f);         \
        break;
    FT_MODES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool FP16>
static hipError_t launch_batch_fast(int mode, bool unroll, const KFrame *frames,
                                    const uint32_t *tiles, uint32_t ntiles, int grid,
                                    hipStream_t s)
{
    switch (mode) {
#define X(M)                                                                               \
    case M:                                                                                \
        if (unroll)                                                                        \
            KLAUNCH((kb_denoise_fast<M, FP16, 7>), dim3(grid), dim3(FT_NT), 0, s, frames,  \
                    tiles, ntiles);                                                        \
        else                                                                               \
            KLAUNCH((kb_denoise_fast<M, FP16, 1>), dim3(grid), dim3(FT_NT), 0, s, frames,  \
                    tiles, ntiles);                                                        \
        break;
    FT_MODES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t craw_launch_denoise_fast(int mode, bool fp16, const KFrame &f, hipStream_t s)
{
    if (fp16 && craw_dn_geom(KFORM_FP16, mode) != KDN_GEOM_FAST)
        return craw_launch_denoise_sym16(mode, f, s);
    return fp16 ? launch_fast<true>(mode, f, s) : launch_fast<false>(mode, f, s);
}

hipError_t craw_launch_batch_denoise_exact(int mode, int shape, const KFrame *frames,
                                           const uint32_t *tiles, uint32_t ntiles, int grid,
                                           hipStream_t s);

hipError_t craw_launch_batch_denoise(int mode, int form, int shape, const KFrame *frames,
                                     const uint32_t *tiles, uint32_t ntiles, int grid,
                                     hipStream_t s)
{
    const bool unroll = (shape & KDN_SHAPE_UNROLL) != 0;

    if (!ntiles)
        return hipSuccess;
    if (form == KFORM_EXACT)
        return craw_launch_batch_denoise_exact(mode, shape, frames, tiles, ntiles, grid, s);
    if (form == KFORM_FP16 && craw_dn_geom(KFORM_FP16, mode) != KDN_GEOM_FAST)
        return craw_launch_batch_denoise_sym16(mode, frames, tiles, ntiles, grid,
                                               (shape & KDN_SHAPE_TALL) != 0, s);
    return form == KFORM_FP16
        ? launch_batch_fast<true>(mode, unroll, frames, tiles, ntiles, grid, s)
        : launch_batch_fast<false>(mode, unroll, frames, tiles, ntiles, grid, s);
}
