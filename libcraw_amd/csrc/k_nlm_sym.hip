/*
 * libcraw for MI355X: the exact spatial NLM with each weight evaluated once
 * for the pair of samples it joins.
 *
 * lib/ref/ref_denoise.c:114-154 of cyanreg/libcraw weighs neighbour q of
 * sample p by exp(-max(d2(p, q) / 9 / h^2 - 2, 0)), d2 being the 3x3 patch
 * distance summed in raster order. The distance is symmetric bit for bit:
 * q's terms are p's negated differences, whose squares are equal, summed
 * in the same raster order; so w(p, q) = w(q, p) exactly. This kernel
 * evaluates the weight of every pair once, for the 24 offsets o after the
 * centre in raster order (W_o(v) = w(v, v + o)), and gives each sample its
 * 48 neighbour weights in the reference's raster order: offset -o of p is
 * W_o(p - o), offset o is W_o(p). The sums are then the reference's, with
 * half the distances and half the exps.
 *
 * Layout: a workgroup's four waves stand side by side over a 224x57 band of
 * the plane, each lane one column, and walk down it a row at a time. A lane
 * keeps, per offset, the squared differences of three rows in registers
 * and forms the weights of one row; the patch sums across columns come from
 * the neighbouring lanes through wave-wide DPP shifts, and the weight of a
 * neighbouring column through ds_bpermute. Each weight row adds its terms
 * at once to the four output rows it feeds (its own and the three below),
 * which carry their sums as running partial sums: every output's terms
 * still arrive in raster order. Of a wave's 64 columns, 62 have weights
 * and the middle 56 are outputs. The stabilised plane is staged in LDS
 * over the band plus 7 columns and 4 rows of reach.
 *
 * Only the modes without a temporal history use it (their weights do not
 * involve the per-sample variance, so they are symmetric).
 */
#include "k_common.h"
#include "kernels.h"

#define XS_NW    4                     /* waves per workgroup */
#define XS_NT    (64 * XS_NW)
#define XS_OUTW  56                    /* output columns per wave */
#define XS_TW    (XS_OUTW * XS_NW)     /* 224 */
#define XS_TH    57
#define XS_AX    7
#define XS_AY    4
#define XS_LW    (XS_TW + 2 * XS_AX)   /* 238 */
#define XS_LH    (XS_TH + 2 * XS_AY)   /* 64 */
#define XS_NO    24                    /* offsets after the centre */

#define DPP_WAVE_SHR1 0x138
#define DPP_WAVE_SHL1 0x130

static_assert(XS_TW == KG_XS_TW && XS_TH == KG_XS_TH, "tile geometry");

/* Offset o (0..23) after the centre, in raster order: dy 0, dx 1..3; then
 * dy 1..3, dx -3..3. */
__host__ __device__ constexpr int xs_dx(int o) { return o < 3 ? o + 1 : (o - 3) % 7 - 3; }
__host__ __device__ constexpr int xs_dy(int o) { return o < 3 ? 0 : (o - 3) / 7 + 1; }

/* The index of offset (dx, dy) after the centre. */
__host__ __device__ constexpr int xs_index(int dx, int dy)
{
    return dy == 0 ? dx - 1 : 3 + (dy - 1) * 7 + (dx + 3);
}

KDEV float xs_shr(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 DPP_WAVE_SHR1, 0xf, 0xf, true));
}

KDEV float xs_shl(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 DPP_WAVE_SHL1, 0xf, 0xf, true));
}

/* lib/ref/ref_denoise.c:36-58 */
KDEV float xs_gat(const KDenoise &p, float u)
{
    if (u >= 0.0f)
        return p.gat_fwd * sqrtf(p.gat_a * u + p.gat_c);
    return p.gat_t0 + u / p.gat_sc;
}

KDEV float xs_igat(const KDenoise &p, float t)
{
    if (t >= p.gat_t0)
        return p.gat_inv * t * t - p.gat_off;
    return (t - p.gat_t0) * p.gat_sc;
}

struct SmXs {
    float plane[XS_LH][XS_LW];
};

/* The weight the reference gives a pair at patch distance d2. */
KDEV float xs_weight(float d2, float inv_h2)
{
    return craw_dm_expf_negmax0(d2 * (1.0f / 9.0f) * inv_h2 - 2.0f);
}

/* A lane's running state down its column. The squared differences of
 * three rows live in three arrays whose roles rotate (the row loop is
 * unrolled by three, so the rotation costs no moves); the sums of the four
 * output rows a weight row feeds are carried as partial sums: output row
 * r + k (k = 0..3) has received its terms from weight rows r - 3 + k .. r - 1,
 * in the reference's raster order. */
struct XsLane {
    float a[4], ws[4];          /* acc and wsum of output rows r .. r + 3 */
};

template <int MODE>
struct XsBand {
    SmXs &sm;
    const KFrame &f;
    int lane, lc, x, y0, px, py;
    bool out_lane;
    float black, range, inv_range, inv_h2;

    /* Weight row r (LDS row r): fills sc with the squared differences of
     * row r + 1, forms the weights W_o(x, r) from rows r - 1 (sa), r (sb) and
     * r + 1, adds every term they give to the four partial sums, and
     * completes output row r. */
    __attribute__((always_inline)) KDEV void step(float (&sa)[XS_NO], float (&sb)[XS_NO],
                                                  float (&sc)[XS_NO], int r, XsLane &L)
    {
        float w[XS_NO];

#pragma unroll
        for (int o = 0; o < XS_NO; o++) {
            const float e = sm.plane[r + 1][lc] - sm.plane[r + 1 + xs_dy(o)][lc + xs_dx(o)];
            float d2;

            sc[o] = e * e;
            /* raster order: row above, own row, row below; columns x - 1,
             * x, x + 1 */
            d2 = xs_shr(sa[o]) + sa[o];
            d2 = d2 + xs_shl(sa[o]);
            d2 = d2 + xs_shr(sb[o]);
            d2 = d2 + sb[o];
            d2 = d2 + xs_shl(sb[o]);
            d2 = d2 + xs_shr(sc[o]);
            d2 = d2 + sc[o];
            d2 = d2 + xs_shl(sc[o]);
            w[o] = xs_weight(d2, inv_h2);
        }

        /* Row r is neighbour row ndy of output row r - ndy. Term (ndx, ndy)
         * of an output at column x is weight W_o(x + ndx, r) with o =
         * (-ndx, -ndy), fetched from lane + ndx. */
        L.a[3] = 0.0f;
        L.ws[3] = 0.0f;
#pragma unroll
        for (int k = 3; k >= 0; k--) {
#pragma unroll
            for (int ndx = -3; ndx <= 3; ndx++) {
                if (k == 0 && ndx >= 0)
                    continue;
                const int o = xs_index(-ndx, k);
                const float wv = ndx == 0 ? w[o] :
                    __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(
                                                  (lane + ndx) << 2, __builtin_bit_cast(int, w[o])));
                const float q = sm.plane[r][lc + ndx];

                L.ws[k] += wv;
                L.a[k] += wv * q;
            }
        }
        /* output row r: the centre (distance 0, weight exp(-0) = 1), then
         * the offsets after it with this lane's own weights */
        {
            const float wv = xs_weight(0.0f, inv_h2);
            L.ws[0] += wv;
            L.a[0] += wv * sm.plane[r][lc];
        }
#pragma unroll
        for (int o = 0; o < XS_NO; o++) {
            const float q = sm.plane[r + xs_dy(o)][lc + xs_dx(o)];

            L.ws[0] += w[o];
            L.a[0] += w[o] * q;
        }

        if (r >= XS_AY) {
            const int y = y0 - XS_AY + r;

            if (out_lane && y < f.dn.h2) {
                const KDenoise &p = f.dn;
                float u, raw0, u0, gr, raw;

                /* highlight guard and the 16-bit re-encode
                 * (lib/ref/ref_denoise.c:298-323) */
                u = xs_igat(p, L.a[0] / L.ws[0]);
                raw0 = (float)f.dn_src[(size_t)(2 * y + py) * p.src_pitch + 2 * x + px] *
                       (1.0f / 65535.0f);
                u0 = (raw0 - black) * inv_range;
                gr = (u0 - p.guard_lo) * p.guard_inv;
                gr = gr > 0.0f ? gr : 0.0f;
                gr = gr < 1.0f ? gr : 1.0f;
                u = u + (u0 - u) * gr;
                raw = black + u * range;
                raw = raw > 0.0f ? raw : 0.0f;
                raw = raw < 1.0f ? raw : 1.0f;
                f.dn_dst[(size_t)(2 * y + py) * p.dst_pitch + 2 * x + px] =
                    (uint16_t)rintf(raw * 65535.0f);
            }
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            L.a[k] = L.a[k + 1];
            L.ws[k] = L.ws[k + 1];
        }
    }
};

template <int MODE>
KDEV void dev_denoise_sym(SmXs &sm, const KFrame &f, int x0, int y0, int pl)
{
    constexpr bool TEMPORAL = MODE & KDN_TEMPORAL;
    static_assert((MODE & KDN_SPATIAL) && !(MODE & KDN_HISTORY), "symmetric weights only");
    const KDenoise &p = f.dn;
    const KDenoiseState &st = f.st;
    const uint16_t *__restrict__ src = f.dn_src;
    const int tid = threadIdx.x;
    const int px = pl & 1, py = pl >> 1;
    const int w2 = p.w2, h2 = p.h2;
    const float black = ksel4(pl, p.black[0], p.black[1], p.black[2], p.black[3]);
    const float range = ksel4(pl, p.range[0], p.range[1], p.range[2], p.range[3]);
    const float inv_range = ksel4(pl, p.inv_range[0], p.inv_range[1], p.inv_range[2],
                                  p.inv_range[3]);
    const size_t psize = (size_t)w2 * h2;
    const bool interior = x0 >= XS_AX && y0 >= XS_AY && x0 + XS_TW + XS_AX <= w2 &&
                          y0 + XS_TH + XS_AY <= h2;

    /* The stabilised plane over the band and its reach, through the plane
     * mirror; a temporal pass without history also starts the state of
     * the tile's own samples from this frame. */
    for (int i = tid; i < XS_LW * XS_LH; i += XS_NT) {
        const int lx = i % XS_LW, ly = i / XS_LW;
        const int vx = x0 - XS_AX + lx, vy = y0 - XS_AY + ly;
        const int mx = interior ? vx : kmirror(vx, w2), my = interior ? vy : kmirror(vy, h2);
        const float raw = (float)src[(size_t)(2 * my + py) * p.src_pitch + 2 * mx + px] *
                          (1.0f / 65535.0f);
        const float t = xs_gat(p, (raw - black) * inv_range);

        sm.plane[ly][lx] = t;
        if (TEMPORAL && vx == mx && vy == my && vx >= x0 && vx < x0 + XS_TW && vy >= y0 &&
            vy < y0 + XS_TH) {
            const size_t o = pl * psize + (size_t)vy * w2 + vx;
            st.next_level[o] = xs_igat(p, t);
            st.next_var[o] = 255;
        }
    }
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6;
    /* LDS column of this lane; plane column x0 - 4 + 56 wave + lane */
    const int lc = XS_AX - 4 + XS_OUTW * wave + lane;
    const int x = x0 - 4 + XS_OUTW * wave + lane;
    XsBand<MODE> band = { sm, f, lane, lc, x, y0, px, py,
                          lane >= 4 && lane < 4 + XS_OUTW && x < w2,
                          black, range, inv_range, p.inv_h2 };
    float s0[XS_NO], s1[XS_NO], s2[XS_NO];
    XsLane L;

    /* squared differences of the two rows above the first weight row
     * (LDS row 1, plane row y0 - 3) */
#pragma unroll
    for (int o = 0; o < XS_NO; o++) {
        float e = sm.plane[0][lc] - sm.plane[xs_dy(o)][lc + xs_dx(o)];
        s0[o] = e * e;
        e = sm.plane[1][lc] - sm.plane[1 + xs_dy(o)][lc + xs_dx(o)];
        s1[o] = e * e;
    }
#pragma unroll
    for (int k = 0; k < 4; k++)
        L.a[k] = L.ws[k] = 0.0f;

    /* weight rows 1 .. XS_TH + 3: the rows of the outputs and the three
     * above them */
    static_assert((XS_TH + 3) % 3 == 0, "the row loop is unrolled by three");
#pragma unroll 1
    for (int r = 1; r <= XS_TH + 3; r += 3) {
        band.step(s0, s1, s2, r, L);
        band.step(s1, s2, s0, r + 1, L);
        band.step(s2, s0, s1, r + 2, L);
    }
}

template <int MODE>
__global__ __launch_bounds__(XS_NT) void k_denoise_sym(KFrame f)
{
    __shared__ SmXs sm;
    dev_denoise_sym<MODE>(sm, f, blockIdx.x * XS_TW, blockIdx.y * XS_TH, blockIdx.z);
}

template <int MODE>
__global__ __launch_bounds__(XS_NT) void kb_denoise_sym(const KFrame *__restrict__ frames,
                                                        const uint32_t *__restrict__ tiles,
                                                        uint32_t ntiles)
{
    __shared__ SmXs sm;

    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t d = tiles[t];
        const KFrame &f = frames[KTILE_FRAME(d)];
        const int ntx = (f.dn.w2 + XS_TW - 1) / XS_TW, nty = (f.dn.h2 + XS_TH - 1) / XS_TH;
        const int i = KTILE_IDX(d), pl = i / (ntx * nty), r = i % (ntx * nty);

        __syncthreads();
        dev_denoise_sym<MODE>(sm, f, (r % ntx) * XS_TW, (r / ntx) * XS_TH, pl);
    }
}

hipError_t craw_launch_denoise_sym(int mode, const KFrame &f, hipStream_t s)
{
    dim3 grid((f.dn.w2 + XS_TW - 1) / XS_TW, (f.dn.h2 + XS_TH - 1) / XS_TH, 4);

    switch (mode) {
    case KDN_SPATIAL:
        KLAUNCH(k_denoise_sym<KDN_SPATIAL>, grid, dim3(XS_NT), 0, s, f);
        break;
    case KDN_SPATIAL | KDN_TEMPORAL:
        KLAUNCH(k_denoise_sym<KDN_SPATIAL | KDN_TEMPORAL>, grid, dim3(XS_NT), 0, s,
                           f);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t craw_launch_batch_denoise_sym(int mode, const KFrame *frames, const uint32_t *tiles,
                                         uint32_t ntiles, int grid, hipStream_t s)
{
    if (!ntiles)
        return hipSuccess;
    switch (mode) {
    case KDN_SPATIAL:
        KLAUNCH(kb_denoise_sym<KDN_SPATIAL>, dim3(grid), dim3(XS_NT), 0, s, frames,
                           tiles, ntiles);
        break;
    case KDN_SPATIAL | KDN_TEMPORAL:
        KLAUNCH(kb_denoise_sym<KDN_SPATIAL | KDN_TEMPORAL>, dim3(grid), dim3(XS_NT),
                           0, s, frames, tiles, ntiles);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
