/*
 * libcraw for MI355X: the temporal filter of the raw-domain denoise alone
 * (no spatial NLM), as lib/ref/ref_denoise.c:156-336 of cyanreg/libcraw
 * defines it; the spatial modes are k_nlm_sym.hip's.
 *
 * One workgroup filters a 32x16 tile of one CFA plane (grid z = plane).
 * Without history the tile's stabilised samples are the output and start
 * the state. With a history the tile's current and previous stabilised
 * planes are staged in LDS one sample beyond it (the 3x3 comparison at the
 * mirrored position of each sample), averaged, and the average's state is
 * written for the next frame. Every operation is the reference's, so the
 * 16-bit output is bit for bit.
 */
#include "k_common.h"
#include "kernels.h"

#define DN_TW 32
#define DN_TH 16
#define DN_NT 256
#define DN_R  4       /* search radius 3 + patch radius 1 */

struct KGat {
    float a, c, sc, fwd, inv, off, t0;
};

/* lib/ref/ref_denoise.c:36-58 */
KDEV float kgat(const KDenoise &p, float u)
{
    if (u >= 0.0f)
        return p.gat_fwd * sqrtf(p.gat_a * u + p.gat_c);
    return p.gat_t0 + u / p.gat_sc;
}

KDEV float kigat(const KDenoise &p, float t)
{
    if (t >= p.gat_t0)
        return p.gat_inv * t * t - p.gat_off;
    return (t - p.gat_t0) * p.gat_sc;
}

/* The variance's state encoding: its square root in eight bits
 * (lib/ref/ref_denoise.c:100-105). */
KDEV uint8_t kvar_code(float v)
{
    v = v > 0.0f ? v : 0.0f;
    v = v < 1.0f ? v : 1.0f;
    return (uint8_t)rintf(sqrtf(v) * 255.0f);
}

KDEV float kvar_decode(uint8_t c)
{
    float q = (float)c * (1.0f / 255.0f);
    return q * q;
}

template <int MODE>
struct SmDn {
    static constexpr bool HISTORY = MODE & KDN_HISTORY;
    static constexpr int PW = DN_TW + 2 * DN_R, PH = DN_TH + 2 * DN_R;
    static constexpr int CW = PW + 2, CH = PH + 2;
    float splane[PH * PW];
    float scur[HISTORY ? CH * CW : 1];
    float stp[HISTORY ? CH * CW : 1];
};

template <int MODE>
KDEV void dev_denoise(SmDn<MODE> &sm, const KFrame &f, int x0, int y0, int pl)
{
    constexpr bool TEMPORAL = MODE & KDN_TEMPORAL;
    static_assert(!(MODE & KDN_SPATIAL), "the spatial modes are k_nlm_sym.hip's");
    constexpr bool HISTORY = MODE & KDN_HISTORY;
    constexpr int PW = SmDn<MODE>::PW;
    constexpr int CW = SmDn<MODE>::CW;
    float *splane = sm.splane, *scur = sm.scur, *stp = sm.stp;
    const KDenoise &p = f.dn;
    const KDenoiseState &st = f.st;
    const uint16_t *__restrict__ src = f.dn_src;
    uint16_t *__restrict__ dst = f.dn_dst;
    const int tid = threadIdx.x;
    const int px = pl & 1, py = pl >> 1;
    const int w2 = p.w2, h2 = p.h2;
    const int xe = min(x0 + DN_TW, w2), ye = min(y0 + DN_TH, h2);
    const int tw = xe - x0, th = ye - y0;
    const float black = ksel4(pl, p.black[0], p.black[1], p.black[2], p.black[3]);
    const float range = ksel4(pl, p.range[0], p.range[1], p.range[2], p.range[3]);
    const float inv_range = ksel4(pl, p.inv_range[0], p.inv_range[1], p.inv_range[2],
                                  p.inv_range[3]);
    const size_t psize = (size_t)w2 * h2;

#define RAW(X, Y) src[(size_t)(2 * (Y) + py) * p.src_pitch + 2 * (X) + px]

    if (HISTORY) {
        const int rw = tw + 2 * (DN_R + 1), rh = th + 2 * (DN_R + 1);

        /* This frame and the previous one, stabilised, one sample beyond
         * the NLM's reach. */
        for (int i = tid; i < rw * rh; i += DN_NT) {
            int lx = i % rw, ly = i / rw;
            int mx = kmirror(x0 - DN_R - 1 + lx, w2), my = kmirror(y0 - DN_R - 1 + ly, h2);
            float raw = (float)RAW(mx, my) * (1.0f / 65535.0f);

            scur[ly * CW + lx] = kgat(p, (raw - black) * inv_range);
            stp[ly * CW + lx] = kgat(p, st.prev_level[pl * psize + (size_t)my * w2 + mx]);
        }
        __syncthreads();

        /* The average and the variance it leaves (lib/ref/ref_denoise.c:82-105)
         * at the mirrored position of every position the filter reads. */
        {
            const int qw = tw + 2 * DN_R, qh = th + 2 * DN_R;

            for (int i = tid; i < qw * qh; i += DN_NT) {
                int lx = i % qw, ly = i / qw;
                int vx = x0 - DN_R + lx, vy = y0 - DN_R + ly;
                int mx = kmirror(vx, w2), my = kmirror(vy, h2);
                /* (mx, my) in the cur/prev tile */
                int cx = mx - (x0 - DN_R - 1), cy = my - (y0 - DN_R - 1);
                float d2 = 0.0f, a, t, pv, v;

                for (int ky = -1; ky <= 1; ky++) {
                    for (int kx = -1; kx <= 1; kx++) {
                        int j = (cy + ky) * CW + cx + kx;
                        float e = stp[j] - scur[j];
                        d2 += e * e;
                    }
                }
                {
                    float arg = d2 * (1.0f / 9.0f) * p.inv_h2 - 2.0f;
                    a = p.alpha_max * craw_dm_expf_nonpos(-(arg > 0.0f ? arg : 0.0f));
                }
                t = scur[cy * CW + cx] * (1.0f - a) + stp[cy * CW + cx] * a;
                pv = kvar_decode(st.prev_var[pl * psize + (size_t)my * w2 + mx]);
                v = (1.0f - a) * (1.0f - a) + a * a * pv;
                v = v > 1.0e-4f ? v : 1.0e-4f;
                splane[ly * PW + lx] = t;
                if (vx == mx && vy == my && vx >= x0 && vx < xe && vy >= y0 && vy < ye) {
                    size_t o = pl * psize + (size_t)vy * w2 + vx;
                    st.next_level[o] = kigat(p, t);
                    st.next_var[o] = kvar_code(v);
                }
            }
        }
        __syncthreads();
    } else {
        const int rw = tw + 2 * DN_R, rh = th + 2 * DN_R;

        for (int i = tid; i < rw * rh; i += DN_NT) {
            int lx = i % rw, ly = i / rw;
            int vx = x0 - DN_R + lx, vy = y0 - DN_R + ly;
            int mx = kmirror(vx, w2), my = kmirror(vy, h2);
            float raw = (float)RAW(mx, my) * (1.0f / 65535.0f);
            float t = kgat(p, (raw - black) * inv_range);

            splane[ly * PW + lx] = t;
            /* A temporal pass without history starts the state from this
             * frame, with a variance of 1. */
            if (TEMPORAL && vx == mx && vy == my && vx >= x0 && vx < xe && vy >= y0 && vy < ye) {
                size_t o = pl * psize + (size_t)vy * w2 + vx;
                st.next_level[o] = kigat(p, t);
                st.next_var[o] = 255;
            }
        }
        __syncthreads();
    }

    /* The tile's own samples. */
#pragma unroll 1
    for (int k = 0; k < DN_TW * DN_TH / DN_NT; k++) {
        int q = tid + k * DN_NT, lx = q % DN_TW, ly = q / DN_TW;
        int x = x0 + lx, y = y0 + ly;
        int c = (ly + DN_R) * PW + lx + DN_R;
        float u, raw0, u0, gr, raw;

        if (x >= w2 || y >= h2)
            continue;

        /* the temporal average alone */
        u = kigat(p, splane[c]);

        /* Highlight guard and the 16-bit re-encode
         * (lib/ref/ref_denoise.c:298-323). */
        raw0 = (float)RAW(x, y) * (1.0f / 65535.0f);
        u0 = (raw0 - black) * inv_range;
        gr = (u0 - p.guard_lo) * p.guard_inv;
        gr = gr > 0.0f ? gr : 0.0f;
        gr = gr < 1.0f ? gr : 1.0f;
        u = u + (u0 - u) * gr;
        raw = black + u * range;
        raw = raw > 0.0f ? raw : 0.0f;
        raw = raw < 1.0f ? raw : 1.0f;
        dst[(size_t)(2 * y + py) * p.dst_pitch + 2 * x + px] = (uint16_t)rintf(raw * 65535.0f);
    }
#undef RAW
}

template <int MODE>
__global__ __launch_bounds__(DN_NT) void k_denoise(KFrame f)
{
    __shared__ SmDn<MODE> sm;
    const int ntx = (f.dn.w2 + DN_TW - 1) / DN_TW, nty = (f.dn.h2 + DN_TH - 1) / DN_TH;
    int t, pl;

    if (!craw_dn_pair_tile(blockIdx.x, ntx * nty, &t, &pl))
        return;
    dev_denoise<MODE>(sm, f, (t % ntx) * DN_TW, (t / ntx) * DN_TH, pl);
}

/* Persistent form over a work-list of (frame, plane, tile) entries. */
template <int MODE>
__global__ __launch_bounds__(DN_NT) void kb_denoise(const KFrame *__restrict__ frames,
                                                    const uint32_t *__restrict__ tiles,
                                                    uint32_t ntiles)
{
    __shared__ SmDn<MODE> sm;

    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t d = tiles[t];
        const KFrame &f = frames[KTILE_FRAME(d)];
        const int ntx = (f.dn.w2 + DN_TW - 1) / DN_TW, nty = (f.dn.h2 + DN_TH - 1) / DN_TH;
        int r, pl;

        if (!craw_dn_pair_tile(KTILE_IDX(d), ntx * nty, &r, &pl))
            continue;
        __syncthreads();
        dev_denoise<MODE>(sm, f, (r % ntx) * DN_TW, (r / ntx) * DN_TH, pl);
    }
}

static_assert(DN_TW == KG_DN_TW && DN_TH == KG_DN_TH, "tile geometry");

/* the modes this kernel serves: the spatial ones go to k_nlm_sym.hip */
#define DN_MODES(X) X(KDN_TEMPORAL) X(KDN_TEMPORAL | KDN_HISTORY)

hipError_t craw_launch_denoise(int mode, const KFrame &f, hipStream_t s)
{
    dim3 grid(craw_dn_pair_count(((f.dn.w2 + DN_TW - 1) / DN_TW) *
                                 ((f.dn.h2 + DN_TH - 1) / DN_TH)));

    if (craw_dn_geom(KFORM_EXACT, mode) != KDN_GEOM_DIRECT)
        return craw_launch_denoise_sym(mode, f, s);

    switch (mode) {
#define X(M)                                                                               \
    case M:                                                                                \
        KLAUNCH(k_denoise<M>, grid, dim3(DN_NT), 0, s, f);                      \
        break;
    DN_MODES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t craw_launch_batch_denoise_exact(int mode, int shape, const KFrame *frames,
                                           const uint32_t *tiles, uint32_t ntiles, int grid,
                                           hipStream_t s)
{
    if (!ntiles)
        return hipSuccess;
    if (craw_dn_geom(KFORM_EXACT, mode) != KDN_GEOM_DIRECT)
        return craw_launch_batch_denoise_sym(mode, frames, tiles, ntiles, grid,
                                             (shape & KDN_SHAPE_TALL) != 0, s);
    switch (mode) {
#define X(M)                                                                               \
    case M:                                                                                \
        KLAUNCH(kb_denoise<M>, dim3(grid), dim3(DN_NT), 0, s, frames, tiles,     \
                           ntiles);                                                        \
        break;
    DN_MODES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
