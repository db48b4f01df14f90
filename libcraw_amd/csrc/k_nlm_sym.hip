/*
 * libcraw for MI355X: the exact spatial NLM with each patch distance
 * evaluated once for the pair of samples it joins.
 *
 * lib/ref/ref_denoise.c:114-154 of cyanreg/libcraw weighs neighbour q of
 * sample p by exp(-max(d2(p, q) / 9 / h^2 - 2, 0)), d2 being the 3x3 patch
 * distance summed in raster order; with a temporal history by
 * exp(-max(d2 / 9 / h^2 - (v_p + v_q), 0) / v_p), v being the variance the
 * temporal average leaves. The distance is symmetric bit for bit: q's
 * terms are p's negated differences, whose squares are equal, summed in
 * the same raster order; v_p + v_q = v_q + v_p as well. So the argument
 * A(p, q) = max(d2 / 9 / h^2 - v, 0) is one value for both samples, and
 * without history so is the weight. This kernel evaluates A once for every
 * pair, for the 24 offsets o after the centre in raster order (p = v,
 * q = v + o), and gives each sample its 48 neighbour weights in the
 * reference's raster order: offset o of p is the pair's weight for p,
 * offset -o of q the pair's weight for q. The sums are then the
 * reference's, with half the distances, and without history half the
 * exps.
 *
 * Layout: a workgroup's four waves stand side by side over a band of
 * 224 columns of the plane (48 rows, 30 with history), each lane one
 * column, and walk down it a row at a time. A lane keeps, per offset, the
 * squared differences of three rows in registers and forms the weights of
 * one row; the patch sums across columns come from the neighbouring lanes
 * through wave-wide DPP shifts, and the weight of a neighbouring column
 * through ds_bpermute. Each weight row adds its terms at once to the four
 * output rows it feeds (its own and the three below), which carry their
 * sums as running partial sums: every output's terms still arrive in
 * raster order. Of a wave's 64 columns, 62 have weights and the middle 56
 * are outputs. The stabilised plane (and with history the variance) is
 * staged in LDS over the band plus 7 columns and 4 rows of reach.
 *
 * With history, the band's temporal average is formed first
 * (lib/ref/ref_denoise.c:82-105, 268-296): this frame and the previous
 * state, stabilised, are staged one sample further, every position of
 * the band and its reach gets its blend factor from the 3x3 comparison at
 * its mirrored position, its average and the variance it leaves, and the
 * band's own samples write the state for the next frame. The lane of p
 * forms A(p, q) and p's weight; the lane of q receives A (not the weight)
 * and forms its own. Each division by a variance is A times the lane's
 * double reciprocal of it (one a row), rounded once to float: the
 * correctly rounded float quotient the reference computes (xs_quot).
 *
 * The precision form (F16) walks the same band with the
 * same pairs: each position's difference for an offset is formed in FP32
 * (two stabilised samples, both large, their difference small) and packed
 * with the next offset's into a half2; squares, the patch sums (across
 * columns through the same DPP shifts, then down three rows of those
 * column sums) and the weight's argument are packed FP16, the weight the
 * hardware exp2, and the sums FP32. A pair's weight is still formed once,
 * for both of its samples: half the distances and half the exps of
 * evaluating every sample's 48 neighbours (with history, half the
 * distances: the argument is shared, each sample divides it by its own
 * variance). The temporal average's weight then uses the hardware exp
 * too: the form has no double arithmetic.
 */
#include "k_common.h"
#include "kernels.h"
#include "detmath_tab.h"

#define XS_NW    4                     /* waves per workgroup */
#define XS_NT    (64 * XS_NW)
#define XS_OUTW  56                    /* output columns per wave */
#define XS_TW    (XS_OUTW * XS_NW)     /* 224 */
#define XS_AX    7
#define XS_AY    4
#define XS_LW    (XS_TW + 2 * XS_AX)   /* 238 */
#define XS_NO    24                    /* offsets after the centre */

#define XS_NH    (XS_NO / 2)           /* packed pairs of offsets */

#define DPP_WAVE_SHR1 0x138
#define DPP_WAVE_SHL1 0x130

typedef _Float16 xh2 __attribute__((ext_vector_type(2)));

/* Band rows and the LDS it stages: without history the plane; with it the
 * plane and the variance, which share the LDS with the staged frames the
 * temporal average is formed from. */
template <int MODE, bool TALL = false, bool F16 = false>
struct XsG {
    static constexpr bool HIST = MODE & KDN_HISTORY;
    /* TALL: the batched form's taller band without history, which its
     * autotuner may pick (fewer band rows of halo, fewer workgroups). With
     * history the precision form's band is shorter than the exact form's:
     * its LDS then lets three workgroups share a CU (its registers allow
     * three waves a SIMD; the exact form's allow two) */
    static constexpr int TH = HIST ? (F16 ? KG_XH_TH16 : KG_XH_TH)
                                   : (TALL ? KG_XS_TH_TALL : KG_XS_TH);
    static constexpr int LH = TH + 2 * XS_AY;
    /* Entries of the exp table the exact form stages in LDS (the precision
     * form uses the hardware exp). With history none: its 48 exps a row
     * with their table reads in flight took the kernel from 220 VGPRs to
     * 256 and a spill. */
    static constexpr int TABN = (F16 || HIST) ? 0 : 1024;
    static constexpr int TABS = TABN ? TABN : 1;
};
static_assert(XS_LW <= XS_NT, "a thread per staged column");

template <int MODE, bool TALL = false, bool F16 = false,
          bool HIST = XsG<MODE, TALL, F16>::HIST>
union SmXs;

template <int MODE, bool TALL, bool F16>
union SmXs<MODE, TALL, F16, false> {
    struct {
        float plane[XsG<MODE, TALL, F16>::LH][XS_LW];
    } m;
};

template <int MODE, bool TALL, bool F16>
union SmXs<MODE, TALL, F16, true> {
    struct {
        float plane[XsG<MODE, TALL, F16>::LH][XS_LW];
        float var[XsG<MODE, TALL, F16>::LH][XS_LW];
    } m;
    struct {
        float cur[XsG<MODE, TALL, F16>::LH + 2][XS_LW + 2];
        float prev[XsG<MODE, TALL, F16>::LH + 2][XS_LW + 2];
    } f;
};

static_assert(XS_TW == KG_XS_TW, "tile geometry");
static_assert((KG_XS_TH + 3) % 3 == 0 && (KG_XS_TH_TALL + 3) % 3 == 0 &&
              (KG_XH_TH + 3) % 3 == 0 && (KG_XH_TH16 + 3) % 3 == 0,
              "the row loop is unrolled by three");

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

KDEV xh2 xs_shr16(xh2 v)
{
    return __builtin_bit_cast(xh2, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                               DPP_WAVE_SHR1, 0xf, 0xf, true));
}

KDEV xh2 xs_shl16(xh2 v)
{
    return __builtin_bit_cast(xh2, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
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

/* exp(-max(a, 0)) as the oracle's craw_dm_expf_negmax0 rounds it, through
 * the LDS copy of the table of N entries. */
template <int N>
KDEV float xs_exp(float a, const double *tab)
{
    if constexpr (N == 1024)
        return craw_dm_expf_negmax0_t1024(a, tab);
    else if constexpr (N == 256)
        return craw_dm_expf_negmax0_t256(a, tab);
    else
        return craw_dm_expf_negmax0(a);
}

/* The weight the reference gives a pair at patch distance d2. */
template <int N>
KDEV float xs_weight(float d2, float inv_h2, const double *tab)
{
    return xs_exp<N>(d2 * (1.0f / 9.0f) * inv_h2 - 2.0f, tab);
}

/* A / v rounded to float, for A >= 0 finite and v > 0, given r = 1 / v
 * correctly rounded to double: the product is within 2^-52 of the
 * quotient, and a quotient of two floats lies at least 2^-49 (relative)
 * from any midpoint of the float grid, so rounding the product gives the
 * correctly rounded float quotient. */
KDEV float xs_quot(float a, double r)
{
    return (float)((double)a * r);
}

/* The workgroup's plane constants, selected by the plane index (the same
 * in every lane), held in scalar registers: in vector registers the
 * precision form's history band spilled them, and reloaded them in its
 * row loop. */
KDEV float xs_uniform(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

/* Sample i of a mosaic row, through a buffer resource on the row (in
 * scalar registers), so that a lane holds only its 32-bit offset rather
 * than a 64-bit pointer (which the same band spilled). */
KDEV uint32_t xs_row_load(const uint16_t *row, int i)
{
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        (void *)row, (short)0, 0x7fffffff, 0x00020000);

    return __builtin_amdgcn_raw_buffer_load_b16(r, 2 * i, 0, 0);
}

/* A lane's running state down its column. The squared differences of
 * three rows live in three arrays whose roles rotate (the row loop is
 * unrolled by three, so the rotation costs no moves); the sums of the four
 * output rows a weight row feeds are carried as partial sums: output row
 * r + k (k = 0..3) has received its terms from weight rows r - 3 + k .. r - 1,
 * in the reference's raster order. */
/* Band rows ahead whose raw samples (the guard's input at the output) a
 * lane loads before it needs them: each output row otherwise waited on its
 * load (the exact NLM 0.586 -> 0.567 ms at 5760x3240, the precision form
 * 0.288 -> 0.272). The precision form's history band loads one row ahead
 * (three would spill its registers): 0.635 -> 0.612 ms a frame with MHC. */
#define XS_RQ 3

struct XsLane {
    float a[4], ws[4];          /* acc and wsum of output rows r .. r + 3 */
    uint32_t rq[XS_RQ];         /* raw samples of band rows r .. r + XS_RQ - 1 */
    double rv[4];               /* with history: 1 / v of rows r .. r + 3 */
    float iv[4];                /* the same in float (precision form) */
};

template <int MODE, bool F16 = false, bool TALL = false>
struct XsBand {
    static constexpr bool HIST = XsG<MODE, TALL>::HIST;
    SmXs<MODE, TALL, F16> &sm;
    const KFrame &f;
    int lane, lc, x, y0, px, py;
    bool out_lane;
    float black, range, inv_range, inv_h2;
    const double *tab;                  /* the exp table in LDS */
    static constexpr int TABN = XsG<MODE, TALL, F16>::TABN;
    static constexpr bool RQ = true;            /* raw samples loaded ahead, */
    static constexpr int RQN = (F16 && HIST) ? 1 : XS_RQ;   /* RQN rows */

    /* The raw mosaic sample at this lane's column of band row rr, the
     * guard's input, clamped into the plane (a clamped position's value
     * is never used). */
    KDEV uint32_t raw_at(int rr) const
    {
        int y = y0 - XS_AY + rr, xc = x;

        y = y < f.dn.h2 ? y : f.dn.h2 - 1;
        y = y > 0 ? y : 0;
        xc = xc < f.dn.w2 ? xc : f.dn.w2 - 1;
        xc = xc > 0 ? xc : 0;
        return xs_row_load(f.dn_src + (size_t)(2 * y + py) * f.dn.src_pitch, 2 * xc + px);
    }

    /* Weight row r (LDS row r): fills sc with the squared differences of
     * row r + 1, forms the pair weights of row r from the squared
     * differences of rows r - 1 (sa), r (sb) and r + 1, adds every term
     * they give to the four partial sums, and completes output row r. */
    __attribute__((always_inline)) KDEV void step(float (&sa)[XS_NO], float (&sb)[XS_NO],
                                                  float (&sc)[XS_NO], int r, XsLane &L)
    {
        auto &plane = sm.m.plane;
        /* Without history, w[o]: the weight of the pair (x, x + o) in both
         * sums. With it, w[o] is the pair's weight in this lane's sum and
         * A[o] the pair's argument, from which the lane of x + o forms the
         * weight in its own sum with its own variance. */
        float w[XS_NO], A[XS_NO];

        if constexpr (HIST)
            L.rv[3] = 1.0 / (double)sm.m.var[r + 3][lc];
        float dd[XS_NO];

        /* The patch distances, three offsets' sums interleaved step by
         * step: a DPP operation must not read a register the instruction
         * just before it wrote, and the interleaving keeps two others
         * between them instead of a wait. Each sum is still in raster
         * order: row above, own row, row below; columns x - 1, x, x + 1. */
#pragma unroll
        for (int o = 0; o < XS_NO; o++) {
            const float e = plane[r + 1][lc] - plane[r + 1 + xs_dy(o)][lc + xs_dx(o)];
            sc[o] = e * e;
        }
#pragma unroll
        for (int g = 0; g < XS_NO; g += 3) {
#pragma unroll
            for (int j = 0; j < 3; j++)
                dd[g + j] = xs_shr(sa[g + j]) + sa[g + j];
#pragma unroll
            for (int j = 0; j < 3; j++)
                dd[g + j] = dd[g + j] + xs_shl(sa[g + j]);
#pragma unroll
            for (int j = 0; j < 3; j++)
                dd[g + j] = dd[g + j] + xs_shr(sb[g + j]);
#pragma unroll
            for (int j = 0; j < 3; j++)
                dd[g + j] = dd[g + j] + sb[g + j];
#pragma unroll
            for (int j = 0; j < 3; j++)
                dd[g + j] = dd[g + j] + xs_shl(sb[g + j]);
#pragma unroll
            for (int j = 0; j < 3; j++)
                dd[g + j] = dd[g + j] + xs_shr(sc[g + j]);
#pragma unroll
            for (int j = 0; j < 3; j++)
                dd[g + j] = dd[g + j] + sc[g + j];
#pragma unroll
            for (int j = 0; j < 3; j++)
                dd[g + j] = dd[g + j] + xs_shl(sc[g + j]);
        }
#pragma unroll
        for (int o = 0; o < XS_NO; o++) {
            const float d2 = dd[o];

            if constexpr (HIST) {
                /* lib/ref/ref_denoise.c:138-145 with v2 = v_p + v_q; the
                 * argument clamped below at 0 (a NaN to 0, as the
                 * reference's comparison does) and above at FLT_MAX,
                 * where either quotient still gives the weight 0 */
                const float vp = sm.m.var[r][lc];
                const float vq = sm.m.var[r + xs_dy(o)][lc + xs_dx(o)];
                const float arg = d2 * (1.0f / 9.0f) * inv_h2 - (vp + vq);

                A[o] = __builtin_fminf(__builtin_fmaxf(arg, 0.0f), 3.40282347e38f);
                w[o] = xs_exp<TABN>(xs_quot(A[o], L.rv[0]), tab);
            } else {
                w[o] = xs_weight<TABN>(d2, inv_h2, tab);
            }
        }

        terms(w, A, r, L);
    }

    /* Adds weight row r's terms (w, and with history the arguments A) to
     * the four output rows it feeds, completes output row r and moves the
     * partial sums up a row. */
    __attribute__((always_inline)) KDEV void terms(const float (&w)[XS_NO],
                                                   const float (&A)[XS_NO], int r, XsLane &L)
    {
        auto &plane = sm.m.plane;

        /* Row r is neighbour row ndy of output row r - ndy. Term (ndx, ndy)
         * of an output at column X is the weight of the pair (X + ndx, X)
         * in X's sum, formed from what lane X + ndx holds for o = (-ndx,
         * -ndy). */
        L.a[3] = 0.0f;
        L.ws[3] = 0.0f;
#pragma unroll
        for (int k = 3; k >= 0; k--) {
#pragma unroll
            for (int ndx = -3; ndx <= 3; ndx++) {
                if (k == 0 && ndx >= 0)
                    continue;
                const int o = xs_index(-ndx, k);
                const float h = HIST ? A[o] : w[o];
                const float hv = ndx == 0 ? h :
                    __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(
                                                  (lane + ndx) << 2, __builtin_bit_cast(int, h)));
                /* with history the lane's own weight from the pair's
                 * argument: exp(-A / v) (in the precision form -A log2 e
                 * is what travels, and the weight is exp2 of it / v) */
                const float wv = !HIST ? hv :
                    (F16 ? __builtin_amdgcn_exp2f(hv * L.iv[k])
                         : xs_exp<TABN>(xs_quot(hv, L.rv[k]), tab));
                const float q = plane[r][lc + ndx];

                L.ws[k] += wv;
                L.a[k] = F16 ? __builtin_fmaf(wv, q, L.a[k]) : L.a[k] + wv * q;
            }
        }
        /* output row r: the centre (distance 0: the argument is at most 0,
         * the weight exp(-0) = 1), then the offsets after it with this
         * lane's own weights */
        {
            /* exp(-max(0 / 9 / h^2 - 2, 0)) = exp(-0) = 1 for the finite
             * positive 1 / h^2 the host forms */
            L.ws[0] += 1.0f;
            L.a[0] += 1.0f * plane[r][lc];
        }
#pragma unroll
        for (int o = 0; o < XS_NO; o++) {
            const float q = plane[r + xs_dy(o)][lc + xs_dx(o)];

            L.ws[0] += w[o];
            L.a[0] = F16 ? __builtin_fmaf(w[o], q, L.a[0]) : L.a[0] + w[o] * q;
        }

        if (r >= XS_AY) {
            const int y = y0 - XS_AY + r;

            if (out_lane && y < f.dn.h2) {
                const KDenoise &p = f.dn;
                float u, raw0, u0, gr, raw;

                /* highlight guard and the 16-bit re-encode
                 * (lib/ref/ref_denoise.c:298-323) */
                u = xs_igat(p, L.a[0] / L.ws[0]);
                raw0 = (float)(RQ ? L.rq[0]
                                  : xs_row_load(f.dn_src + (size_t)(2 * y + py) * p.src_pitch,
                                                2 * x + px)) *
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
        if constexpr (RQ) {
#pragma unroll
            for (int k = 0; k < RQN - 1; k++)
                L.rq[k] = L.rq[k + 1];
            L.rq[RQN - 1] = raw_at(r + RQN);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            L.a[k] = L.a[k + 1];
            L.ws[k] = L.ws[k + 1];
            if constexpr (HIST) {
                L.rv[k] = L.rv[k + 1];
                L.iv[k] = L.iv[k + 1];
            }
        }
    }

    /* The horizontal patch sums of LDS row r's squared differences for
     * every offset, packed by pairs of offsets (precision form). */
    __attribute__((always_inline)) KDEV void hrow16(int r, xh2 (&h)[XS_NH])
    {
        auto &plane = sm.m.plane;
        const float cen = plane[r][lc];
        xh2 sq[XS_NH];

#pragma unroll
        for (int j = 0; j < XS_NH; j++) {
            const float e0 = cen - plane[r + xs_dy(2 * j)][lc + xs_dx(2 * j)];
            const float e1 = cen - plane[r + xs_dy(2 * j + 1)][lc + xs_dx(2 * j + 1)];
            const xh2 e = __builtin_bit_cast(xh2, __builtin_amdgcn_cvt_pkrtz(e0, e1));

            sq[j] = e * e;
        }
#pragma unroll
        for (int j = 0; j < XS_NH; j++)
            h[j] = (xs_shr16(sq[j]) + sq[j]) + xs_shl16(sq[j]);
    }

    /* Weight row r of the precision form: the column sums of row r + 1
     * into hc, the patch distances from those of rows r - 1 (ha), r (hb)
     * and r + 1, the argument d2 / 9 / h^2 - 2 clamped at 0 (times log2 e)
     * in FP16, the weight exp2 of it in FP32. */
    __attribute__((always_inline)) KDEV void step16(const xh2 (&ha)[XS_NH],
                                                    const xh2 (&hb)[XS_NH], xh2 (&hc)[XS_NH],
                                                    int r, XsLane &L)
    {
        const float l2e = 1.44269504088896341f;
        const xh2 kk = { (_Float16)(-(1.0f / 9.0f) * inv_h2 * l2e),
                         (_Float16)(-(1.0f / 9.0f) * inv_h2 * l2e) };
        const xh2 c2 = { (_Float16)(2.0f * l2e), (_Float16)(2.0f * l2e) };
        const xh2 zero = { (_Float16)0.0f, (_Float16)0.0f };
        float w[XS_NO], A[XS_NO];

        if constexpr (HIST)
            L.iv[3] = 1.0f / sm.m.var[r + 3][lc];
        hrow16(r + 1, hc);
#pragma unroll
        for (int j = 0; j < XS_NH; j++) {
            xh2 d2 = (ha[j] + hb[j]) + hc[j];

            if constexpr (HIST) {
                /* the argument with the two samples' variances in FP32:
                 * t = min(d2 kd + (v_p + v_q) log2 e, 0) = -A log2 e, the
                 * weight exp2(t / v_p) here and exp2(t / v_q) in q's lane */
                const float kd = -(1.0f / 9.0f) * inv_h2 * l2e;
                const float vp = sm.m.var[r][lc];
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int o = 2 * j + h;
                    const float vq = sm.m.var[r + xs_dy(o)][lc + xs_dx(o)];
                    const float t = __builtin_fminf(
                        __builtin_fmaf((float)(h ? d2.y : d2.x), kd, (vp + vq) * l2e), 0.0f);

                    A[o] = t;
                    w[o] = __builtin_amdgcn_exp2f(t * L.iv[0]);
                }
            } else {
                d2 = __builtin_elementwise_min(__builtin_elementwise_fma(d2, kk, c2), zero);
                w[2 * j] = __builtin_amdgcn_exp2f((float)d2.x);
                w[2 * j + 1] = __builtin_amdgcn_exp2f((float)d2.y);
            }
        }
        terms(w, A, r, L);
    }
};

/* The variance's state encoding: its square root in eight bits
 * (lib/ref/ref_denoise.c:100-105). */
KDEV uint8_t xs_var_code(float v)
{
    v = v > 0.0f ? v : 0.0f;
    v = v < 1.0f ? v : 1.0f;
    return (uint8_t)rintf(sqrtf(v) * 255.0f);
}

KDEV float xs_var_decode(uint8_t c)
{
    float q = (float)c * (1.0f / 255.0f);
    return q * q;
}

template <int MODE, bool F16, bool TALL>
KDEV void dev_denoise_sym(SmXs<MODE, TALL, F16> &sm, double *tab, const KFrame &f, int x0,
                          int y0, int pl)
{
    using G = XsG<MODE, TALL, F16>;
    constexpr bool TEMPORAL = MODE & KDN_TEMPORAL;
    constexpr bool HIST = G::HIST;
    constexpr int TH = G::TH, LH = G::LH;
    static_assert(MODE & KDN_SPATIAL, "the spatial modes");
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
                          y0 + TH + XS_AY <= h2;

    /* the exact form's exp table into LDS; the staging's barrier below
     * comes before its first read */
    if constexpr (G::TABN == 1024) {
#pragma unroll
        for (int i = 0; i < 1024 / XS_NT; i++)
            tab[tid + i * XS_NT] = craw_dm_exp2_tab1024[tid + i * XS_NT];
    } else if constexpr (G::TABN == 256) {
        static_assert(XS_NT == 256, "a thread per entry");
        tab[tid] = craw_dm_exp2_tab256[tid];
    }

#define RAW(X, Y) ((float)src[(size_t)(2 * (Y) + py) * p.src_pitch + 2 * (X) + px] *          \
                   (1.0f / 65535.0f))

    if constexpr (HIST) {
        /* This frame and the previous state, stabilised, one sample beyond
         * the band's reach; staged value i of a row is position
         * x0 - 8 + i through the plane mirror. */
        constexpr int SW = XS_LW + 2, SH = LH + 2;

        for (int i = tid; i < SW * SH; i += XS_NT) {
            const int lx = i % SW, ly = i / SW;
            const int mx = kmirror(x0 - XS_AX - 1 + lx, w2), my = kmirror(y0 - XS_AY - 1 + ly, h2);

            sm.f.cur[ly][lx] = xs_gat(p, (RAW(mx, my) - black) * inv_range);
            sm.f.prev[ly][lx] = xs_gat(p, st.prev_level[pl * psize + (size_t)my * w2 + mx]);
        }
        __syncthreads();

        /* The temporal average and the variance it leaves at every
         * position of the band and its reach (lib/ref/ref_denoise.c:82-105,
         * 268-296), each from the 3x3 comparison at its mirrored position;
         * thread lx takes column lx, down the rows. A position more than 3
         * columns right of the plane or 3 rows below it reaches no output:
         * it is formed at the nearest that does, which keeps every staged
         * read inside the staged window. */
        float tv[LH], vv[LH];
        const int lx = tid;
        const int vx = x0 - XS_AX + lx;
        const int mx = kmirror(vx < w2 + 3 ? vx : w2 + 3, w2);
        /* (mx, my) in the staged window */
        const int cx = mx - (x0 - XS_AX - 1);

#pragma unroll
        for (int ly = 0; ly < LH; ly++) {
            const int vy = y0 - XS_AY + ly;
            const int my = kmirror(vy < h2 + 3 ? vy : h2 + 3, h2);
            const int cy = my - (y0 - XS_AY - 1);

            tv[ly] = vv[ly] = 0.0f;
            if (lx < XS_LW) {
                float d2 = 0.0f, a, t, v;

                for (int ky = -1; ky <= 1; ky++) {
#pragma unroll
                    for (int kx = -1; kx <= 1; kx++) {
                        const float e = sm.f.prev[cy + ky][cx + kx] - sm.f.cur[cy + ky][cx + kx];
                        d2 += e * e;
                    }
                }
                if constexpr (F16)
                    a = p.alpha_max * kf_exp(-__builtin_fmaxf(d2 * (1.0f / 9.0f) * p.inv_h2 -
                                                              2.0f, 0.0f));
                else
                    a = p.alpha_max * xs_exp<G::TABN>(d2 * (1.0f / 9.0f) * p.inv_h2 - 2.0f,
                                                      tab);
                t = sm.f.cur[cy][cx] * (1.0f - a) + sm.f.prev[cy][cx] * a;
                v = (1.0f - a) * (1.0f - a) +
                    a * a * xs_var_decode(st.prev_var[pl * psize + (size_t)my * w2 + mx]);
                v = v > 1.0e-4f ? v : 1.0e-4f;
                tv[ly] = t;
                vv[ly] = v;
                if (vx == mx && vy == my && vx >= x0 && vx < x0 + XS_TW && vy >= y0 &&
                    vy < y0 + TH) {
                    const size_t o = pl * psize + (size_t)vy * w2 + vx;
                    st.next_level[o] = xs_igat(p, t);
                    st.next_var[o] = xs_var_code(v);
                }
            }
        }
        __syncthreads();
        if (lx < XS_LW) {
#pragma unroll
            for (int ly = 0; ly < LH; ly++) {
                sm.m.plane[ly][lx] = tv[ly];
                sm.m.var[ly][lx] = vv[ly];
            }
        }
    } else {
        /* The stabilised plane over the band and its reach, through the
         * plane mirror; a temporal pass without history also starts the
         * state of the band's own samples from this frame. */
        for (int i = tid; i < XS_LW * LH; i += XS_NT) {
            const int lx = i % XS_LW, ly = i / XS_LW;
            const int vx = x0 - XS_AX + lx, vy = y0 - XS_AY + ly;
            const int mx = interior ? vx : kmirror(vx, w2), my = interior ? vy : kmirror(vy, h2);
            const float t = xs_gat(p, (RAW(mx, my) - black) * inv_range);

            sm.m.plane[ly][lx] = t;
            if (TEMPORAL && vx == mx && vy == my && vx >= x0 && vx < x0 + XS_TW && vy >= y0 &&
                vy < y0 + TH) {
                const size_t o = pl * psize + (size_t)vy * w2 + vx;
                st.next_level[o] = xs_igat(p, t);
                st.next_var[o] = 255;
            }
        }
    }
    __syncthreads();
#undef RAW

    const int lane = tid & 63, wave = tid >> 6;
    /* LDS column of this lane; plane column x0 - 4 + 56 wave + lane */
    const int lc = XS_AX - 4 + XS_OUTW * wave + lane;
    const int x = x0 - 4 + XS_OUTW * wave + lane;
    XsBand<MODE, F16, TALL> band = { sm, f, lane, lc, x, y0, px, py,
                               lane >= 4 && lane < 4 + XS_OUTW && x < w2,
                               xs_uniform(black), xs_uniform(range), xs_uniform(inv_range),
                               p.inv_h2, tab };
    XsLane L;

    if constexpr (XsBand<MODE, F16, TALL>::RQ) {
#pragma unroll
        for (int k = 0; k < XsBand<MODE, F16, TALL>::RQN; k++)
            L.rq[k] = band.raw_at(1 + k);
    }
    if constexpr (F16) {
        xh2 h0[XS_NH], h1[XS_NH], h2[XS_NH];

        /* the column sums of the two rows above the first weight row */
        band.hrow16(0, h0);
        band.hrow16(1, h1);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            L.a[k] = L.ws[k] = 0.0f;
            L.rv[k] = 0.0;
            L.iv[k] = 0.0f;
        }
        if constexpr (HIST) {
#pragma unroll
            for (int k = 0; k < 3; k++)
                L.iv[k] = 1.0f / sm.m.var[1 + k][lc];
        }
#pragma unroll 1
        for (int r = 1; r <= TH + 3; r += 3) {
            band.step16(h0, h1, h2, r, L);
            band.step16(h1, h2, h0, r + 1, L);
            band.step16(h2, h0, h1, r + 2, L);
        }
        return;
    }
    float s0[XS_NO], s1[XS_NO], s2[XS_NO];

    /* squared differences of the two rows above the first weight row
     * (LDS row 1, plane row y0 - 3) */
#pragma unroll
    for (int o = 0; o < XS_NO; o++) {
        float e = sm.m.plane[0][lc] - sm.m.plane[xs_dy(o)][lc + xs_dx(o)];
        s0[o] = e * e;
        e = sm.m.plane[1][lc] - sm.m.plane[1 + xs_dy(o)][lc + xs_dx(o)];
        s1[o] = e * e;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        L.a[k] = L.ws[k] = 0.0f;
        L.rv[k] = 0.0;
    }
    /* the reciprocal variances of rows 1 .. 3, before the first weight row
     * forms that of row 4 */
    if constexpr (HIST) {
#pragma unroll
        for (int k = 0; k < 3; k++)
            L.rv[k] = 1.0 / (double)sm.m.var[1 + k][lc];
    }

    /* weight rows 1 .. TH + 3: the rows of the outputs and the three above
     * them */
#pragma unroll 1
    for (int r = 1; r <= TH + 3; r += 3) {
        band.step(s0, s1, s2, r, L);
        band.step(s1, s2, s0, r + 1, L);
        band.step(s2, s0, s1, r + 2, L);
    }
}

/* Waves a SIMD the kernels are compiled for (the register budget: 512 /
 * waves VGPRs a lane; the LDS of a 48-row band allows three). The exact
 * form at three spills 128 bytes a lane and runs 0.679 ms at 5760x3240
 * against 0.603 at two; the precision form fits three without a spill. */
#ifndef KG_XS_WAVES
#define KG_XS_WAVES 2
#endif
#ifndef KG_XH16_WAVES
#define KG_XH16_WAVES 3
#endif

template <int MODE, bool F16>
__global__ __launch_bounds__(XS_NT) __attribute__((amdgpu_waves_per_eu(
    F16 ? ((MODE & KDN_HISTORY) ? KG_XH16_WAVES : 3)
        : ((MODE & KDN_HISTORY) ? 2 : KG_XS_WAVES)))) void k_denoise_sym(KFrame f)
{
    constexpr int TH = XsG<MODE, false, F16>::TH;
    __shared__ SmXs<MODE, false, F16> sm;
    __shared__ double tab[XsG<MODE, false, F16>::TABS];
    const int ntx = (f.dn.w2 + XS_TW - 1) / XS_TW, nty = (f.dn.h2 + TH - 1) / TH;
    int t, pl;

    if (!craw_dn_pair_tile(blockIdx.x, ntx * nty, &t, &pl))
        return;
    dev_denoise_sym<MODE, F16, false>(sm, tab, f, (t % ntx) * XS_TW, (t / ntx) * TH, pl);
}

template <int MODE, bool F16, bool TALL>
__global__ __launch_bounds__(XS_NT) __attribute__((amdgpu_waves_per_eu(
    (F16 && !TALL) ? ((MODE & KDN_HISTORY) ? KG_XH16_WAVES : 3) : 1)))
void kb_denoise_sym(const KFrame *__restrict__ frames, const uint32_t *__restrict__ tiles,
                    uint32_t ntiles)
{
    constexpr int TH = XsG<MODE, TALL, F16>::TH;
    __shared__ SmXs<MODE, TALL, F16> sm;
    __shared__ double tab[XsG<MODE, TALL, F16>::TABS];

    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t d = tiles[t];
        const KFrame &f = frames[KTILE_FRAME(d)];
        const int ntx = (f.dn.w2 + XS_TW - 1) / XS_TW, nty = (f.dn.h2 + TH - 1) / TH;
        int r, pl;

        if (!craw_dn_pair_tile(KTILE_IDX(d), ntx * nty, &r, &pl))
            continue;
        __syncthreads();
        dev_denoise_sym<MODE, F16, TALL>(sm, tab, f, (r % ntx) * XS_TW, (r / ntx) * TH, pl);
    }
}

#define XS_MODES(X)                                                                        \
    X(KDN_SPATIAL) X(KDN_SPATIAL | KDN_TEMPORAL) X(KDN_SPATIAL | KDN_TEMPORAL | KDN_HISTORY)

hipError_t craw_launch_denoise_sym(int mode, const KFrame &f, hipStream_t s)
{
    switch (mode) {
#define X(M)                                                                               \
    case M: {                                                                              \
        const dim3 grid(craw_dn_pair_count(((f.dn.w2 + XS_TW - 1) / XS_TW) *                \
                                           ((f.dn.h2 + XsG<M>::TH - 1) / XsG<M>::TH)));    \
        KLAUNCH((k_denoise_sym<M, false>), grid, dim3(XS_NT), 0, s, f);                    \
        break;                                                                             \
    }
    XS_MODES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t craw_launch_batch_denoise_sym(int mode, const KFrame *frames, const uint32_t *tiles,
                                         uint32_t ntiles, int grid, bool tall, hipStream_t s)
{
    if (!ntiles)
        return hipSuccess;
    switch (mode) {
#define X(M)                                                                               \
    case M:                                                                                \
        if (tall && !((M) & KDN_HISTORY))                                                  \
            KLAUNCH((kb_denoise_sym<M, false, true>), dim3(grid), dim3(XS_NT), 0, s, frames, \
                    tiles, ntiles);                                                        \
        else                                                                               \
            KLAUNCH((kb_denoise_sym<M, false, false>), dim3(grid), dim3(XS_NT), 0, s, frames, \
                    tiles, ntiles);                                                        \
        break;
    XS_MODES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

/* The precision form's spatial modes (no history): packed-FP16 distances,
 * each pair's weight formed once. */
hipError_t craw_launch_denoise_sym16(int mode, const KFrame &f, hipStream_t s)
{
    switch (mode) {
#define X(M)                                                                               \
    case M: {                                                                              \
        const dim3 grid(craw_dn_pair_count(((f.dn.w2 + XS_TW - 1) / XS_TW) *                \
                                           ((f.dn.h2 + XsG<M, false, true>::TH - 1) /      \
                                            XsG<M, false, true>::TH)));                    \
        KLAUNCH((k_denoise_sym<M, true>), grid, dim3(XS_NT), 0, s, f);                     \
        break;                                                                             \
    }
    XS_MODES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t craw_launch_batch_denoise_sym16(int mode, const KFrame *frames, const uint32_t *tiles,
                                           uint32_t ntiles, int grid, bool tall, hipStream_t s)
{
    if (!ntiles)
        return hipSuccess;
    switch (mode) {
#define X(M)                                                                               \
    case M:                                                                                \
        if (tall && !((M) & KDN_HISTORY))                                                  \
            KLAUNCH((kb_denoise_sym<M, true, true>), dim3(grid), dim3(XS_NT), 0, s, frames, \
                    tiles, ntiles);                                                        \
        else                                                                               \
            KLAUNCH((kb_denoise_sym<M, true, false>), dim3(grid), dim3(XS_NT), 0, s, frames, \
                    tiles, ntiles);                                                        \
        break;
    XS_MODES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
