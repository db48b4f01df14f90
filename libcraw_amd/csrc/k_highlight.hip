/*
 * libcraw for MI355X: highlight reconstruction (lib/ref/ref_highlight.c:
 * 91-335 of cyanreg/libcraw) before the debayer: a balanced, unclipped
 * mosaic rebuilt where the sensor clipped, stored as R16F, which the
 * debayer then reads with an identity prep.
 *
 *   u = (raw - black) / (white - black) a site (the range as a reciprocal),
 *   a = smoothstep(0.92, 0.98, u), its clip ramp.
 *
 * k_hl_field: the colour field's level 0, a texel per 4x4 block, and its
 * pull to levels 1-4, one workgroup per 32x16 texels (128x64 pixels). The
 * sites' u and each quad's largest a are staged in LDS, the quads' a with
 * the dilation's reach of 3 quads (quads past the frame count as 0, as the
 * reference skips them); the 7x7 maximum is a row pass then a column pass
 * (a maximum is exact in any order). A texel sums its quads top row first,
 * left to right: weight (1 - a_dilated) min(u_max, 1)^2, the weighted
 * clamped logs of R/G and B/G. The pull sums 2x2 texels (in the same
 * order) inside the tile: a tile of 32x16 at level 0 is whole 16x8, 8x4,
 * 4x2 and 2x1 tiles of levels 1-4.
 *
 * k_hl_coarse: one workgroup pulls levels 5 and up and pushes them back
 * down to level 4; k_hl_push pushes levels 3 to 0, a texel a thread: the
 * level's own colour where it has weight, blended by min(w, 1) towards the
 * bilinear of the pushed level above (the top level divided on the way).
 *
 * k_hl_fill: a 64x16 tile of sites with u staged over a 1-site mirror
 * reach; a site with a > 0 is raised towards the level its unclipped
 * neighbours and the field's colour predict, never lowered. Every site is
 * stored as R16F (round to nearest, ties to even).
 *
 * Every expression keeps the reference's order and its ternary min / max /
 * clamp; logf / expf are detmath.h's, as the oracle's deterministic mode.
 */
#include "k_common.h"
#include "kernels.h"

#define HL_U_LO    0.92f
#define HL_U_HI    0.98f
#define HL_EPS     1e-4f
#define HL_LOG_MAX 2.0f
#define HL_W_SAT   1.0f
#define HL_DILATE  3

/* the reference's utils.h ternaries */
#define HMAX(a, b) ((a) > (b) ? (a) : (b))
#define HMIN(a, b) ((a) < (b) ? (a) : (b))
#define HCLAMP(v, lo, hi) HMIN(HMAX((v), (lo)), (hi))

KDEV int hl_chan(int x, int y)
{
    const int s = (y & 1) * 2 + (x & 1);
    return s == 0 ? 0 : (s == 3 ? 2 : 1);
}

KDEV float hl_u(const KHighlight &h, int x, int y)
{
    const int s = (y & 1) * 2 + (x & 1);
    const float raw = (float)h.src[(size_t)y * h.src_pitch + x] * (1.0f / 65535.0f);

    return (raw - ksel4(s, h.black[0], h.black[1], h.black[2], h.black[3])) *
           ksel4(s, h.inv_range[0], h.inv_range[1], h.inv_range[2], h.inv_range[3]);
}

KDEV float hl_ramp(float u)
{
    float t = (u - HL_U_LO) / (HL_U_HI - HL_U_LO);

    t = HCLAMP(t, 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

/* log and exp: correctly rounded (detmath.h), or in the precision form
 * the hardware's float ones */
KDEV float hl_log(const KHighlight &h, float x)
{
    return h.float_math ? kf_log(x) : craw_dm_logf(x);
}

KDEV float hl_exp(const KHighlight &h, float x)
{
    return h.float_math ? kf_exp(x) : craw_dm_expf(x);
}

KDEV float hl_wb(const KHighlight &h, int c)
{
    return c == 0 ? h.wb[0] : (c == 1 ? h.wb[1] : h.wb[2]);
}

/* ------------------------------------------------------------ field */

#define HF_TX 32
#define HF_TY 16
#define HF_QX (2 * HF_TX)
#define HF_QY (2 * HF_TY)
#define HF_AX (HF_QX + 2 * HL_DILATE)
#define HF_AY (HF_QY + 2 * HL_DILATE)
#define HF_NT 256

struct SmField {
    float u[2 * HF_QY][2 * HF_QX];          /* the tile's sites */
    float qa[HF_AY][HF_AX];                 /* quad ramp maxima with the reach */
    float hm[HF_AY][HF_QX];                 /* their 7-wide row maxima */
    float4 l0[HF_TY][HF_TX];
    float4 l1[HF_TY / 2][HF_TX / 2];
    float4 l2[HF_TY / 4][HF_TX / 4];
    float4 l3[HF_TY / 8][HF_TX / 8];
};

/* Level l's texel (x, y) of the tile (origin ox, oy in that level) as the
 * 2x2 sum of level l - 1's, skipping those past the level. */
KDEV float4 hl_pull4(const float4 *fine, int fstride, int fw, int fh, int fox, int foy, int x,
                     int y)
{
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);

    for (int dy = 0; dy < 2; dy++)
        for (int dx = 0; dx < 2; dx++) {
            const int fx = 2 * x + dx, fy = 2 * y + dy;

            if (fox + fx < fw && foy + fy < fh) {
                const float4 v = fine[fy * fstride + fx];
                acc.x += v.x;
                acc.y += v.y;
                acc.z += v.z;
            }
        }
    return acc;
}

__global__ __launch_bounds__(HF_NT) void k_hl_field(KHighlight h)
{
    __shared__ SmField sm;
    const int tid = threadIdx.x;
    const int bx0 = blockIdx.x * HF_TX, by0 = blockIdx.y * HF_TY;
    const int qx0 = 2 * bx0, qy0 = 2 * by0, px0 = 2 * qx0, py0 = 2 * qy0;
    const int qw = h.W / 2, qh = h.H / 2;

    /* the tile's sites */
    for (int i = tid; i < 4 * HF_QX * HF_QY; i += HF_NT) {
        const int lx = i % (2 * HF_QX), ly = i / (2 * HF_QX);
        const int x = px0 + lx, y = py0 + ly;

        sm.u[ly][lx] = x < h.W && y < h.H ? hl_u(h, x, y) : 0.0f;
    }
    /* each quad's largest ramp, 0 past the frame */
    for (int i = tid; i < HF_AX * HF_AY; i += HF_NT) {
        const int lx = i % HF_AX, ly = i / HF_AX;
        const int qx = qx0 - HL_DILATE + lx, qy = qy0 - HL_DILATE + ly;
        float m = 0.0f;

        if (qx >= 0 && qy >= 0 && qx < qw && qy < qh)
            for (int s = 0; s < 4; s++)
                m = HMAX(m, hl_ramp(hl_u(h, 2 * qx + (s & 1), 2 * qy + (s >> 1))));
        sm.qa[ly][lx] = m;
    }
    __syncthreads();
    for (int i = tid; i < HF_QX * HF_AY; i += HF_NT) {
        const int lx = i % HF_QX, ly = i / HF_QX;
        float m = sm.qa[ly][lx];

        for (int k = 1; k <= 2 * HL_DILATE; k++)
            m = HMAX(m, sm.qa[ly][lx + k]);
        sm.hm[ly][lx] = m;
    }
    __syncthreads();

    /* level 0 */
    for (int t = tid; t < HF_TX * HF_TY; t += HF_NT) {
        const int tx = t % HF_TX, ty = t / HF_TX;
        float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);

        if (bx0 + tx >= h.lw[0] || by0 + ty >= h.lh[0])
            continue;
        for (int j = 0; j < 2; j++) {
            for (int i = 0; i < 2; i++) {
                const int lqx = 2 * tx + i, lqy = 2 * ty + j;
                float amax = 0.0f, umax = 0.0f, rgb[3] = { 0.0f, 0.0f, 0.0f }, m, wq, g, r, b;

                if (qx0 + lqx >= qw || qy0 + lqy >= qh)
                    continue;
                for (int k = 0; k <= 2 * HL_DILATE; k++)
                    amax = HMAX(amax, sm.hm[lqy + k][lqx]);
                for (int s = 0; s < 4; s++) {
                    const float v = sm.u[2 * lqy + (s >> 1)][2 * lqx + (s & 1)];
                    const int c = s == 0 ? 0 : (s == 3 ? 2 : 1);

                    rgb[c] += v * hl_wb(h, c);
                    umax = HMAX(umax, v);
                }
                rgb[1] *= 0.5f;
                m = HMIN(umax, 1.0f);
                wq = (1.0f - amax) * m * m;
                g = HMAX(rgb[1], HL_EPS);
                r = HCLAMP(hl_log(h, HMAX(rgb[0], HL_EPS) / g), -HL_LOG_MAX, HL_LOG_MAX);
                b = HCLAMP(hl_log(h, HMAX(rgb[2], HL_EPS) / g), -HL_LOG_MAX, HL_LOG_MAX);
                acc.x += wq * r;
                acc.y += wq * b;
                acc.z += wq;
            }
        }
        sm.l0[ty][tx] = acc;
        h.level[0][(size_t)(by0 + ty) * h.lw[0] + bx0 + tx] = acc;
    }
    __syncthreads();

    /* levels 1-4 of the tile */
    if (h.nl > 1) {
        for (int t = tid; t < (HF_TX / 2) * (HF_TY / 2); t += HF_NT) {
            const int x = t % (HF_TX / 2), y = t / (HF_TX / 2), ox = bx0 / 2, oy = by0 / 2;

            if (ox + x < h.lw[1] && oy + y < h.lh[1]) {
                const float4 v = hl_pull4(&sm.l0[0][0], HF_TX, h.lw[0], h.lh[0], bx0, by0, x, y);
                sm.l1[y][x] = v;
                h.level[1][(size_t)(oy + y) * h.lw[1] + ox + x] = v;
            }
        }
    }
    __syncthreads();
    if (h.nl > 2 && tid < (HF_TX / 4) * (HF_TY / 4)) {
        const int x = tid % (HF_TX / 4), y = tid / (HF_TX / 4), ox = bx0 / 4, oy = by0 / 4;

        if (ox + x < h.lw[2] && oy + y < h.lh[2]) {
            const float4 v = hl_pull4(&sm.l1[0][0], HF_TX / 2, h.lw[1], h.lh[1], bx0 / 2,
                                      by0 / 2, x, y);
            sm.l2[y][x] = v;
            h.level[2][(size_t)(oy + y) * h.lw[2] + ox + x] = v;
        }
    }
    __syncthreads();
    if (h.nl > 3 && tid < (HF_TX / 8) * (HF_TY / 8)) {
        const int x = tid % (HF_TX / 8), y = tid / (HF_TX / 8), ox = bx0 / 8, oy = by0 / 8;

        if (ox + x < h.lw[3] && oy + y < h.lh[3]) {
            const float4 v = hl_pull4(&sm.l2[0][0], HF_TX / 4, h.lw[2], h.lh[2], bx0 / 4,
                                      by0 / 4, x, y);
            sm.l3[y][x] = v;
            h.level[3][(size_t)(oy + y) * h.lw[3] + ox + x] = v;
        }
    }
    __syncthreads();
    if (h.nl > 4 && tid < (HF_TX / 16) * (HF_TY / 16)) {
        const int x = tid % (HF_TX / 16), y = tid / (HF_TX / 16), ox = bx0 / 16, oy = by0 / 16;

        if (ox + x < h.lw[4] && oy + y < h.lh[4])
            h.level[4][(size_t)(oy + y) * h.lw[4] + ox + x] =
                hl_pull4(&sm.l3[0][0], HF_TX / 8, h.lw[3], h.lh[3], bx0 / 8, by0 / 8, x, y);
    }
}

/* ------------------------------------------------------ pull / push */

/* A texel's colour, clamped to the level; the top level still holds sums
 * and is divided on the way (lib/ref/ref_highlight.c:59-71). */
KDEV void hl_texel(const float4 *lv, int w, int hh, int x, int y, bool top, float c[2])
{
    float4 t;

    x = x < 0 ? 0 : (x >= w ? w - 1 : x);
    y = y < 0 ? 0 : (y >= hh ? hh - 1 : y);
    t = lv[(size_t)y * w + x];
    c[0] = top ? (t.z > 0.0f ? t.x / t.z : 0.0f) : t.x;
    c[1] = top ? (t.z > 0.0f ? t.y / t.z : 0.0f) : t.y;
}

/* lib/ref/ref_highlight.c:73-89: x first, then y, each a (1 - f) + b f */
KDEV void hl_bilinear(const float4 *lv, int w, int hh, float px, float py, bool top, float out[2])
{
    const float fx0 = floorf(px), fy0 = floorf(py), fx = px - fx0, fy = py - fy0;
    const int x0 = (int)fx0, y0 = (int)fy0;
    float c00[2], c10[2], c01[2], c11[2];

    hl_texel(lv, w, hh, x0, y0, top, c00);
    hl_texel(lv, w, hh, x0 + 1, y0, top, c10);
    hl_texel(lv, w, hh, x0, y0 + 1, top, c01);
    hl_texel(lv, w, hh, x0 + 1, y0 + 1, top, c11);
    for (int i = 0; i < 2; i++) {
        const float a = c00[i] * (1.0f - fx) + c10[i] * fx;
        const float b = c01[i] * (1.0f - fx) + c11[i] * fx;

        out[i] = a * (1.0f - fy) + b * fy;
    }
}

KDEV void hl_pull_texel(const KHighlight &h, int l, int x, int y)
{
    const float4 *f = h.level[l - 1];
    const int fw = h.lw[l - 1], fh = h.lh[l - 1];
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);

    for (int dy = 0; dy < 2; dy++)
        for (int dx = 0; dx < 2; dx++) {
            const int fx = 2 * x + dx, fy = 2 * y + dy;

            if (fx < fw && fy < fh) {
                const float4 v = f[(size_t)fy * fw + fx];
                acc.x += v.x;
                acc.y += v.y;
                acc.z += v.z;
            }
        }
    h.level[l][(size_t)y * h.lw[l] + x] = acc;
}

/* lib/ref/ref_highlight.c:254-281: level l rewritten in place from its own
 * sums and the pushed level above */
KDEV void hl_push_texel(const KHighlight &h, int l, int x, int y)
{
    const bool top = l + 2 == h.nl;
    float4 *t = &h.level[l][(size_t)y * h.lw[l] + x];
    float4 v = *t;
    float up[2];

    hl_bilinear(h.level[l + 1], h.lw[l + 1], h.lh[l + 1], ((float)x + 0.5f) * 0.5f - 0.5f,
                ((float)y + 0.5f) * 0.5f - 0.5f, top, up);
    if (v.z > 0.0f) {
        const float s = HMIN(v.z / HL_W_SAT, 1.0f);

        v.x = up[0] + s * (v.x / v.z - up[0]);
        v.y = up[1] + s * (v.y / v.z - up[1]);
    } else {
        v.x = up[0];
        v.y = up[1];
    }
    *t = v;
}

#define HC_NT 1024

/* Levels 5 and up: pulled, then pushed back down to level 4, one
 * workgroup (they hold at most a few thousand texels at any frame size). */
__global__ __launch_bounds__(HC_NT) void k_hl_coarse(KHighlight h)
{
    const int tid = threadIdx.x;

    for (int l = 5; l < h.nl; l++) {
        for (int i = tid; i < h.lw[l] * h.lh[l]; i += HC_NT)
            hl_pull_texel(h, l, i % h.lw[l], i / h.lw[l]);
        __syncthreads();
    }
    for (int l = h.nl - 2; l >= 4; l--) {
        for (int i = tid; i < h.lw[l] * h.lh[l]; i += HC_NT)
            hl_push_texel(h, l, i % h.lw[l], i / h.lw[l]);
        __syncthreads();
    }
}

#define HP_NT 256

__global__ __launch_bounds__(HP_NT) void k_hl_push(KHighlight h, int l)
{
    const int i = blockIdx.x * HP_NT + threadIdx.x;

    if (i < h.lw[l] * h.lh[l])
        hl_push_texel(h, l, i % h.lw[l], i / h.lw[l]);
}

/* ------------------------------------------------------------- fill */

#define HL_FW 64
#define HL_FH 16
#define HL_FNT 256

/* lib/ref/ref_highlight.c:25-32 */
KDEV int hl_mirror(int v, int n)
{
    return v < 0 ? -v : (v >= n ? 2 * n - 2 - v : v);
}

__global__ __launch_bounds__(HL_FNT) void k_hl_fill(KHighlight h)
{
    __shared__ float su[HL_FH + 2][HL_FW + 2];
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * HL_FW, y0 = blockIdx.y * HL_FH;

    for (int i = tid; i < (HL_FW + 2) * (HL_FH + 2); i += HL_FNT) {
        const int lx = i % (HL_FW + 2), ly = i / (HL_FW + 2);
        const int x = hl_mirror(min(x0 - 1 + lx, h.W), h.W);
        const int y = hl_mirror(min(y0 - 1 + ly, h.H), h.H);

        su[ly][lx] = hl_u(h, x, y);
    }
    __syncthreads();

    for (int p = tid; p < HL_FW * HL_FH; p += HL_FNT) {
        const int px = p % HL_FW, py = p / HL_FW, x = x0 + px, y = y0 + py;
        const int c = hl_chan(x, y);
        float u, xc, a, res;

        if (x >= h.W || y >= h.H)
            continue;
        u = su[py + 1][px + 1];
        xc = u * hl_wb(h, c);
        a = hl_ramp(u);
        res = xc;
        if (a > 0.0f) {
            float sum[3] = { 0.0f, 0.0f, 0.0f }, n[3] = { 0.0f, 0.0f, 0.0f };
            float am[3] = { 0.0f, 0.0f, 0.0f }, lf[2], lrho[3];
            float a1, a2, l1, l2, lc, A, lfull, lpart, L;
            int k1, k2;

            hl_bilinear(h.level[0], h.lw[0], h.lh[0], (float)(x >> 1) * 0.5f - 0.25f,
                        (float)(y >> 1) * 0.5f - 0.25f, false, lf);
            lrho[0] = lf[0];
            lrho[1] = 0.0f;
            lrho[2] = lf[1];
            for (int dy = -1; dy <= 1; dy++) {
                for (int dx = -1; dx <= 1; dx++) {
                    const int nx = hl_mirror(x + dx, h.W), ny = hl_mirror(y + dy, h.H);
                    const int k = hl_chan(nx, ny);
                    const float nu = su[py + 1 + dy][px + 1 + dx];

                    if (k == c)
                        continue;
                    sum[k] += nu * hl_wb(h, k);
                    n[k] += 1.0f;
                    am[k] = HMAX(am[k], hl_ramp(nu));
                }
            }
            k1 = c == 0 ? 1 : 0;
            k2 = c == 2 ? 1 : 2;
            a1 = 1.0f - am[k1];
            a2 = 1.0f - am[k2];
            l1 = hl_log(h, HMAX(sum[k1] / n[k1], HL_EPS)) - lrho[k1];
            l2 = hl_log(h, HMAX(sum[k2] / n[k2], HL_EPS)) - lrho[k2];
            lc = hl_log(h, HMAX(xc, HL_EPS)) - lrho[c];
            A = a1 + a2;
            lfull = hl_exp(h, HMAX(lc, HMAX(l1, l2)));
            lpart = A > 0.0f ? hl_exp(h, (a1 * l1 + a2 * l2) / A) : lfull;
            L = lfull + HMIN(A, 1.0f) * (lpart - lfull);
            res = xc + a * HMAX(L * hl_exp(h, lrho[c]) - xc, 0.0f);
        }
        h.mosaic[(size_t)y * h.W + x] = __half_as_ushort(__float2half_rn(res));
    }
}

/* ------------------------------------------------------------ launch */

hipError_t craw_launch_hl_field(const KHighlight &h, hipStream_t s)
{
    const dim3 grid((h.lw[0] + HF_TX - 1) / HF_TX, (h.lh[0] + HF_TY - 1) / HF_TY);

    KLAUNCH(k_hl_field, grid, dim3(HF_NT), 0, s, h);
    return hipGetLastError();
}

hipError_t craw_launch_hl_push(const KHighlight &h, hipStream_t s)
{
    hipError_t e;

    if (h.nl > 5) {
        KLAUNCH(k_hl_coarse, dim3(1), dim3(HC_NT), 0, s, h);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    for (int l = (h.nl - 2 < 3 ? h.nl - 2 : 3); l >= 0; l--) {
        KLAUNCH(k_hl_push, dim3((h.lw[l] * h.lh[l] + HP_NT - 1) / HP_NT), dim3(HP_NT), 0, s, h,
                l);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    return hipSuccess;
}

/* Launches craw_launch_hl_push makes for `h`: the coarse levels' one
 * workgroup when there are more than five levels, then a push a level from
 * min(nl - 2, 3) down to 0. */
int craw_hl_push_launches(const KHighlight &h)
{
    return (h.nl > 5 ? 1 : 0) + ((h.nl - 2 < 3 ? h.nl - 2 : 3) + 1);
}

const void *craw_hl_kernel(int which)
{
    switch (which) {
    case 0: return (const void *)k_hl_field;
    case 1: return (const void *)k_hl_coarse;
    case 2: return (const void *)k_hl_push;
    case 3: return (const void *)k_hl_fill;
    default: return NULL;
    }
}

hipError_t craw_launch_hl_fill(const KHighlight &h, hipStream_t s)
{
    const dim3 grid((h.W + HL_FW - 1) / HL_FW, (h.H + HL_FH - 1) / HL_FH);

    KLAUNCH(k_hl_fill, grid, dim3(HL_FNT), 0, s, h);
    return hipGetLastError();
}
