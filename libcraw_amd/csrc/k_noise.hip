/*
 * libcraw for MI355X: the noise fit's measurement of a frame in device
 * memory (lib/noise.c:394-455 of cyanreg/libcraw measures it on the host).
 *
 * k_noise_measure: one workgroup per 64x60 samples of one CFA plane. The
 * plane is normalised in double and stored as float in LDS with the 9x9
 * mean's reach; each lane walks one column down 15 rows, holding the 9x9
 * window in registers (one new row of nine a sample), and gives every
 * interior sample the reference's high-pass residual and 9x9 mean (summed
 * in raster order) and so its bin. It writes the sample's key,
 * (bin << 32) | bits(|r|) (all ones for a sample the measurement skips),
 * and adds to its bin's count, to the bin's exact sum of means and to the
 * first-level histogram of |r|:
 *   - a mean is 24 mantissa bits times 2^(e - 150); added at bit e - 1 of
 *     a fixed point in units of 2^-149 split in 32-bit limbs, each
 *     accumulated in 64 bits, the sum is exact in any order. A lane sums
 *     its run of samples of one bin (and two neighbouring limbs) in
 *     registers and adds the run to the workgroup's tables in LDS when the
 *     bin changes; the tables go to memory with global atomics at the end;
 *   - the histogram is keyed by bits 30..20 of |r| (sign 0), in a 4096-entry
 *     LDS hash whose entries hold (bin, digit) << 12 | count (a tile's 3840
 *     samples cannot overflow the count; digits of finite |r| stay below
 *     2040, so the all-ones word is free to mark an empty entry), with a
 *     global fallback after 32 probes.
 * The median of each bin (rank n / 2) is then found by radix selection:
 * k_noise_select picks each bin's digit holding the rank (and clears the
 * row for the next level), k_noise_level histograms the next digit (bits
 * 19..9, then 8..0) of the keys whose higher digits were selected.
 */
#include "k_common.h"
#include "kernels.h"

#define NZ_TW 64
#define NZ_WR 15
#define NZ_TH (4 * NZ_WR)
#define NZ_R  4
#define NZ_NT 256
#define NZ_LW (NZ_TW + 2 * NZ_R)
#define NZ_LH (NZ_TH + 2 * NZ_R)
#define NZ_HASH 4096
#define NZ_PROBES 32
#define NZ_EMPTY 0xffffffffu

struct SmNoise {
    float plane[NZ_LH][NZ_LW];
    uint32_t hash[NZ_HASH];
};

/* The per-bin tables after SmNoise: limbs (8-byte aligned) then counts. */
static size_t nz_dyn_bytes(uint32_t nb_bins)
{
    return (size_t)nb_bins * (KNZ_LIMBS * sizeof(unsigned long long) + sizeof(uint32_t));
}

KDEV void nz_hist_add(uint32_t *hash, uint32_t *ghist, uint32_t key)
{
    uint32_t h = (key * 0x9e3779b1u) >> 20;

#pragma unroll 1
    for (int probe = 0; probe < NZ_PROBES; probe++) {
        uint32_t v = __hip_atomic_load(&hash[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);

        if (v == NZ_EMPTY) {
            v = atomicCAS(&hash[h], NZ_EMPTY, (key << 12) | 1u);
            if (v == NZ_EMPTY)
                return;
        }
        if ((v >> 12) == key) {
            atomicAdd(&hash[h], 1u);
            return;
        }
        h = (h + 1) & (NZ_HASH - 1);
    }
    atomicAdd(&ghist[(key >> 11) * KNZ_D + (key & 2047u)], 1u);
}

/* A run's count and limb sums l, l + 1, l + 2 into the workgroup's tables.
 * A sum is non-zero only for a limb some sample of the run reached, and
 * every sample's limbs lie below KNZ_LIMBS. */
KDEV void nz_run_flush(uint32_t *count, unsigned long long *limb, int b, uint32_t cnt, int l,
                       unsigned long long a0, unsigned long long a1, unsigned long long a2)
{
    unsigned long long *lb = limb + (size_t)b * KNZ_LIMBS + l;

    atomicAdd(&count[b], cnt);
    if (a0)
        atomicAdd(lb, a0);
    if (a1)
        atomicAdd(lb + 1, a1);
    if (a2)
        atomicAdd(lb + 2, a2);
}

__global__ __launch_bounds__(NZ_NT) void k_noise_measure(KNoise n)
{
    __shared__ SmNoise sm;
    extern __shared__ unsigned long long nz_dyn[];
    unsigned long long *limb = nz_dyn;
    uint32_t *count = (uint32_t *)(nz_dyn + (size_t)n.nb_bins * KNZ_LIMBS);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int pl = blockIdx.z, px = pl & 1, py = pl >> 1;
    const int i0 = blockIdx.x * NZ_TW, k0 = blockIdx.y * NZ_TH;
    const double black = n.black[pl], range = n.range[pl];
    uint64_t *keys = n.keys + (size_t)pl * n.w2 * n.h2;

    for (int b = tid; b < (int)(n.nb_bins * KNZ_LIMBS); b += NZ_NT)
        limb[b] = 0ull;
    for (int b = tid; b < (int)n.nb_bins; b += NZ_NT)
        count[b] = 0u;
    for (int i = tid; i < NZ_HASH; i += NZ_NT)
        sm.hash[i] = NZ_EMPTY;
    /* the plane over the tile and its reach; positions past the plane
     * (which no measured sample reaches) repeat its edge */
    for (int i = tid; i < NZ_LW * NZ_LH; i += NZ_NT) {
        const int lx = i % NZ_LW, ly = i / NZ_LW;
        const int x = min(max(i0 - NZ_R + lx, 0), n.w2 - 1);
        const int y = min(max(k0 - NZ_R + ly, 0), n.h2 - 1);
        const double raw = n.src[(size_t)(2 * y + py) * n.pitch + 2 * x + px] * (1.0 / 65535.0);

        sm.plane[ly][lx] = (float)((raw - black) / range);
    }
    __syncthreads();

    {
        const int i = i0 + lane, kb = k0 + wave * NZ_WR, cyb = wave * NZ_WR;
        const bool col_in = i < n.w2, col_mid = i >= NZ_R && i < n.w2 - NZ_R;
        int run_bin = -1, run_l = 0;
        uint32_t run_n = 0u;
        unsigned long long a0 = 0ull, a1 = 0ull, a2 = 0ull;
        float win[2 * NZ_R + 1][2 * NZ_R + 1];

        /* win[j][d]: plane row (sample row - 4 + j), column (lane - 4 + d) */
#pragma unroll
        for (int j = 0; j < 2 * NZ_R; j++)
#pragma unroll
            for (int d = 0; d <= 2 * NZ_R; d++)
                win[j][d] = sm.plane[cyb + j][lane + d];

#pragma unroll
        for (int s = 0; s < NZ_WR; s++) {
            const int k = kb + s;
            uint64_t key = ~0ull;

#pragma unroll
            for (int d = 0; d <= 2 * NZ_R; d++)
                win[2 * NZ_R][d] = sm.plane[cyb + s + 2 * NZ_R][lane + d];

            if (col_mid && k >= NZ_R && k < n.h2 - NZ_R) {
                float r, mean = 0.0f;

                /* lib/noise.c:415-426: [1 -2 1; -2 4 -2; 1 -2 1] / 6, and the
                 * 9x9 mean in raster order */
                r = (win[3][3] + win[3][5] + win[5][3] + win[5][5] -
                     2.0f * (win[3][4] + win[5][4] + win[4][3] + win[4][5]) + 4.0f * win[4][4]) *
                    (1.0f / 6.0f);
#pragma unroll
                for (int dy = 0; dy <= 2 * NZ_R; dy++)
#pragma unroll
                    for (int dx = 0; dx <= 2 * NZ_R; dx++)
                        mean += win[dy][dx];
                mean *= 1.0f / ((2 * NZ_R + 1) * (2 * NZ_R + 1));

                if (mean > 0.0f && !(mean > n.max_level)) {
                    int bin = (int)(sqrtf(mean) * (float)n.nb_bins);
                    const uint32_t rb = __float_as_uint(fabsf(r));
                    const uint32_t mb = __float_as_uint(mean);
                    const uint32_t e = mb >> 23;
                    const uint64_t m24 = (mb & 0x7fffffu) | (e ? 0x800000u : 0u);
                    const int bit = (e ? (int)e : 1) - 1, l = bit >> 5;
                    const uint64_t v = m24 << (bit & 31);
                    const unsigned long long lo = v & 0xffffffffull, hi = v >> 32;
                    bool up;

                    bin = bin < 0 ? 0 : (bin >= (int)n.nb_bins ? (int)n.nb_bins - 1 : bin);
                    key = ((uint64_t)bin << 32) | rb;
                    if (bin != run_bin || (l != run_l && l != run_l + 1)) {
                        if (run_n)
                            nz_run_flush(count, limb, run_bin, run_n, run_l, a0, a1, a2);
                        run_bin = bin;
                        run_l = l;
                        run_n = 0u;
                        a0 = a1 = a2 = 0ull;
                    }
                    up = l != run_l;
                    run_n++;
                    a0 += up ? 0ull : lo;
                    a1 += up ? lo : hi;
                    a2 += up ? hi : 0ull;
                    nz_hist_add(sm.hash, n.hist, ((uint32_t)bin << 11) | (rb >> 20));
                }
            }
            if (col_in && k < n.h2)
                keys[(size_t)k * n.w2 + i] = key;

#pragma unroll
            for (int j = 0; j < 2 * NZ_R; j++)
#pragma unroll
                for (int d = 0; d <= 2 * NZ_R; d++)
                    win[j][d] = win[j + 1][d];
        }
        if (run_n)
            nz_run_flush(count, limb, run_bin, run_n, run_l, a0, a1, a2);
    }
    __syncthreads();

    for (int b = tid; b < (int)n.nb_bins; b += NZ_NT) {
        if (!count[b])
            continue;
        atomicAdd(&n.count[b], count[b]);
        for (int l = 0; l < KNZ_LIMBS; l++)
            if (limb[(size_t)b * KNZ_LIMBS + l])
                atomicAdd(&n.limbs[(size_t)b * KNZ_LIMBS + l], limb[(size_t)b * KNZ_LIMBS + l]);
    }
    for (int i = tid; i < NZ_HASH; i += NZ_NT) {
        const uint32_t v = sm.hash[i];

        if (v != NZ_EMPTY)
            atomicAdd(&n.hist[(v >> 23) * KNZ_D + ((v >> 12) & 2047u)], v & 4095u);
    }
}

/* Level `level` (1..3) of bin blockIdx.x's selection: the digit of its
 * histogram row holding the rank sought (n / 2 at the first level), the
 * selected prefix and the rank left within the digit; the row is cleared
 * for the next level. */
__global__ __launch_bounds__(NZ_NT) void k_noise_select(KNoise n, int level)
{
    __shared__ uint32_t part[NZ_NT];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int nd = level == 3 ? 512 : 2048, per = nd / NZ_NT;
    uint32_t *row = n.hist + (size_t)b * KNZ_D;
    const uint32_t cnt = n.count[b];
    const uint32_t rank = level == 1 ? cnt / 2 : n.sel[2 * b + 1];
    uint32_t mine[8], sum = 0u;

    for (int j = 0; j < per; j++) {
        mine[j] = row[tid * per + j];
        sum += mine[j];
    }
    part[tid] = sum;
    __syncthreads();
    /* inclusive scan of the threads' sums (Hillis-Steele) */
    for (int off = 1; off < NZ_NT; off <<= 1) {
        const uint32_t add = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += add;
        __syncthreads();
    }
    if (cnt) {
        const uint32_t before = tid ? part[tid - 1] : 0u;

        if (rank >= before && rank < part[tid]) {
            uint32_t acc = before;

            for (int j = 0; j < per; j++) {
                if (rank < acc + mine[j]) {
                    const uint32_t digit = (uint32_t)(tid * per + j);
                    const uint32_t prefix = level == 1 ? 0u : n.sel[2 * b];

                    n.sel[2 * b] = level == 3 ? (prefix << 9) | digit : (prefix << 11) | digit;
                    n.sel[2 * b + 1] = rank - acc;
                    break;
                }
                acc += mine[j];
            }
        }
    }
    for (int j = 0; j < per; j++)
        row[tid * per + j] = 0u;
}

/* The next digit of the keys whose higher digits were selected: bits 19..9
 * (level 2) or 8..0 (level 3) of |r|. */
__global__ __launch_bounds__(NZ_NT) void k_noise_level(KNoise n, int level, size_t nkeys)
{
    for (size_t i = (size_t)blockIdx.x * NZ_NT + threadIdx.x; i < nkeys;
         i += (size_t)gridDim.x * NZ_NT) {
        const uint64_t key = n.keys[i];
        uint32_t b, rb;

        if (key == ~0ull)
            continue;
        b = (uint32_t)(key >> 32);
        rb = (uint32_t)key;
        if (level == 2 ? (rb >> 20) == n.sel[2 * b] : (rb >> 9) == n.sel[2 * b])
            atomicAdd(&n.hist[(size_t)b * KNZ_D + (level == 2 ? (rb >> 9) & 2047u : rb & 511u)],
                      1u);
    }
}

hipError_t craw_launch_noise_measure(const KNoise &n, hipStream_t s)
{
    const size_t nkeys = (size_t)4 * n.w2 * n.h2;
    const int grid_keys = (int)((nkeys + NZ_NT - 1) / NZ_NT < 65536 ? (nkeys + NZ_NT - 1) / NZ_NT
                                                                     : 65536);
    hipError_t e;

    KLAUNCH(k_noise_measure, dim3((n.w2 + NZ_TW - 1) / NZ_TW, (n.h2 + NZ_TH - 1) / NZ_TH, 4),
            dim3(NZ_NT), nz_dyn_bytes(n.nb_bins), s, n);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    for (int level = 1; level <= 3; level++) {
        if (level > 1) {
            KLAUNCH(k_noise_level, dim3(grid_keys), dim3(NZ_NT), 0, s, n, level, nkeys);
            if ((e = hipGetLastError()) != hipSuccess)
                return e;
        }
        KLAUNCH(k_noise_select, dim3(n.nb_bins), dim3(NZ_NT), 0, s, n, level);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    return hipSuccess;
}
