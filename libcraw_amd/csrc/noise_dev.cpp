/*
 * libcraw for MI355X: craw_noise_estimate_image(), the noise fit with the
 * frame measured on the device (k_noise.hip) and fitted on the host
 * (noise_host.c).
 */
#include <stdlib.h>
#include <string.h>

#include "internal.h"
#include "noise_fit.h"

/* The exact sum held in 32-bit limbs of 64-bit accumulators, in units of
 * 2^-149, correctly rounded to double. */
static double limbs_to_double(const unsigned long long *l)
{
    uint32_t w[KNZ_LIMBS + 2];
    unsigned __int128 acc = 0, top;
    int nw = KNZ_LIMBS + 2, bits, shift;

    for (int k = 0; k < nw; k++) {
        if (k < KNZ_LIMBS)
            acc += l[k];
        w[k] = (uint32_t)acc;
        acc >>= 32;
    }
    while (nw > 0 && !w[nw - 1])
        nw--;
    if (!nw)
        return 0.0;
    bits = 32 * (nw - 1) + (32 - __builtin_clz(w[nw - 1]));
    if (bits <= 128) {
        top = 0;
        for (int k = nw - 1; k >= 0; k--)
            top = (top << 32) | w[k];
        return ldexp((double)top, -149);
    }
    /* the top 128 bits and a sticky bit for everything below them: the
     * conversion then rounds to nearest as the exact value would */
    shift = bits - 128;
    top = 0;
    for (int b = bits - 1; b >= shift; b--)
        top = (top << 1) | ((w[b >> 5] >> (b & 31)) & 1u);
    for (int b = shift - 1; b >= 0; b--)
        if ((w[b >> 5] >> (b & 31)) & 1u) {
            top |= 1;
            break;
        }
    return ldexp((double)top, shift - 149);
}

extern "C" CRAW_API int craw_noise_estimate_image(CRAWContext *ctx, CRAWStream stream,
                                                  const CRAWImage *src,
                                                  const CRAWProfile *profile,
                                                  const CRAWNoiseEstimateParams *params,
                                                  CRAWNoiseEstimate *out)
{
    const hipStream_t s = (hipStream_t)stream;
    uint32_t nb_bins, min_samples;
    float max_level;
    KNoise n;
    size_t nkeys, off_count, off_limbs, off_hist, off_sel, total;
    char *scratch = NULL;
    uint32_t *h_count = NULL, *h_sel = NULL;
    unsigned long long *h_limbs = NULL;
    CRAWNoiseBin *bins = NULL;
    int ret;

    if (!ctx || !src || !profile || !out || !src->data)
        return CRAW_ERROR(EINVAL);
    if (src->format != CRAW_FORMAT_R16_UNORM || src->pitch % 2)
        return CRAW_ERROR(EINVAL);
    if ((ret = craw_noise_check(src->width, src->height, src->pitch / 2, profile)) < 0)
        return ret;
    if ((ret = craw_noise_resolve(params, &nb_bins, &min_samples, &max_level)) < 0)
        return ret;
    if (hipSetDevice(ctx->device) != hipSuccess)
        return CRAW_ERROR(ENODEV);

    memset(&n, 0, sizeof(n));
    n.src = (const uint16_t *)src->data;
    n.pitch = src->pitch / 2;
    n.w2 = (int)(src->width / 2);
    n.h2 = (int)(src->height / 2);
    for (int i = 0; i < 4; i++) {
        n.black[i] = profile->black_level[i];
        n.range[i] = profile->white_level[i] - profile->black_level[i];
    }
    n.nb_bins = nb_bins;
    n.max_level = max_level;

    nkeys = (size_t)4 * n.w2 * n.h2;
    off_count = nkeys * 8;
    off_limbs = off_count + (((size_t)nb_bins * 4 + 255) & ~(size_t)255);
    off_hist = off_limbs + (size_t)nb_bins * KNZ_LIMBS * 8;
    off_sel = off_hist + (size_t)nb_bins * KNZ_D * 4;
    total = off_sel + (size_t)nb_bins * 8;
    pthread_mutex_lock(&ctx->noise_lock);
    if (ctx->noise_scratch_size < total) {
        /* the previous measurement's work must be done before its memory
         * goes; the stream it ran on was synchronised before it returned */
        CRAW_IGNORE(hipFree(ctx->noise_scratch));
        ctx->noise_scratch = NULL;
        ctx->noise_scratch_size = 0;
        if (hipMalloc(&ctx->noise_scratch, total) != hipSuccess) {
            pthread_mutex_unlock(&ctx->noise_lock);
            return CRAW_ERROR(ENOMEM);
        }
        ctx->noise_scratch_size = total;
    }
    scratch = (char *)ctx->noise_scratch;
    n.keys = (uint64_t *)scratch;
    n.count = (uint32_t *)(scratch + off_count);
    n.limbs = (unsigned long long *)(scratch + off_limbs);
    n.hist = (uint32_t *)(scratch + off_hist);
    n.sel = (uint32_t *)(scratch + off_sel);

    h_count = (uint32_t *)malloc(nb_bins * 4);
    h_sel = (uint32_t *)malloc(nb_bins * 8);
    h_limbs = (unsigned long long *)malloc((size_t)nb_bins * KNZ_LIMBS * 8);
    bins = (CRAWNoiseBin *)calloc(nb_bins, sizeof(*bins));
    if (!h_count || !h_sel || !h_limbs || !bins) {
        ret = CRAW_ERROR(ENOMEM);
        goto done;
    }

    /* counts, sums, histograms and selection start at zero */
    if (craw_launch_zero_u32(n.count, (total - off_count) / 4, s) != hipSuccess ||
        craw_launch_noise_measure(n, s) != hipSuccess ||
        hipMemcpyAsync(h_count, n.count, nb_bins * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(h_limbs, n.limbs, (size_t)nb_bins * KNZ_LIMBS * 8, hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipMemcpyAsync(h_sel, n.sel, nb_bins * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        craw_log(ctx, CRAW_LOG_ERROR, "Noise measurement failed: %s\n",
                 hipGetErrorString(hipGetLastError()));
        ret = CRAW_ERROR(EIO);
        goto done;
    }

    for (uint32_t b = 0; b < nb_bins; b++) {
        bins[b].n = h_count[b];
        bins[b].mean_sum = limbs_to_double(h_limbs + (size_t)b * KNZ_LIMBS);
        if (h_count[b])
            memcpy(&bins[b].median, &h_sel[2 * b], 4);
    }
    ret = craw_noise_fit(bins, nb_bins, min_samples, profile, out);

done:
    pthread_mutex_unlock(&ctx->noise_lock);
    free(h_count);
    free(h_sel);
    free(h_limbs);
    free(bins);
    return ret;
}
