/*
 * craw_bench: develop frames through libcraw's C API and report each pass's
 * GPU time, as cineraw's --bench does (tools/cineraw.c:1156-1290 of
 * cyanreg/libcraw): per pass the dispatches and the average and median
 * time over the frames, their sum, and the whole recording's time.
 *
 *   craw_bench [--raw FILE] [--width W] [--height H] [--frames N]
 *              [--warmup K] [--debayer rcd|mhc|bilinear]
 *              [--space xyz|709|2020|p3|ap0|ap1] [--denoise S]
 *              [--chroma S] [--temporal A] [--false-color N] [--highlights]
 *              [--awb] [--fp16] [--fast] [--half] [--cold] [--estimate-noise]
 *              [--out FILE]
 *
 * The frame is a raw little-endian uint16 RGGB dump of W x H samples
 * (12-bit data, black 512, white 4095 of 65535, the format of the
 * reference's test frames), or without --raw a synthetic one: smooth
 * gradients with Gaussian noise of variance 2e-4 u + 4e-6 (with
 * --highlights, 1.7 times brighter, so that its brighter part clips and
 * the highlight stage has work). It is uploaded
 * once; the frames are recorded back to back on one stream (warm), or
 * with --cold each waited for before the next. --estimate-noise fits the
 * noise model on the device first and develops with it.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include <craw/pipeline.h>

#define BLACK (512.0f / 65535.0f)
#define WHITE (4095.0f / 65535.0f)
#define MAX_PASSES 16

struct series {
    double *v;
    size_t n;
};

struct pass {
    const char *name;
    unsigned count;
    struct series ms;
};

static void series_push(struct series *s, double v)
{
    if (!(s->n & (s->n - 1))) {
        double *nv = realloc(s->v, (s->n ? 2 * s->n : 1) * sizeof(*nv));

        if (!nv) {
            fprintf(stderr, "out of memory\n");
            exit(1);
        }
        s->v = nv;
    }
    s->v[s->n++] = v;
}

static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static void series_stats(struct series *s, double *avg, double *median)
{
    double sum = 0.0;

    *avg = *median = 0.0;
    if (!s->n)
        return;
    for (size_t i = 0; i < s->n; i++)
        sum += s->v[i];
    qsort(s->v, s->n, sizeof(*s->v), cmp_double);
    *avg = sum / (double)s->n;
    *median = s->n & 1 ? s->v[s->n / 2] : 0.5 * (s->v[s->n / 2 - 1] + s->v[s->n / 2]);
}

/* xorshift64* and Box-Muller: the synthetic frame's noise */
static uint64_t rng_state = 0x9e3779b97f4a7c15ull;

static double uniform01(void)
{
    rng_state ^= rng_state >> 12;
    rng_state ^= rng_state << 25;
    rng_state ^= rng_state >> 27;
    return ((rng_state * 0x2545f4914f6cdd1dull) >> 11) * (1.0 / 9007199254740992.0);
}

static double gaussian(void)
{
    const double u = uniform01() + 1e-300, v = uniform01();
    return sqrt(-2.0 * log(u)) * cos(2.0 * M_PI * v);
}

static void synthetic_frame(uint16_t *m, int w, int h, double gain)
{
    static const double wb[4] = { 2.0, 1.0, 1.0, 1.6 };

    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const int s = (y & 1) * 2 + (x & 1);
            const double fx = (double)x / w, fy = (double)y / h;
            double u = gain * (0.05 + 0.6 * fx * (0.4 + 0.6 * fy) +
                               0.05 * sin(x * 0.21) * sin(y * 0.13)) / wb[s];
            double raw;

            u += gaussian() * sqrt(2e-4 * (u > 0 ? u : 0) + 4e-6);
            raw = rint((BLACK + u * (WHITE - BLACK)) * 65535.0);
            m[(size_t)y * w + x] = (uint16_t)(raw < 0 ? 0 : (raw > 4095 ? 4095 : raw));
        }
}

static void usage(void)
{
    fprintf(stderr, "usage: craw_bench [--raw FILE] [--width W] [--height H] [--frames N] "
                    "[--warmup K] [--debayer rcd|mhc|bilinear] [--space xyz|709|2020|p3|ap0|ap1] "
                    "[--denoise S] [--chroma S] [--temporal A] [--false-color N] [--highlights] [--awb] "
                    "[--fp16] [--fast] [--half] "
                    "[--cold] [--estimate-noise] [--out FILE]\n");
}

#define CHECK_HIP(x)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                        \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

int main(int argc, char **argv)
{
    const char *raw_path = NULL, *out_path = NULL;
    int w = 5760, h = 3240, frames = 100, warmup = 50, cold = 0, awb = 0, fp16 = 0, half = 0;
    int fast = 0;
    int estimate_noise = 0, false_color = 0, highlights = 0;
    float denoise = 1.0f, chroma = 0.5f, temporal = 0.0f;
    enum CRAWDebayer debayer = CRAW_DEBAYER_RCD;
    enum CRAWColorSpace space = CRAW_CS_ACES_AP1_LINEAR;
    CRAWContext *ctx = NULL;
    CRAWPipeline *pipe = NULL, *plain = NULL;
    static const char *const debayer_names[] = { "bilinear", "mhc", "rcd" };
    CRAWPipelineParams pp;
    CRAWFrameParams fp;
    CRAWProfile prof;
    CRAWHipInitParams hp;
    CRAWImage src, dst;
    struct pass passes[MAX_PASSES];
    struct series sum = { 0 }, whole = { 0 };
    int nb_passes = 0, ret;
    uint16_t *mosaic;
    void *d_src = NULL, *d_dst = NULL;
    size_t out_bpp;
    hipStream_t stream;
    hipEvent_t ev[CRAW_PIPELINE_PROFILE_SLOTS][2];

    for (int i = 1; i < argc; i++) {
        const char *a = argv[i], *v = i + 1 < argc ? argv[i + 1] : NULL;

        if (!strcmp(a, "--raw") && v) raw_path = argv[++i];
        else if (!strcmp(a, "--out") && v) out_path = argv[++i];
        else if (!strcmp(a, "--width") && v) w = atoi(argv[++i]);
        else if (!strcmp(a, "--height") && v) h = atoi(argv[++i]);
        else if (!strcmp(a, "--frames") && v) frames = atoi(argv[++i]);
        else if (!strcmp(a, "--warmup") && v) warmup = atoi(argv[++i]);
        else if (!strcmp(a, "--denoise") && v) denoise = (float)atof(argv[++i]);
        else if (!strcmp(a, "--chroma") && v) chroma = (float)atof(argv[++i]);
        else if (!strcmp(a, "--temporal") && v) temporal = (float)atof(argv[++i]);
        else if (!strcmp(a, "--false-color") && v) false_color = atoi(argv[++i]);
        else if (!strcmp(a, "--awb")) awb = 1;
        else if (!strcmp(a, "--highlights")) highlights = 1;
        else if (!strcmp(a, "--fp16")) fp16 = 1;
        else if (!strcmp(a, "--fast")) fast = 1;
        else if (!strcmp(a, "--half")) half = 1;
        else if (!strcmp(a, "--cold")) cold = 1;
        else if (!strcmp(a, "--estimate-noise")) estimate_noise = 1;
        else if (!strcmp(a, "--debayer") && v) {
            i++;
            debayer = !strcmp(v, "mhc") ? CRAW_DEBAYER_MHC :
                      (!strcmp(v, "bilinear") ? CRAW_DEBAYER_BILINEAR : CRAW_DEBAYER_RCD);
        } else if (!strcmp(a, "--space") && v) {
            i++;
            space = !strcmp(v, "xyz") ? CRAW_CS_XYZ_D65 :
                    !strcmp(v, "709") ? CRAW_CS_BT709_LINEAR :
                    !strcmp(v, "2020") ? CRAW_CS_BT2020_LINEAR :
                    !strcmp(v, "p3") ? CRAW_CS_P3D65_LINEAR :
                    !strcmp(v, "ap0") ? CRAW_CS_ACES_AP0_LINEAR : CRAW_CS_ACES_AP1_LINEAR;
        } else {
            usage();
            return 2;
        }
    }
    if (w <= 0 || h <= 0 || (w & 1) || (h & 1) || frames <= 0 || warmup < 0) {
        usage();
        return 2;
    }

    mosaic = malloc((size_t)w * h * sizeof(*mosaic));
    if (!mosaic)
        return 1;
    if (raw_path) {
        FILE *f = fopen(raw_path, "rb");

        if (!f || fread(mosaic, sizeof(*mosaic), (size_t)w * h, f) != (size_t)w * h) {
            fprintf(stderr, "%s: not a %dx%d uint16 frame\n", raw_path, w, h);
            return 1;
        }
        fclose(f);
    } else {
        synthetic_frame(mosaic, w, h, highlights ? 1.7 : 1.0);
    }

    memset(&hp, 0, sizeof(hp));
    hp.device = 0;
    if ((ret = craw_context_create(&ctx, &hp, NULL)) < 0) {
        fprintf(stderr, "craw_context_create: %d\n", ret);
        return 1;
    }

    craw_profile_init(&prof);
    for (int s = 0; s < 4; s++) {
        prof.black_level[s] = BLACK;
        prof.white_level[s] = WHITE;
    }
    prof.as_shot_wb[0] = 2.0f;
    prof.as_shot_wb[1] = 1.0f;
    prof.as_shot_wb[2] = 1.6f;
    {
        /* balanced camera RGB to XYZ D65: rows sum to D65 white's XYZ */
        static const float cam[3][3] = { { 0.6300f, 0.2400f, 0.0805f },
                                         { 0.2600f, 0.8300f, -0.0900f },
                                         { 0.0300f, -0.1000f, 1.1590f } };
        memcpy(prof.as_shot_matrix, cam, sizeof(cam));
    }

    out_bpp = half ? 8 : 16;
    CHECK_HIP(hipStreamCreate(&stream));
    for (int k = 0; k < CRAW_PIPELINE_PROFILE_SLOTS; k++) {
        CHECK_HIP(hipEventCreate(&ev[k][0]));
        CHECK_HIP(hipEventCreate(&ev[k][1]));
    }
    CHECK_HIP(hipMalloc(&d_src, (size_t)w * h * 2));
    CHECK_HIP(hipMalloc(&d_dst, (size_t)w * h * out_bpp));
    CHECK_HIP(hipMemcpy(d_src, mosaic, (size_t)w * h * 2, hipMemcpyHostToDevice));
    src = (CRAWImage) { d_src, (size_t)w * 2, CRAW_FORMAT_R16_UNORM, (uint32_t)w, (uint32_t)h };
    dst = (CRAWImage) { d_dst, (size_t)w * out_bpp,
                        half ? CRAW_FORMAT_R16G16B16A16_SFLOAT : CRAW_FORMAT_R32G32B32A32_SFLOAT,
                        (uint32_t)w, (uint32_t)h };

    craw_frame_params_init(&fp);
    fp.profile = &prof;
    fp.denoise_strength = denoise;
    fp.denoise_chroma_strength = chroma;
    fp.denoise_temporal_strength = temporal;
    fp.false_color_passes = (uint32_t)false_color;
    fp.flags = highlights ? CRAW_FRAME_FLAG_HIGHLIGHT_RECONSTRUCTION : 0;
    fp.noise_override.a = 2e-4f;
    fp.noise_override.b = 4e-6f;
    if (estimate_noise && (denoise > 0.0f || temporal > 0.0f)) {
        CRAWNoiseEstimate est;
        hipEvent_t n0, n1;
        float nms = 0.0f;

        CHECK_HIP(hipEventCreate(&n0));
        CHECK_HIP(hipEventCreate(&n1));
        CHECK_HIP(hipEventRecord(n0, stream));
        ret = craw_noise_estimate_image(ctx, stream, &src, &prof, NULL, &est);
        CHECK_HIP(hipEventRecord(n1, stream));
        CHECK_HIP(hipEventSynchronize(n1));
        CHECK_HIP(hipEventElapsedTime(&nms, n0, n1));
        if (ret < 0) {
            fprintf(stderr, "craw_noise_estimate_image: %d\n", ret);
        } else {
            printf("noise model: a %.4e b %.4e, %u bins, fit error %.1f%%, read noise %.2f "
                   "code values (%.2f ms)\n", (double)est.model.a, (double)est.model.b,
                   est.nb_bins, 100.0 * (double)est.fit_rms, (double)est.read_noise_lsb,
                   (double)nms);
            fp.noise_override = est.model;
        }
    }

    craw_pipeline_params_init(&pp);
    pp.width = (uint32_t)w;
    pp.height = (uint32_t)h;
    pp.debayer = debayer;
    pp.out_space = space;
    pp.out_format = dst.format;
    pp.flags = CRAW_PIPELINE_FLAG_PROFILE | (awb ? CRAW_PIPELINE_FLAG_AWB_HISTOGRAM : 0) |
               (denoise > 0.0f ? CRAW_PIPELINE_FLAG_DENOISE : 0) |
               (chroma > 0.0f ? CRAW_PIPELINE_FLAG_CHROMA_DENOISE : 0) |
               (temporal > 0.0f ? CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE : 0) |
               (false_color > 0 ? CRAW_PIPELINE_FLAG_FALSE_COLOR : 0) |
               (highlights ? CRAW_PIPELINE_FLAG_HIGHLIGHT_RECONSTRUCTION : 0) |
               (fp16 ? CRAW_PIPELINE_FLAG_NLM_FP16 : 0) | (fast ? CRAW_PIPELINE_FLAG_NLM_FAST : 0);
    if ((ret = craw_pipeline_create(ctx, &pp, &pipe)) < 0) {
        fprintf(stderr, "craw_pipeline_create: %d\n", ret);
        return 1;
    }

    /* Profiled frames, recorded in groups of the profiler's slots (back to
     * back) or one at a time (--cold), each group waited for and read. */
    memset(passes, 0, sizeof(passes));
    for (int i = 0; i < warmup + frames;) {
        const int group = cold ? 1 : CRAW_PIPELINE_PROFILE_SLOTS;
        int n = warmup + frames - i < group ? warmup + frames - i : group;

        for (int k = 0; k < n; k++) {
            CHECK_HIP(hipEventRecord(ev[k][0], stream));
            if ((ret = craw_pipeline_record(pipe, stream, &src, &dst, &fp)) < 0) {
                fprintf(stderr, "craw_pipeline_record: %d\n", ret);
                return 1;
            }
            CHECK_HIP(hipEventRecord(ev[k][1], stream));
        }
        CHECK_HIP(hipEventSynchronize(ev[n - 1][1]));
        for (int k = 0; k < n; k++, i++) {
            CRAWStageTime st[MAX_PASSES];
            float ms = 0.0f;
            double s = 0.0;
            int nb;

            if (i < warmup)
                continue;
            CHECK_HIP(hipEventElapsedTime(&ms, ev[k][0], ev[k][1]));
            nb = craw_pipeline_get_stage_times_n(pipe, (uint32_t)(n - 1 - k), st, MAX_PASSES);
            if (nb < 0) {
                fprintf(stderr, "craw_pipeline_get_stage_times_n: %d\n", nb);
                return 1;
            }
            for (int q = 0; q < nb && q < MAX_PASSES; q++) {
                int j;

                for (j = 0; j < nb_passes; j++)
                    if (!strcmp(passes[j].name, st[q].name))
                        break;
                if (j == nb_passes) {
                    if (nb_passes == MAX_PASSES)
                        continue;
                    passes[nb_passes++].name = st[q].name;
                }
                passes[j].count = st[q].count;
                series_push(&passes[j].ms, st[q].ms);
                s += st[q].ms;
            }
            series_push(&sum, s);
            series_push(&whole, ms);
        }
    }

    /* Throughput: the same frames back to back on a pipeline without
     * timestamps, after the profiled run has brought the clocks up. */
    pp.flags &= ~(uint32_t)CRAW_PIPELINE_FLAG_PROFILE;
    if ((ret = craw_pipeline_create(ctx, &pp, &plain)) < 0) {
        fprintf(stderr, "craw_pipeline_create: %d\n", ret);
        return 1;
    }
    for (int pass = 0; pass < 2; pass++) {
        const int n = pass ? frames : warmup;
        float ms = 0.0f;

        CHECK_HIP(hipEventRecord(ev[0][0], stream));
        for (int i = 0; i < n; i++)
            if ((ret = craw_pipeline_record(plain, stream, &src, &dst, &fp)) < 0) {
                fprintf(stderr, "craw_pipeline_record: %d\n", ret);
                return 1;
            }
        CHECK_HIP(hipEventRecord(ev[0][1], stream));
        CHECK_HIP(hipEventSynchronize(ev[0][1]));
        CHECK_HIP(hipEventElapsedTime(&ms, ev[0][0], ev[0][1]));
        if (pass && n)
            printf("%dx%d %s, %d frames back to back: %.3f ms a frame, %.1f Mpix/s\n", w, h,
                   debayer_names[debayer], n, ms / n, (double)w * h / 1e6 / (ms / n / 1e3));
    }

    printf("\nGPU time per pass over %zu frame%s (%s):\n", sum.n, sum.n == 1 ? "" : "s",
           cold ? "each waited for" : "back to back, in groups of four");
    printf("  %-20s %10s %10s %10s\n", "pass", "dispatches", "avg ms", "median ms");
    for (int j = 0; j < nb_passes; j++) {
        double avg, med;

        series_stats(&passes[j].ms, &avg, &med);
        printf("  %-20s %10u %10.3f %10.3f\n", passes[j].name, passes[j].count, avg, med);
        free(passes[j].ms.v);
    }
    {
        double avg, med;

        series_stats(&sum, &avg, &med);
        printf("  %-20s %10s %10.3f %10.3f\n", "sum of the passes", "", avg, med);
        series_stats(&whole, &avg, &med);
        printf("  %-20s %10s %10.3f %10.3f\n", "whole recording", "", avg, med);
    }

    if (out_path) {
        void *host = malloc((size_t)w * h * out_bpp);
        FILE *f = fopen(out_path, "wb");

        if (host && f) {
            CHECK_HIP(hipMemcpy(host, d_dst, (size_t)w * h * out_bpp, hipMemcpyDeviceToHost));
            fwrite(host, out_bpp, (size_t)w * h, f);
        }
        if (f)
            fclose(f);
        free(host);
    }

    free(sum.v);
    free(whole.v);
    free(mosaic);
    craw_pipeline_free(&plain);
    craw_pipeline_free(&pipe);
    CHECK_HIP(hipFree(d_src));
    CHECK_HIP(hipFree(d_dst));
    CHECK_HIP(hipStreamDestroy(stream));
    craw_context_free(&ctx);
    return 0;
}
