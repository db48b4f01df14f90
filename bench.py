"""Benchmark: developed Mpix/s of the 5.8K full pipeline on one MI355X.

One step develops one 5760x3240 RGGB 12-bit frame (black 512, white 4095)
through the whole path (config 2 of the workload list): variance-
stabilised NLM on the four CFA planes at strength 1.0, RCD debayer, chroma
denoise 0.5, the AWB histogram, camera matrix to ACES AP1 linear with the
ACES 1.3 gamut compression, RGBA32F out; FP32 throughout, the plain form:
one craw_pipeline_record() per frame on one stream. The input frame is
resident in HBM before timing starts.

Prints one JSON line. `roofline` is the dominant kernel's: algorithmic
bytes per launch over its average launch time, taken from the library's
own hipEvent pair around each launch (CRAW_PIPELINE_FLAG_PROFILE, on the
stream the kernels run on) over the last recordings of the timed region.
`cpu_baseline` times the oracle (the CPU restatement of lib/ref) on the
host: single-threaded on a bounded sample and on all cores on the full
frame, both on the same seeded scene.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 39.3   # fp32 lane operations/s, non-packed: 256 CU x 64 lanes x 2.4 GHz

W, H = 5760, 3240
NOISE = (2e-4, 4e-6)

# Algorithmic HBM bytes per pixel of each kernel at its API formats (u16
# mosaic, float camera RGB planes, quarter-grid planes, RGBA32F out);
# halo re-reads and LDS traffic excluded.
ALGO_BYTES_PER_PX = {
    "denoise_nlm": 2 + 2,                    # mosaic in, denoised mosaic out
    "rcd": 2 + 12 + 12.0 / 16,               # mosaic in, RGB planes + block means out
    "chroma_blur": 5 * 4 / 16.0,             # 3 quarter planes in, 2 out
    "chroma_up_color": 12 + 20.0 / 16 + 16,  # RGB planes + quarter planes in, RGBA32F out
    "awb_clear": 0.0,
}
# VALU operations per sample of the FP32 NLM (7x7 search, 3x3 patch):
# 49 x (9 differences, 9 squares, 8 adds for the distance; 4 for the
# weight's argument; ~22 for the deterministic exp; 3 to accumulate).
NLM_OPS_PER_SAMPLE = 49 * (9 + 9 + 8 + 4 + 22 + 3)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def cpu_baseline(threads):
    """The oracle's float/glibc mode (lib/ref as written) on the host."""
    import oracle
    from libcraw_amd import capi, scene
    from libcraw_amd.pipeline import frame_params, pipeline_params

    prof = scene.make_profile()
    res = {}
    for label, (w, h, n) in (("single", (2880, 1620, 1)), ("all", (W, H, threads))):
        bayer = scene.make_frame(w, h, seed=1, a=NOISE[0], b=NOISE[1])
        pp = pipeline_params(w, h, debayer=capi.CRAW_DEBAYER_RCD,
                             out_space=capi.CRAW_CS_ACES_AP1_LINEAR,
                             flags=capi.CRAW_PIPELINE_FLAG_DENOISE |
                             capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE |
                             capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM)
        fp = frame_params(prof, denoise_strength=1.0, denoise_chroma_strength=0.5,
                          noise_override=NOISE)
        oracle.f32l.set_threads(n)
        t0 = time.perf_counter()
        r = oracle.f32l.process(bayer, pp, fp, want_hist=True)
        dt = time.perf_counter() - t0
        assert r["ret"] == 0
        res[label] = (w * h / 1e6 / dt, n, w, h, dt)
    mp_all, n_all, _, _, dt_all = res["all"]
    mp_1, _, w1, h1, dt1 = res["single"]
    return {
        "value": round(mp_all, 3), "unit": "Mpix/s", "cores": n_all, "kind": "port",
        "sample": "oracle float/glibc mode (lib/ref restated), full path of the same "
                  "workload on one 5760x3240 frame, %d OpenMP threads, %.1f s; "
                  "single-threaded on a 2880x1620 frame of the same scene generator: "
                  "%.3f Mpix/s (%.1f s)" % (n_all, dt_all, mp_1, dt1),
        "single_thread_value": round(mp_1, 3),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; the development path shards by frames "
              "and clip streams across GPUs, and no multi-GPU figure is measured. "
              "Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from libcraw_amd import capi, hip, scene
    from libcraw_amd.pipeline import Context, Pipeline, frame_params, image, pipeline_params

    prof = scene.make_profile()
    bayer = scene.make_frame(W, H, seed=1, a=NOISE[0], b=NOISE[1])
    fmt = capi.CRAW_FORMAT_R32G32B32A32_SFLOAT
    flags = (capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE |
             capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM | capi.CRAW_PIPELINE_FLAG_PROFILE)
    ctx = Context(0)
    pp = pipeline_params(W, H, debayer=capi.CRAW_DEBAYER_RCD,
                         out_space=capi.CRAW_CS_ACES_AP1_LINEAR, out_format=fmt, flags=flags)
    pipe = Pipeline(ctx, pp)
    src, dst = hip.DeviceBuffer(W * H * 2), hip.DeviceBuffer(W * H * 16)
    src.upload(bayer)
    stream = hip.Stream()
    fp = frame_params(prof, denoise_strength=1.0, denoise_chroma_strength=0.5,
                      noise_override=NOISE)
    si = image(src, W, H, capi.CRAW_FORMAT_R16_UNORM)
    di = image(dst, W, H, fmt)

    for _ in range(args.warmup):
        assert pipe.record(stream, si, di, fp) == 0
    stream.synchronize()
    print("bench: warmup done", file=sys.stderr, flush=True)

    e0, e1 = hip.Event(), hip.Event()
    hip.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        assert pipe.record(stream, si, di, fp) == 0
    e1.record(stream)
    stream.synchronize()
    hip.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_ms(e1)
    ms_per_step = gpu_ms / args.steps
    mpix = W * H / 1e6
    value = mpix * args.steps / (gpu_ms / 1e3)

    # per-kernel averages over the last recordings of the timed region
    stages = {}
    for back in range(min(args.steps, capi_slots())):
        for name, count, ms in pipe.stage_times(back):
            s = stages.setdefault(name, [0, 0.0])
            s[0] += count
            s[1] += ms
    stage_ms = {k: v[1] / v[0] for k, v in stages.items()}
    dom = max(stage_ms, key=stage_ms.get)
    algo = ALGO_BYTES_PER_PX.get(dom, 0.0) * W * H
    achieved = algo / (stage_ms[dom] / 1e3) / 1e9
    roofline = {
        "kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
        "algo_bytes_per_launch": int(algo), "avg_launch_ms": round(stage_ms[dom], 4),
    }
    if dom == "denoise_nlm":
        ops = NLM_OPS_PER_SAMPLE * W * H
        tops = ops / (stage_ms[dom] / 1e3) / 1e12
        roofline["valu"] = {"achieved": round(tops, 2), "peak": VALU_PEAK_TOPS,
                            "unit": "T fp32 lane-ops/s", "frac": round(tops / VALU_PEAK_TOPS, 3),
                            "ops_per_sample": NLM_OPS_PER_SAMPLE}

    # plain-form latency: each frame recorded and waited for alone
    lat = []
    for _ in range(max(10, min(args.steps, 50))):
        a, b = hip.Event(), hip.Event()
        a.record(stream)
        assert pipe.record(stream, si, di, fp) == 0
        b.record(stream)
        b.synchronize()
        lat.append(a.elapsed_ms(b))
    lat = np.array(lat)

    out = {
        "metric": "developed Mpix/s, 5.8K full pipeline",
        "value": round(value, 2), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain form, one 5760x3240 RGGB 12-bit frame per step: "
                               "VST+NLM 7x7/3x3 strength 1.0, RCD, chroma denoise 0.5, AWB "
                               "histogram, ACES AP1 + RGC, RGBA32F",
                   "width": W, "height": H, "noise": list(NOISE), "form": "plain"},
        "wall_s": round(wall, 3),
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                       "p99": round(float(np.percentile(lat, 99)), 4), "n": int(lat.size),
                       "form": "plain"},
        "roofline": roofline,
    }
    print("bench: gpu timing done", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        out["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(out))


def capi_slots():
    return 4


if __name__ == "__main__":
    main()
