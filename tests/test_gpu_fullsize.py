"""Correctness at the benchmarked sizes, bit for bit against the oracle.

The smaller parity tests cover every tiling case; these run the workloads
the benchmark times at their own sizes, so that full-size-only indexing is
checked too: 40-row RCD strips over 3240 rows, 224-column NLM bands over
2880x1620 planes, 6000x4000 frames inside one batch arena beside smaller
ones. The oracle (oracle.f32d) runs on the host's cores (OpenMP, at most 16
threads):

  - the 5.8K full pipeline, the benchmark's own frame: VST + NLM 1.0, RCD,
    chroma 0.5, AWB histogram, ACES AP1 with gamut compression, RGBA32F;
  - the 1920x1080 plain MHC form to BT.709 without denoise;
  - the temporal clip's composition at 5760x3240: MHC, BT.2020, temporal
    0.8 + NLM 1.0, four frames with the state threaded through;
  - a batch recording of 1920x1080, 4096x2160, 5760x3240 and 6000x4000
    frames with per-frame noise and white balance: each frame equal to the
    plain form, the 6000x4000 one to the oracle as well.
"""
import os

import numpy as np
import pytest

import oracle
from libcraw_amd import capi, hip, scene
from libcraw_amd.pipeline import Batch, batch_frame, batch_params, frame_params, image

from gpu_util import Runner, assert_bit_exact, context

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
F32 = capi.CRAW_FORMAT_R32G32B32A32_SFLOAT
FULL = (capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE |
        capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM)


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    oracle.f32d.set_threads(THREADS)
    yield
    oracle.f32d.set_threads(1)


def check(run, bayer, fp, what, hist=False, prev=None):
    print("fullsize: %s %dx%d" % (what, run.w, run.h), flush=True)
    assert run.run(bayer, fp) == 0
    ref = oracle.f32d.process(bayer, run.pp, fp, prev=prev, want_hist=hist,
                              want_state=prev is not None or bool(
                                  run.pp.flags & capi.CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE))
    assert ref["ret"] == 0
    assert_bit_exact(run.output(), ref["out"], what)
    if hist:
        for name, a, b in zip(("count", "weight", "bright"), run.pipe.histogram(), ref["hist"]):
            assert np.array_equal(a, b), "%s: histogram %s differs in %d bins" % (
                what, name, int((a != b).sum()))
    return ref


def test_full_pipeline_5760x3240(gpu_available):
    w, h = 5760, 3240
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_RCD, out_space=capi.CRAW_CS_ACES_AP1_LINEAR,
                 flags=FULL)
    fp = frame_params(scene.make_profile(), denoise_strength=1.0, denoise_chroma_strength=0.5,
                      noise_override=(2e-4, 4e-6))
    ref = check(run, scene.make_frame(w, h, seed=1, a=2e-4, b=4e-6), fp, "5.8K full pipeline",
                hist=True)
    # most of the 720 x 405 blocks are samples (the rest clipped or black)
    assert ref["hist"][0].sum() > 0.5 * (w // 8) * (h // 8)


def test_plain_mhc_1920x1080(gpu_available):
    w, h = 1920, 1080
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_MHC, out_space=capi.CRAW_CS_BT709_LINEAR)
    check(run, scene.make_frame(w, h, seed=2), frame_params(scene.make_profile()),
          "1080p MHC BT.709")


def test_temporal_clip_5760x3240(gpu_available):
    w, h = 5760, 3240
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_MHC, out_space=capi.CRAW_CS_BT2020_LINEAR,
                 flags=capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE)
    fp = frame_params(scene.make_profile(), denoise_strength=1.0, denoise_temporal_strength=0.8,
                      noise_override=(4e-4, 1.6e-5))
    state = None
    for i in range(4):
        bayer = scene.clip_frame(w, h, i, a=4e-4, b=1.6e-5)
        state = check(run, bayer, fp, "5.8K temporal clip frame %d" % i, prev=state)["state"]


def test_batch_with_6000x4000(gpu_available):
    sizes = [(6000, 4000), (1920, 1080), (5760, 3240), (4096, 2160)]
    rng = np.random.Generator(np.random.Philox(31))
    fps, mos = [], []
    for i, (w, h) in enumerate(sizes):
        wb = (float(rng.uniform(1.2, 2.6)), 1.0, float(rng.uniform(1.2, 2.6)))
        a, b = float(rng.uniform(1e-4, 8e-4)), float(rng.uniform(2e-6, 3e-5))
        fps.append(frame_params(scene.make_profile(wb=wb), denoise_strength=1.0,
                                denoise_chroma_strength=0.5, noise_override=(a, b)))
        mos.append(scene.make_frame(w, h, seed=60 + i, wb=wb, a=a, b=b))
    src = [hip.DeviceBuffer(w * h * 2) for w, h in sizes]
    dst = [hip.DeviceBuffer(w * h * 16) for w, h in sizes]
    for s, m in zip(src, mos):
        s.upload(m)
    bp = batch_params(len(sizes), 6000, 4000, sum(w * h for w, h in sizes),
                      debayer=capi.CRAW_DEBAYER_RCD, out_space=capi.CRAW_CS_ACES_AP1_LINEAR,
                      out_format=F32, flags=FULL, use_graphs=True)
    b = Batch(context(), bp)
    stream = hip.Stream()
    frames = [batch_frame(image(src[i], w, h, capi.CRAW_FORMAT_R16_UNORM),
                          image(dst[i], w, h, F32), fps[i]) for i, (w, h) in enumerate(sizes)]
    print("fullsize: batch of %s" % sizes, flush=True)
    for _ in range(2):      # the second recording replays the captured graph
        assert b.record(stream, frames) == 0
        stream.synchronize()
    for i, (w, h) in enumerate(sizes):
        out = dst[i].download(np.float32, (h, w, 4))
        run = Runner(w, h, debayer=capi.CRAW_DEBAYER_RCD, out_space=capi.CRAW_CS_ACES_AP1_LINEAR,
                     flags=FULL)
        assert run.run(mos[i], fps[i]) == 0
        assert_bit_exact(out, run.output(), "batch frame %dx%d against plain" % (w, h))
        for name, x, y in zip(("count", "weight", "bright"), b.histogram(i), run.pipe.histogram()):
            assert np.array_equal(x, y), "batch frame %dx%d histogram %s" % (w, h, name)
        if (w, h) == (6000, 4000):
            ref = oracle.f32d.process(mos[i], run.pp, fps[i], want_hist=True)
            assert ref["ret"] == 0
            assert_bit_exact(out, ref["out"], "batch frame 6000x4000 against the oracle")
        del run
