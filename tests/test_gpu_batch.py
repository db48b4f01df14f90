"""The batched form (include/craw/batch.h) against the plain pipeline.

Bar: every frame of a batch recording equal, bit for bit, to what
craw_pipeline_record() develops from the same frame with the same
parameters (output and AWB histogram), whichever frames share the
recording, in every NLM form, with graphs on or off, and after the
autotuner has changed the launch shapes; temporal denoise per clip stream
equal to a plain pipeline running that clip alone. The plain pipeline is
itself bit-exact to the oracle (test_gpu_parity.py); one case here also
checks a batch frame against the oracle directly.
"""
import numpy as np
import pytest

import oracle
from libcraw_amd import capi, hip, scene
from libcraw_amd.pipeline import (Batch, CrawError, TemporalState, batch_frame,
                                  batch_params, frame_params, image)

from gpu_util import Runner, assert_bit_exact, context, half_to_float

pytestmark = pytest.mark.gpu

F32 = capi.CRAW_FORMAT_R32G32B32A32_SFLOAT
F16 = capi.CRAW_FORMAT_R16G16B16A16_SFLOAT
FULL = (capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE |
        capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM)
# uneven sizes: none a multiple of every tile, the smallest near the minimum
SIZES = [(256, 160), (96, 64), (202, 122), (40, 24), (130, 66)]


class Frames:
    """Device source and destination buffers for a list of frame sizes."""

    def __init__(self, sizes, fmt=F32):
        self.sizes, self.fmt = sizes, fmt
        bpp = 16 if fmt == F32 else 8
        self.src = [hip.DeviceBuffer(w * h * 2) for w, h in sizes]
        self.dst = [hip.DeviceBuffer(w * h * bpp) for w, h in sizes]

    def upload(self, mosaics):
        for buf, m in zip(self.src, mosaics):
            buf.upload(m)

    def images(self, i):
        w, h = self.sizes[i]
        return (image(self.src[i], w, h, capi.CRAW_FORMAT_R16_UNORM),
                image(self.dst[i], w, h, self.fmt))

    def output(self, i):
        w, h = self.sizes[i]
        if self.fmt == F32:
            return self.dst[i].download(np.float32, (h, w, 4))
        return half_to_float(self.dst[i].download(np.uint16, (h, w, 4)))


def plain(size, mosaic, fp, flags, debayer=capi.CRAW_DEBAYER_RCD, half=False,
          space=capi.CRAW_CS_XYZ_D65):
    w, h = size
    run = Runner(w, h, debayer=debayer, flags=flags, half=half, out_space=space)
    assert run.run(mosaic, fp) == 0
    hist = run.pipe.histogram() if flags & capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM else None
    return run.output(), hist


def mosaics(sizes, seed):
    return [scene.make_frame(w, h, seed=seed + i) for i, (w, h) in enumerate(sizes)]


def params_for(prof, i):
    # different strengths per frame: the frame table carries each frame's own
    return frame_params(prof, denoise_strength=0.8 + 0.15 * (i % 3),
                        denoise_chroma_strength=0.5 + 0.1 * (i % 4),
                        exposure_stops=0.25 * (i % 2),
                        noise_override=(2e-4 * (1 + i % 3), 4e-6 * (1 + i % 2)))


def new_batch(sizes, flags, debayer=capi.CRAW_DEBAYER_RCD, graphs=False, fmt=F32,
              max_frames=None, space=capi.CRAW_CS_XYZ_D65):
    mw = max(w for w, _ in sizes)
    mh = max(h for _, h in sizes)
    bp = batch_params(max_frames or len(sizes), mw, mh, sum(w * h for w, h in sizes),
                      debayer=debayer, out_space=space, out_format=fmt, flags=flags,
                      use_graphs=graphs)
    return Batch(context(), bp)


def check_against_plain(batch, frames, fps, mos, flags, debayer, what, half=False,
                        space=capi.CRAW_CS_XYZ_D65):
    for i, size in enumerate(frames.sizes):
        ref, hist = plain(size, mos[i], fps[i], flags, debayer=debayer, half=half, space=space)
        assert_bit_exact(frames.output(i), ref, "%s frame %d %s" % (what, i, size))
        if hist is not None:
            got = batch.histogram(i)
            for name, a, b in zip(("count", "weight", "bright"), got, hist):
                assert np.array_equal(a, b), "%s frame %d: histogram %s differs" % (
                    what, i, name)


@pytest.mark.parametrize("debayer", [capi.CRAW_DEBAYER_BILINEAR, capi.CRAW_DEBAYER_MHC,
                                     capi.CRAW_DEBAYER_RCD])
def test_uneven_frames_equal_plain(gpu_available, debayer):
    prof = scene.make_profile()
    fr = Frames(SIZES)
    mos = mosaics(SIZES, 40 + debayer)
    fr.upload(mos)
    fps = [params_for(prof, i) for i in range(len(SIZES))]
    b = new_batch(SIZES, FULL, debayer=debayer)
    stream = hip.Stream()
    assert b.record(stream, [batch_frame(*fr.images(i), fps[i]) for i in range(len(SIZES))]) == 0
    stream.synchronize()
    check_against_plain(b, fr, fps, mos, FULL, debayer, "debayer %d" % debayer)


def test_batch_frame_equals_oracle(gpu_available):
    prof = scene.make_profile()
    sizes = [(128, 96), (72, 40)]
    fr = Frames(sizes)
    mos = mosaics(sizes, 7)
    fr.upload(mos)
    fps = [params_for(prof, i) for i in range(2)]
    b = new_batch(sizes, FULL)
    stream = hip.Stream()
    assert b.record(stream, [batch_frame(*fr.images(i), fps[i]) for i in range(2)]) == 0
    stream.synchronize()
    for i, (w, h) in enumerate(sizes):
        run = Runner(w, h, flags=FULL)
        ref = oracle.f32d.process(mos[i], run.pp, fps[i], want_hist=True)
        assert ref["ret"] == 0
        assert_bit_exact(fr.output(i), ref["out"], "frame %d vs oracle" % i)
        for a, r in zip(b.histogram(i), ref["hist"]):
            assert np.array_equal(a, r)


@pytest.mark.parametrize("form", ["fast", "fp16"])
def test_nlm_forms_equal_plain(gpu_available, form):
    extra = (capi.CRAW_PIPELINE_FLAG_NLM_FAST if form == "fast"
             else capi.CRAW_PIPELINE_FLAG_NLM_FP16)
    flags = FULL | extra
    prof = scene.make_profile()
    fr = Frames(SIZES, fmt=F16)
    mos = mosaics(SIZES, 90)
    fr.upload(mos)
    fps = [params_for(prof, i) for i in range(len(SIZES))]
    b = new_batch(SIZES, flags, fmt=F16)
    stream = hip.Stream()
    assert b.record(stream, [batch_frame(*fr.images(i), fps[i]) for i in range(len(SIZES))]) == 0
    stream.synchronize()
    check_against_plain(b, fr, fps, mos, flags, capi.CRAW_DEBAYER_RCD, form, half=True)


def test_graph_replays_follow_new_inputs(gpu_available):
    prof = scene.make_profile()
    sizes = SIZES[:3]
    fr = Frames(sizes)
    fps = [params_for(prof, i) for i in range(3)]
    b = new_batch(sizes, FULL, graphs=True)
    stream = hip.Stream()
    frames = [batch_frame(*fr.images(i), fps[i]) for i in range(3)]
    for rep in range(3):
        mos = mosaics(sizes, 200 + 10 * rep)
        fr.upload(mos)
        assert b.record(stream, frames) == 0
        stream.synchronize()
        check_against_plain(b, fr, fps, mos, FULL, capi.CRAW_DEBAYER_RCD, "replay %d" % rep)
    # a different shape captures a new graph; the old shape still replays right
    assert b.record(stream, frames[:2]) == 0
    stream.synchronize()
    mos = mosaics(sizes, 300)
    fr.upload(mos)
    assert b.record(stream, frames) == 0
    stream.synchronize()
    check_against_plain(b, fr, fps, mos, FULL, capi.CRAW_DEBAYER_RCD, "after reshape")


def test_graph_replays_follow_new_buffers_and_params(gpu_available):
    """Replays of one captured shape with other destination buffers and
    other frame parameters each time (what a server's batches see): every
    frame still develops as the plain form develops it."""
    prof = scene.make_profile()
    sizes = SIZES[:3]
    b = new_batch(sizes, FULL, graphs=True)
    stream = hip.Stream()
    for rep in range(3):
        fr = Frames(sizes)
        fps = [params_for(prof, i + rep) for i in range(3)]
        mos = mosaics(sizes, 400 + 10 * rep)
        fr.upload(mos)
        frames = [batch_frame(*fr.images(i), fps[i]) for i in range(3)]
        assert b.record(stream, frames) == 0
        stream.synchronize()
        check_against_plain(b, fr, fps, mos, FULL, capi.CRAW_DEBAYER_RCD, "replay %d" % rep)


@pytest.mark.parametrize("graphs", [False, True])
def test_highlights_and_false_colour(gpu_available, graphs):
    """Frames of uneven sizes with and without highlight reconstruction and
    with 0-2 false colour passes, before and without the chroma stage, in
    one recording: each frame as the plain form develops it. With graphs
    the recording is captured once and replayed with other destination
    buffers, mosaics and frame parameters each time (the highlight and
    false colour launches take a frame's constants as arguments, set again
    before each replay)."""
    flags = (capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE | capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM |
             capi.CRAW_PIPELINE_FLAG_HIGHLIGHT_RECONSTRUCTION |
             capi.CRAW_PIPELINE_FLAG_FALSE_COLOR)
    prof = scene.make_profile()
    sizes = SIZES[:4]
    b = new_batch(sizes, flags, graphs=graphs)
    stream = hip.Stream()
    for rep in range(3):
        fr = Frames(sizes)
        mos = mosaics(sizes, 500 + 10 * rep)
        for m in mos:
            h, w = m.shape
            m[h // 4:h // 2, w // 3:w // 2 + 4 * rep] = 4095           # blown
        fps = [frame_params(prof,
                            flags=capi.CRAW_FRAME_FLAG_HIGHLIGHT_RECONSTRUCTION if i % 2 == 0
                            else 0, false_color_passes=i % 3,
                            denoise_chroma_strength=(0.5 + 0.1 * rep) if i != 1 else 0.0,
                            exposure_stops=0.25 * rep)
               for i in range(len(sizes))]
        fr.upload(mos)
        frames = [batch_frame(*fr.images(i), fps[i]) for i in range(len(sizes))]
        assert b.record(stream, frames) == 0
        stream.synchronize()
        check_against_plain(b, fr, fps, mos, flags, capi.CRAW_DEBAYER_RCD, "rep %d" % rep)


def test_temporal_streams(gpu_available):
    """Two clips of different sizes, interleaved in one batch per time step,
    each equal to a plain pipeline developing that clip alone; a reset and a
    replay on one clip leave the other untouched."""
    flags = capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE
    sizes = [(128, 96), (96, 64)]
    prof = scene.make_profile()
    fr = Frames(sizes)
    b = new_batch(sizes, flags)
    states = [TemporalState(context(), w, h) for w, h in sizes]
    runs = [Runner(w, h, flags=flags) for w, h in sizes]
    stream = hip.Stream()
    # (clip 0 frame flags, clip 1 frame flags) per step
    steps = [(0, 0), (0, 0), ("reset", 0), (0, "replay"), (0, 0)]
    for t, kinds in enumerate(steps):
        mos = [scene.clip_frame(w, h, t, seed=60 + c) for c, (w, h) in enumerate(sizes)]
        fr.upload(mos)
        fps = []
        for c, kind in enumerate(kinds):
            fp = frame_params(prof, denoise_strength=1.0, denoise_temporal_strength=0.6,
                              noise_override=(4e-4, 1.6e-5))
            if kind == "reset":
                fp.flags |= capi.CRAW_FRAME_FLAG_TEMPORAL_RESET
            if kind == "replay":
                fp.denoise_temporal_replay = 1
            fps.append(fp)
        assert b.record(stream, [batch_frame(*fr.images(c), fps[c], states[c])
                                 for c in range(2)]) == 0
        stream.synchronize()
        for c in range(2):
            assert runs[c].run(mos[c], fps[c]) == 0
            assert_bit_exact(fr.output(c), runs[c].output(), "step %d clip %d" % (t, c))


def test_state_untouched_by_a_batch_without_the_pass(gpu_available):
    """A clip's state attached to a recording of a batch built without the
    temporal pass is left as it was (lib/pipeline.c:2308 moves the state
    only on a pipeline with the pass): the clip's next frame still reads the
    frame before."""
    flags = capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE
    size = (128, 96)
    prof = scene.make_profile()
    fr = Frames([size])
    with_pass, without = new_batch([size], flags), new_batch([size], capi.CRAW_PIPELINE_FLAG_DENOISE)
    state = TemporalState(context(), *size)
    run = Runner(*size, flags=flags)
    stream = hip.Stream()
    fp = frame_params(prof, denoise_strength=1.0, denoise_temporal_strength=0.7,
                      noise_override=(4e-4, 1.6e-5))
    for t, b in enumerate((with_pass, without, with_pass)):
        mos = [scene.clip_frame(*size, t, seed=61)]
        fr.upload(mos)
        assert b.record(stream, [batch_frame(*fr.images(0), fp, state)]) == 0
        stream.synchronize()
        if b is with_pass:
            assert run.run(mos[0], fp) == 0
            assert_bit_exact(fr.output(0), run.output(), "step %d" % t)


FORMS = {"exact": 0, "fp16": capi.CRAW_PIPELINE_FLAG_NLM_FP16,
         "fast": capi.CRAW_PIPELINE_FLAG_NLM_FAST}
DN_STAGE = {"exact": "denoise_nlm", "fp16": "denoise_nlm_fp16", "fast": "denoise_nlm_fast"}


@pytest.mark.parametrize("form", sorted(FORMS))
def test_autotune_keeps_outputs(gpu_available, form):
    flags = FULL | FORMS[form]
    prof = scene.make_profile()
    fr = Frames(SIZES)
    mos = mosaics(SIZES, 500)
    fr.upload(mos)
    fps = [params_for(prof, i) for i in range(len(SIZES))]
    b = new_batch(SIZES, flags | capi.CRAW_PIPELINE_FLAG_PROFILE)
    stream = hip.Stream()
    frames = [batch_frame(*fr.images(i), fps[i]) for i in range(len(SIZES))]
    us = b.autotune(stream, frames, iterations=2)
    assert us > 0
    wg, opt = b.launch_shape()
    print("%s: autotuned workgroups per CU %s, denoise options %d" % (form, wg, opt))
    assert b.record(stream, frames) == 0
    stream.synchronize()
    names = [n for n, _, _ in b.stage_times()]
    assert "rcd" in names and DN_STAGE[form] in names
    check_against_plain(b, fr, fps, mos, flags, capi.CRAW_DEBAYER_RCD, "autotuned")


@pytest.mark.parametrize("form", sorted(FORMS))
def test_every_launch_shape_equals_plain(gpu_available, form):
    """Each denoise option (taller pairwise bands, unrolled fast search)
    with one-workgroup-per-tile and persistent grids: the same bits as the
    plain pipeline, which has neither option."""
    flags = FULL | FORMS[form]
    prof = scene.make_profile()
    fr = Frames(SIZES)
    mos = mosaics(SIZES, 520)
    fr.upload(mos)
    fps = [params_for(prof, i) for i in range(len(SIZES))]
    refs = [plain(size, mos[i], fps[i], flags) for i, size in enumerate(SIZES)]
    b = new_batch(SIZES, flags)
    stream = hip.Stream()
    frames = [batch_frame(*fr.images(i), fps[i]) for i in range(len(SIZES))]
    tall, unroll = capi.CRAW_BATCH_DENOISE_TALL, capi.CRAW_BATCH_DENOISE_UNROLL
    for opt in (0, tall, unroll, tall | unroll):
        for wg in ((0, 0, 0, 0), (1, 2, 4, 8), (3, 1, 1, 1)):
            b.set_launch_shape(wg, opt)
            assert b.launch_shape() == (list(wg), opt)
            assert b.record(stream, frames) == 0
            stream.synchronize()
            for i, size in enumerate(SIZES):
                what = "%s options %d grid %s frame %d %s" % (form, opt, wg, i, size)
                assert_bit_exact(fr.output(i), refs[i][0], what)
                got = b.histogram(i)
                for a, r in zip(got, refs[i][1]):
                    assert np.array_equal(a, r), what + ": histogram differs"
    for bad in (((0, 0, 0, 65), 0), ((-1, 0, 0, 0), 0), ((0, 0, 0, 0), 4)):
        with pytest.raises(CrawError):
            b.set_launch_shape(*bad)


@pytest.mark.parametrize("graphs", [False, True])
def test_rcd_colour_from_planes_equals_plain(gpu_available, graphs):
    """RCD to ACES AP1: frames without chroma (strength 0) take the colour
    stage from the debayer's planes, beside frames with chroma and frames
    whose gamut compression is vetoed (fused colour) in one recording;
    each equal to the plain form, over recordings whose mix changes."""
    flags = FULL
    prof = scene.make_profile()
    fr = Frames(SIZES)
    mos = mosaics(SIZES, 700)
    fr.upload(mos)
    b = new_batch(SIZES, flags, graphs=graphs, space=capi.CRAW_CS_ACES_AP1_LINEAR)
    stream = hip.Stream()
    for rec in range(3):
        fps = []
        for i in range(len(SIZES)):
            kind = (i + rec) % 3
            fps.append(frame_params(prof, denoise_strength=1.0,
                                    denoise_chroma_strength=0.6 if kind == 0 else 0.0,
                                    flags=capi.CRAW_FRAME_FLAG_NO_GAMUT_COMPRESSION
                                    if kind == 2 else 0,
                                    noise_override=(2e-4, 4e-6)))
        assert b.record(stream, [batch_frame(*fr.images(i), fps[i])
                                 for i in range(len(SIZES))]) == 0
        stream.synchronize()
        check_against_plain(b, fr, fps, mos, flags, capi.CRAW_DEBAYER_RCD,
                            "recording %d graphs %d" % (rec, graphs),
                            space=capi.CRAW_CS_ACES_AP1_LINEAR)


def test_limits_and_errors(gpu_available):
    prof = scene.make_profile()
    sizes = [(64, 48), (64, 48)]
    fr = Frames(sizes + [(128, 96)])
    b = new_batch(sizes, capi.CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE, max_frames=2)
    stream = hip.Stream()
    fp = frame_params(prof, denoise_temporal_strength=0.5, noise_override=(2e-4, 4e-6))
    ok = [batch_frame(*fr.images(i), fp) for i in range(2)]
    assert b.record(stream, ok) == 0
    # more frames, more pixels, a larger frame than created for
    assert b.record(stream, ok + ok[:1]) == -capi.ENOSPC
    assert b.record(stream, [batch_frame(*fr.images(2), fp)]) == -capi.ENOSPC
    assert b.record(stream, []) == -capi.ENOSPC
    # one state twice in a recording; a state of the wrong size
    st = TemporalState(context(), 64, 48)
    assert b.record(stream, [batch_frame(*fr.images(i), fp, st) for i in range(2)]) == \
        -capi.EINVAL
    wrong = TemporalState(context(), 96, 64)
    assert b.record(stream, [batch_frame(*fr.images(0), fp, wrong)]) == -capi.EINVAL
    # stages this library does not have
    bp = batch_params(2, 64, 48, 64 * 48 * 2, flags=capi.CRAW_PIPELINE_FLAG_SHARPEN)
    with pytest.raises(Exception):
        Batch(context(), bp)
    stream.synchronize()
