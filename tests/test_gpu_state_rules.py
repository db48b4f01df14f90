"""The plain pipeline's temporal-state rules, validation, error codes and
profiling, through the C API, against the oracle.

These are the reference's own checks (tests/test_pipeline.c of
cyanreg/libcraw: the temporal rules 1408-1632 with the in-flight recordings
of 1177-1226, the rejections 3779-3854, profiling 5391-5515), with every
output compared bit for bit with the oracle (oracle.f32d) run on the state
the rule says the recording reads, and not only with the GPU's own run:

  - off both ways: strength 0, a pipeline built without the pass, and the
    first frame of the average are the plain chain;
  - a sequence of eight frames threads the state (next frame reads the set
    the last recording wrote);
  - a replay reads the set the last recording read: frame 4 developed six
    times over at other settings and at its own, each the oracle's frame 4
    on frame 3's state; frame 5 afterwards unchanged;
  - a reset, and the frame after a frame without the pass, are first frames;
  - four recordings in flight on one stream without a wait between them
    equal the sequential run;
  - a recording that fails (EINVAL, ENOTSUP) changes nothing: the next one
    is the frame the sequence would have developed;
  - creation and record-time validation errors;
  - CRAW_PIPELINE_FLAG_PROFILE: output bit for bit the unprofiled one, the
    passes named in the order they ran, EINVAL without the flag or for a
    recording never made.
"""
import ctypes

import numpy as np
import pytest

import oracle
from libcraw_amd import capi, hip, scene
from libcraw_amd.pipeline import Pipeline, frame_params, image, pipeline_params

from gpu_util import Runner, assert_bit_exact, context

pytestmark = pytest.mark.gpu

W, H = 192, 128
A, B = 4e-4, 1.6e-5
ALPHA = 0.8
BOTH = capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE
METHOD = capi.CRAW_DEBAYER_MHC
N = 8


def frames():
    return [scene.clip_frame(W, H, i, square=48, speed=5, a=A, b=B) for i in range(N)]


def params(prof, h=1.0, alpha=ALPHA, **kw):
    return frame_params(prof, denoise_strength=h, denoise_temporal_strength=alpha,
                        noise_override=(A, B), **kw)


class Oracle:
    """The oracle's developments, each on the state a rule names."""

    def __init__(self, flags=BOTH):
        self.pp = pipeline_params(W, H, debayer=METHOD, flags=flags)

    def run(self, bayer, fp, prev=None):
        r = oracle.f32d.process(bayer, self.pp, fp, prev=prev, want_state=True)
        assert r["ret"] == 0
        return r["out"], r["state"]


@pytest.fixture(scope="module")
def seq():
    """Frames, their oracle outputs and the states after each, for the
    plain sequence at h 1, alpha 0.8."""
    prof = scene.make_profile()
    fr = frames()
    orc = Oracle()
    fp = params(prof)
    outs, states, prev = [], [], None
    for f in fr:
        out, prev = orc.run(f, fp, prev)
        outs.append(out)
        states.append(prev)
    return dict(prof=prof, frames=fr, outs=outs, states=states, fp=fp, orc=orc)


def test_off_both_ways(gpu_available, seq):
    prof, f0 = seq["prof"], seq["frames"][0]
    plain = Oracle(flags=0).run(f0, params(prof, 0.0, 0.0))[0]
    run = Runner(W, H, debayer=METHOD, flags=capi.CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE)
    assert run.run(f0, params(prof, 0.0, 0.0)) == 0
    assert_bit_exact(run.output(), plain, "temporal strength 0")
    assert run.run(f0, params(prof, 0.0, ALPHA)) == 0
    assert_bit_exact(run.output(), plain, "first frame of the average alone")
    run = Runner(W, H, debayer=METHOD, flags=0)
    assert run.run(f0, params(prof, 0.0, ALPHA)) == 0
    assert_bit_exact(run.output(), plain, "a pipeline without the flag")


def test_sequence(gpu_available, seq):
    run = Runner(W, H, debayer=METHOD, flags=BOTH)
    for i, f in enumerate(seq["frames"]):
        assert run.run(f, seq["fp"]) == 0
        assert_bit_exact(run.output(), seq["outs"][i], "sequence frame %d" % i)
    # the sequence carried state: frame 7 is not its own first frame
    first = seq["orc"].run(seq["frames"][7], seq["fp"])[0]
    assert not np.array_equal(first, seq["outs"][7])


def test_replay(gpu_available, seq):
    prof, fr, fp = seq["prof"], seq["frames"], seq["fp"]
    run = Runner(W, H, debayer=METHOD, flags=BOTH)
    for i in range(5):
        assert run.run(fr[i], fp) == 0
    # the reference's six replays: (alpha, h) = (0.8, 1), (0.5, 2), (0, 1),
    # (0.8, 1), (1, 0), (0.8, 1); each reads frame 3's state
    for alpha, h in ((0.8, 1.0), (0.5, 2.0), (0.0, 1.0), (0.8, 1.0), (1.0, 0.0), (0.8, 1.0)):
        alt = params(prof, h, alpha, denoise_temporal_replay=True)
        assert run.run(fr[4], alt) == 0
        want = seq["orc"].run(fr[4], alt, prev=seq["states"][3] if alpha > 0 else None)[0]
        assert_bit_exact(run.output(), want, "replay of frame 4 at alpha %g h %g" % (alpha, h))
        original = alpha == ALPHA and h == 1.0
        assert np.array_equal(run.output(), seq["outs"][4]) == original
    assert run.run(fr[5], fp) == 0
    assert_bit_exact(run.output(), seq["outs"][5], "frame 5 after the replays")


def test_reset_and_pass_off(gpu_available, seq):
    prof, fr, fp = seq["prof"], seq["frames"], seq["fp"]
    run = Runner(W, H, debayer=METHOD, flags=BOTH)
    for i in range(6):
        assert run.run(fr[i], fp) == 0
    reset = params(prof, flags=capi.CRAW_FRAME_FLAG_TEMPORAL_RESET)
    assert run.run(fr[6], reset) == 0
    first6 = seq["orc"].run(fr[6], fp)[0]
    assert_bit_exact(run.output(), first6, "a reset frame")
    assert not np.array_equal(run.output(), seq["outs"][6])
    # a frame without the pass leaves nothing for the frame after it
    assert run.run(fr[7], params(prof, 1.0, 0.0)) == 0
    assert_bit_exact(run.output(), seq["orc"].run(fr[7], params(prof, 1.0, 0.0))[0],
                     "a frame without the pass")
    assert run.run(fr[0], fp) == 0
    assert_bit_exact(run.output(), seq["orc"].run(fr[0], fp)[0],
                     "the frame after one without the pass")


def test_four_in_flight(gpu_available, seq):
    """Four recordings back to back on one stream, each with its own source
    and destination, no wait between them."""
    fr, fp = seq["frames"], seq["fp"]
    pp = pipeline_params(W, H, debayer=METHOD, flags=BOTH)
    pipe = Pipeline(context(), pp)
    stream = hip.Stream()
    srcs = [hip.DeviceBuffer(W * H * 2) for _ in range(4)]
    dsts = [hip.DeviceBuffer(W * H * 16) for _ in range(4)]
    for i in range(4):
        srcs[i].upload(fr[i])
    hip.synchronize()
    for i in range(4):
        assert pipe.record(stream, image(srcs[i], W, H, capi.CRAW_FORMAT_R16_UNORM),
                           image(dsts[i], W, H, capi.CRAW_FORMAT_R32G32B32A32_SFLOAT), fp) == 0
    stream.synchronize()
    for i in range(4):
        assert_bit_exact(dsts[i].download(np.float32, (H, W, 4)), seq["outs"][i],
                         "in-flight frame %d" % i)


def test_failed_records_change_nothing(gpu_available, seq):
    prof, fr, fp = seq["prof"], seq["frames"], seq["fp"]
    run = Runner(W, H, debayer=METHOD, flags=BOTH)
    for i in range(3):
        assert run.run(fr[i], fp) == 0
    bare = scene.make_profile()
    for bad, code in ((params(prof, 1.0, 1.5), -capi.EINVAL),
                      (frame_params(bare, denoise_strength=1.0,
                                    denoise_temporal_strength=ALPHA), -capi.ENOTSUP),
                      (params(prof, denoise_temporal_replay=True,
                              flags=capi.CRAW_FRAME_FLAG_TEMPORAL_RESET,
                              exposure_stops=0.0, out_space=99, out_space_set=True),
                       -capi.EINVAL)):
        assert run.run(fr[3], bad) == code
    # a source of the wrong format, a destination of the wrong size
    img_bad = image(run.src, W, H, capi.CRAW_FORMAT_R16G16B16A16_SFLOAT)
    dst = image(run.dst, W, H, run.fmt)
    assert run.pipe.record(run.stream, img_bad, dst, fp) == -capi.EINVAL
    src = image(run.src, W, H, capi.CRAW_FORMAT_R16_UNORM)
    assert run.pipe.record(run.stream, src, image(run.dst, W, H - 2, run.fmt), fp) == -capi.EINVAL
    # the source aliasing the destination
    both = hip.DeviceBuffer(W * H * 16)
    assert run.pipe.record(run.stream, image(both, W, H, capi.CRAW_FORMAT_R16_UNORM),
                           image(both, W, H, run.fmt), fp) == -capi.EINVAL
    assert run.run(fr[3], fp) == 0
    assert_bit_exact(run.output(), seq["outs"][3], "frame 3 after failed recordings")
    assert run.run(fr[4], fp) == 0
    assert_bit_exact(run.output(), seq["outs"][4], "frame 4 after failed recordings")


def test_noise_model_from_the_profile(gpu_available):
    """No noise_override: the profile's table at its exposure index gives
    the model (lib/pipeline.c:1391-1409), on the GPU as in the oracle."""
    prof = scene.make_profile()
    for i, (iso, a, b) in enumerate(((400.0, 1.1e-4, 2.5e-6), (1600.0, 3.9e-4, 1.2e-5),
                                     (6400.0, 1.5e-3, 6.0e-5))):
        prof.noise_knots[i].iso, prof.noise_knots[i].a, prof.noise_knots[i].b = iso, a, b
    prof.nb_noise_knots = 3
    orc = Oracle()
    run = Runner(W, H, debayer=METHOD, flags=BOTH)
    prev = None
    for i, ei in enumerate((800.0, 800.0, 3200.0)):
        prof.as_shot_ei = ei
        fp = frame_params(prof, denoise_strength=1.0, denoise_temporal_strength=ALPHA)
        bayer = scene.clip_frame(W, H, i, square=48, speed=5, a=3e-4, b=9e-6)
        assert run.run(bayer, fp) == 0
        want, prev = orc.run(bayer, fp, prev)
        assert_bit_exact(run.output(), want, "profile noise table, frame %d ei %g" % (i, ei))


def test_creation_validation(gpu_available):
    lib = capi.lib()
    ctx = context()
    h = ctypes.c_void_p()

    def create(**kw):
        pp = pipeline_params(W, H)
        for k, v in kw.items():
            setattr(pp, k, v)
        ret = lib.craw_pipeline_create(ctx.handle, ctypes.byref(pp), ctypes.byref(h))
        if ret == 0:
            lib.craw_pipeline_free(ctypes.byref(h))
        return ret

    assert create(cfa=2) == -capi.ENOTSUP                        # BGGR
    assert create(width=W + 1) == -capi.EINVAL                   # odd width
    assert create(height=6) == -capi.EINVAL
    assert create(out_format=capi.CRAW_FORMAT_R16_UNORM) == -capi.EINVAL
    assert create(out_format=77) == -capi.EINVAL
    assert create(width=8, height=8, flags=capi.CRAW_PIPELINE_FLAG_DENOISE) == -capi.EINVAL
    assert create(width=10, height=10, flags=BOTH) == -capi.EINVAL
    assert create(width=12, height=12, flags=BOTH) == 0    # 2 (S + P + 2), the minimum
    assert create(width=16, height=16, flags=capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE) == \
        -capi.EINVAL
    assert create(flags=capi.CRAW_PIPELINE_FLAG_SHARPEN) == -capi.ENOTSUP
    assert create(nb_awb_slots=5, flags=capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM) == -capi.EINVAL
    assert create() == 0
    # the histogram of a pipeline without the flag
    pipe = Pipeline(ctx, pipeline_params(W, H))
    hist = capi.CRAWAwbHistogram()
    assert lib.craw_pipeline_get_awb_histogram(pipe.handle, ctypes.byref(hist)) == -capi.EINVAL
    # and one with it, before any recording
    pipe = Pipeline(ctx, pipeline_params(W, H, flags=capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM))
    assert lib.craw_pipeline_get_awb_histogram(pipe.handle, ctypes.byref(hist)) == -capi.EINVAL


FULL = (capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM | capi.CRAW_PIPELINE_FLAG_DENOISE |
        capi.CRAW_PIPELINE_FLAG_FALSE_COLOR | capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE |
        capi.CRAW_PIPELINE_FLAG_HIGHLIGHT_RECONSTRUCTION)


def test_profiling(gpu_available):
    """Every pass of the chain once, in the order it ran; the output is the
    unprofiled pipeline's bit for bit."""
    lib = capi.lib()
    prof = scene.make_profile()
    bayer = scene.make_frame(W, H, seed=77)
    fp = frame_params(prof, flags=capi.CRAW_FRAME_FLAG_HIGHLIGHT_RECONSTRUCTION,
                      noise_override=(2e-4, 4e-6), denoise_strength=1.0, false_color_passes=2,
                      denoise_chroma_strength=0.5)
    plain = Runner(W, H, debayer=capi.CRAW_DEBAYER_RCD, flags=FULL)
    st = (capi.CRAWStageTime * 32)()
    assert lib.craw_pipeline_get_stage_times(plain.pipe.handle, st, 32) == -capi.EINVAL
    assert plain.run(bayer, fp) == 0
    want = plain.output()
    pp = pipeline_params(W, H, debayer=capi.CRAW_DEBAYER_RCD,
                         flags=FULL | capi.CRAW_PIPELINE_FLAG_PROFILE)
    r = oracle.f32d.process(bayer, pp, fp, want_hist=True)
    assert_bit_exact(want, r["out"], "the unprofiled chain")

    run = Runner(W, H, debayer=capi.CRAW_DEBAYER_RCD, flags=FULL | capi.CRAW_PIPELINE_FLAG_PROFILE)
    assert lib.craw_pipeline_get_stage_times(run.pipe.handle, st, 32) == -capi.EINVAL
    chain = ["denoise_nlm", "highlight_field", "highlight_push", "highlight_fill", "rcd",
             "false_color", "chroma_blur", "chroma_up_color", "awb_hist", "awb_export"]
    for k in range(3):
        assert run.run(bayer, fp) == 0
        assert np.array_equal(run.output(), want), "profiling changed the frame"
        times = run.pipe.stage_times(0)
        names = [n for n, _, _ in times]
        # the first recording also clears the histogram it starts from
        assert names == (["awb_clear"] if k == 0 else []) + chain, names
        assert all(c == 1 for _, c, _ in times)
        assert all(ms >= 0.0 for _, _, ms in times) and sum(ms for _, _, ms in times) > 0.0
        g = run.pipe.histogram()
        for a_, b_ in zip(g, r["hist"]):
            assert np.array_equal(a_, b_)
    assert lib.craw_pipeline_get_stage_times_n(run.pipe.handle, 4, st, 32) == -capi.EINVAL
    n = lib.craw_pipeline_get_stage_times_n(run.pipe.handle, 2, st, 32)
    assert n == len(chain) + 1
    # EAGAIN before completion, never a wait: a frame recorded and asked for
    # at once is either still running or done
    big = Runner(2048, 1536, flags=capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_PROFILE)
    assert big.run(scene.make_frame(2048, 1536, seed=3),
                   frame_params(prof, denoise_strength=1.0, noise_override=(2e-4, 4e-6)),
                   sync=False) == 0
    ret = lib.craw_pipeline_get_stage_times(big.pipe.handle, st, 32)
    assert ret == -capi.EAGAIN or ret > 0
    big.stream.synchronize()
    assert lib.craw_pipeline_get_stage_times(big.pipe.handle, st, 32) > 0
