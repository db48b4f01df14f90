"""GPU parity of the false colour stage (lib/ref/ref_false_color.c of the
reference, after the debayer and before the chroma denoise): the output and
the AWB histogram bit for bit the oracle's (oracle.f32d.process) for one to
five passes, every debayer, with and without the chroma stage after it, at
sizes that leave partial 64x32 tiles and at the 8x8 minimum (the apron then
mirrors more than once); both ways of switching it off reproduce the plain
chain; six passes are refused (the batched form: test_gpu_batch.py)."""
import ctypes

import numpy as np
import pytest

from libcraw_amd import capi, scene
from libcraw_amd.pipeline import frame_params

from gpu_util import Runner, assert_bit_exact
from test_gpu_parity import _check

pytestmark = pytest.mark.gpu

METHODS = [capi.CRAW_DEBAYER_BILINEAR, capi.CRAW_DEBAYER_MHC, capi.CRAW_DEBAYER_RCD]
FC = capi.CRAW_PIPELINE_FLAG_FALSE_COLOR
AWB = capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM
CHROMA = capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("passes", [1, 2, 3, 4, 5])
def test_passes_to_the_output(gpu_available, method, passes):
    w, h = 200, 136
    run = Runner(w, h, debayer=method, flags=FC | AWB)
    bayer = scene.make_frame(w, h, seed=20 + passes)
    fp = frame_params(scene.make_profile(), false_color_passes=passes)
    _check(run, bayer, fp, "method %d passes %d" % (method, passes), hist=True)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("passes", [1, 2, 5])
def test_passes_then_chroma(gpu_available, method, passes):
    w, h = 200, 136
    run = Runner(w, h, debayer=method, out_space=capi.CRAW_CS_ACES_AP1_LINEAR,
                 flags=FC | CHROMA | AWB)
    bayer = scene.make_frame(w, h, seed=40 + passes)
    fp = frame_params(scene.make_profile(), false_color_passes=passes,
                      denoise_chroma_strength=0.6)
    _check(run, bayer, fp, "method %d passes %d + chroma" % (method, passes), hist=True)


@pytest.mark.parametrize("size", [(8, 8), (24, 16), (70, 34), (130, 66), (66, 130)])
@pytest.mark.parametrize("passes", [1, 5])
def test_sizes(gpu_available, size, passes):
    w, h = size
    prof = scene.make_profile()
    bayer = scene.make_frame(w, h, seed=w + h)
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_MHC, flags=FC | AWB)
    _check(run, bayer, frame_params(prof, false_color_passes=passes),
           "%dx%d passes %d" % (w, h, passes), hist=True)
    if min(w, h) >= 20:
        run = Runner(w, h, debayer=capi.CRAW_DEBAYER_RCD, flags=FC | CHROMA, half=True)
        _check(run, bayer, frame_params(prof, false_color_passes=passes,
                                        denoise_chroma_strength=1.0),
               "%dx%d passes %d + chroma, RGBA16F" % (w, h, passes))


def test_off_both_ways(gpu_available):
    """tests/test_pipeline.c:3908-3935: passes on a pipeline without the
    flag, and no passes on one with it, give the plain chain."""
    w, h = 128, 96
    prof = scene.make_profile()
    bayer = scene.make_frame(w, h, seed=4)
    plain = Runner(w, h, debayer=capi.CRAW_DEBAYER_MHC)
    assert plain.run(bayer, frame_params(prof)) == 0
    want = plain.output()
    for flags, passes in ((0, 5), (FC, 0)):
        run = Runner(w, h, debayer=capi.CRAW_DEBAYER_MHC, flags=flags)
        out = _check(run, bayer, frame_params(prof, false_color_passes=passes),
                     "flags %d passes %d" % (flags, passes))
        assert_bit_exact(out, want, "off: flags %d passes %d" % (flags, passes))
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_MHC, flags=FC)
    out = _check(run, bayer, frame_params(prof, false_color_passes=2), "on")
    assert not np.array_equal(out, want)


def test_too_many_passes(gpu_available):
    """tests/test_pipeline.c:4053-4080: more passes than the stage has are an
    error, not a clamp; the pipeline records normally afterwards."""
    w, h = 64, 64
    prof = scene.make_profile()
    bayer = scene.make_frame(w, h, seed=6)
    run = Runner(w, h, flags=FC)
    assert run.run(bayer, frame_params(prof, false_color_passes=6)) == -22
    _check(run, bayer, frame_params(prof, false_color_passes=5), "after a refused record")
