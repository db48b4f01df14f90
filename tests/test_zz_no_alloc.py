"""craw_develop_noalloc(): the allocation-free development against
craw_pipeline_record() of a pipeline built for the same stages, bit for bit
(output and AWB histogram), two calls in a row on one device, the second on
a smaller frame (the scratch array keeps the first call's contents), with
and without the histogram, every debayer; the limits it refuses."""
import ctypes

import numpy as np
import pytest

import oracle
from libcraw_amd import capi, hip, scene
from libcraw_amd.pipeline import frame_params, image

from gpu_util import Runner, assert_bit_exact, context

pytestmark = pytest.mark.gpu

F32 = capi.CRAW_FORMAT_R32G32B32A32_SFLOAT
F16 = capi.CRAW_FORMAT_R16G16B16A16_SFLOAT
WORDS = 3 * 128 * 128


def noalloc(bayer, fp, debayer, space, fmt, with_hist, stream):
    lib = capi.lib()
    h, w = bayer.shape
    bpp = 16 if fmt == F32 else 8
    src, dst = hip.DeviceBuffer(w * h * 2), hip.DeviceBuffer(w * h * bpp)
    src.upload(bayer)
    hist = hip.DeviceBuffer(WORDS * 4) if with_hist else None
    if hist is not None:
        hist.zero()
    hip.synchronize()
    ret = lib.craw_develop_noalloc(context().handle, stream.handle,
                                   ctypes.byref(image(src, w, h, capi.CRAW_FORMAT_R16_UNORM)),
                                   ctypes.byref(image(dst, w, h, fmt)), ctypes.byref(fp),
                                   debayer, space, hist.value if hist is not None else None)
    assert ret == 0, ret
    stream.synchronize()
    if fmt == F32:
        out = dst.download(np.float32, (h, w, 4))
    else:
        out = dst.download(np.uint16, (h, w, 4)).view(np.float16).astype(np.float32)
    hs = hist.download(np.uint32, (3, 128 * 128)) if hist is not None else None
    return out, hs


def reference(bayer, fp, debayer, space, fmt, with_hist):
    h, w = bayer.shape
    run = Runner(w, h, debayer=debayer, out_space=space, half=fmt == F16,
                 flags=capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM if with_hist else 0)
    assert run.run(bayer, fp) == 0
    return run.output(), (np.stack(run.pipe.histogram()) if with_hist else None), run.pp


@pytest.mark.parametrize("debayer", [capi.CRAW_DEBAYER_BILINEAR, capi.CRAW_DEBAYER_MHC,
                                     capi.CRAW_DEBAYER_RCD])
def test_two_calls_equal_the_pipeline(gpu_available, debayer):
    stream = hip.Stream()
    prof = scene.make_profile()
    cases = [((1920, 1080), capi.CRAW_CS_ACES_AP1_LINEAR, F32, True),
             ((250, 98), capi.CRAW_CS_BT709_LINEAR, F32, True),
             ((130, 66), capi.CRAW_CS_BT2020_LINEAR, F16, False)]
    for k, ((w, h), space, fmt, with_hist) in enumerate(cases):
        bayer = scene.make_frame(w, h, seed=90 + k + debayer)
        fp = frame_params(prof, exposure_stops=0.25 * k,
                          denoise_strength=1.0, denoise_chroma_strength=0.5)  # ignored
        got, ghist = noalloc(bayer, fp, debayer, space, fmt, with_hist, stream)
        want, whist, pp = reference(bayer, fp, debayer, space, fmt, with_hist)
        what = "call %d: %dx%d debayer %d" % (k, w, h, debayer)
        assert_bit_exact(got, want, what)
        if with_hist:
            assert np.array_equal(ghist, whist), what + ": histogram"
            assert ghist[0].sum() > 0
        if k == 1:
            ref = oracle.f32d.process(bayer, pp, fp, want_hist=True)
            assert_bit_exact(got, ref["out"], what + " against the oracle")
            assert np.array_equal(ghist, np.stack(ref["hist"]))


def test_limits(gpu_available):
    lib = capi.lib()
    stream = hip.Stream()
    fp = frame_params(scene.make_profile())
    src, dst = hip.DeviceBuffer(64 * 48 * 2), hip.DeviceBuffer(64 * 48 * 16)
    hist = hip.DeviceBuffer(WORDS * 4 + 64)

    def call(w, h, hist_ptr, fmt=F32, space=capi.CRAW_CS_XYZ_D65):
        return lib.craw_develop_noalloc(context().handle, stream.handle,
                                        ctypes.byref(image(src, w, h, capi.CRAW_FORMAT_R16_UNORM,
                                                           pitch=w * 2)),
                                        ctypes.byref(image(dst, w, h, fmt, pitch=w * 16)),
                                        ctypes.byref(fp), capi.CRAW_DEBAYER_MHC, space, hist_ptr)

    # more 8x8 blocks than the scratch holds (refused before anything is read)
    assert call(8200, 4200, hist.value) == -capi.EINVAL
    # a histogram that is not 16-byte aligned
    assert call(64, 48, hist.value + 4) == -capi.EINVAL
    # the pipeline's own checks: an odd width, an output format it cannot write
    assert call(63, 48, hist.value) == -capi.EINVAL
    assert call(64, 48, hist.value, fmt=capi.CRAW_FORMAT_R16_UNORM) == -capi.EINVAL
    assert lib.craw_develop_noalloc(None, None, None, None, None, 0, 0, None) == -capi.EINVAL
    stream.synchronize()
