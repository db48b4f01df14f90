"""GPU parity of the highlight reconstruction (lib/ref/ref_highlight.c of
the reference, before the debayer): the developed output and the AWB
histogram bit for bit the oracle's (oracle.f32d.process) on the
reference's highlight scene (G-only and fully blown blobs), with every
debayer, after the denoise, before the false colour and chroma stages, at
sizes from the 8x8 minimum (a two-level field) to ones with ten levels
and partial tiles everywhere; a scene with nothing clipped; both ways of
switching it off reproduce the plain chain (the batched form:
test_gpu_batch.py)."""
import numpy as np
import pytest

from libcraw_amd import capi, scene
from libcraw_amd.pipeline import frame_params

import hl_scene as sc
from gpu_util import Runner, assert_bit_exact
from test_gpu_parity import _check

pytestmark = pytest.mark.gpu

METHODS = [capi.CRAW_DEBAYER_BILINEAR, capi.CRAW_DEBAYER_MHC, capi.CRAW_DEBAYER_RCD]
HL = capi.CRAW_PIPELINE_FLAG_HIGHLIGHT_RECONSTRUCTION
FHL = capi.CRAW_FRAME_FLAG_HIGHLIGHT_RECONSTRUCTION
AWB = capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM


@pytest.mark.parametrize("method", METHODS)
def test_highlight_scene(gpu_available, method):
    m = sc.mosaic(1.0)
    run = Runner(sc.W, sc.H, debayer=method, flags=HL | AWB)
    out = _check(run, m, frame_params(sc.identity_profile(), flags=FHL),
                 "highlights, method %d" % method, hist=True)
    assert out[..., 1].max() > 1.5                 # the blown core rebuilt above the clip


@pytest.mark.parametrize("scale", [0.2, 0.7, 1.0, 3.0])
def test_scales(gpu_available, scale):
    m = sc.mosaic(scale)
    run = Runner(sc.W, sc.H, debayer=capi.CRAW_DEBAYER_RCD,
                 out_space=capi.CRAW_CS_ACES_AP1_LINEAR, flags=HL)
    _check(run, m, frame_params(sc.identity_profile(), flags=FHL), "scale %g" % scale)


def test_with_the_other_stages(gpu_available):
    """Denoise before (the field is built from the denoised mosaic), false
    colour and chroma after, RGBA16F out."""
    m = sc.mosaic(1.0)
    flags = (HL | capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_FALSE_COLOR |
             capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE | AWB)
    run = Runner(sc.W, sc.H, debayer=capi.CRAW_DEBAYER_RCD, flags=flags, half=True)
    fp = frame_params(sc.identity_profile(), flags=FHL, denoise_strength=1.0,
                      noise_override=(2e-4, 4e-6), false_color_passes=2,
                      denoise_chroma_strength=0.5)
    _check(run, m, fp, "highlights + denoise + false colour + chroma", hist=True)


@pytest.mark.parametrize("size", [(8, 8), (24, 16), (70, 34), (130, 66), (520, 300),
                                  (1032, 776)])
def test_sizes(gpu_available, size):
    w, h = size
    rng = np.random.default_rng(w * h)
    bayer = scene.make_frame(w, h, seed=w + h, a=2e-4, b=4e-6)
    # blown patches: whole quads at and above white, some single sites
    bayer[h // 4:h // 4 + max(2, h // 8), w // 3:w // 3 + max(2, w // 6)] = 4095
    idx = rng.integers(0, bayer.size, max(1, bayer.size // 300))
    bayer.reshape(-1)[idx] = 4095
    prof = scene.make_profile()
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_MHC, flags=HL | AWB)
    _check(run, bayer, frame_params(prof, flags=FHL), "%dx%d" % (w, h), hist=True)


def test_off_both_ways(gpu_available):
    m = sc.mosaic(1.0)
    prof = sc.identity_profile()
    plain = Runner(sc.W, sc.H, debayer=capi.CRAW_DEBAYER_MHC)
    assert plain.run(m, frame_params(prof)) == 0
    want = plain.output()
    for flags, fflags in ((HL, 0), (0, FHL)):
        run = Runner(sc.W, sc.H, debayer=capi.CRAW_DEBAYER_MHC, flags=flags)
        assert run.run(m, frame_params(prof, flags=fflags)) == 0
        assert_bit_exact(run.output(), want, "off: flags %d frame flags %d" % (flags, fflags))
