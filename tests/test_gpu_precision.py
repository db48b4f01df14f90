"""The precision form (CRAW_PIPELINE_FLAG_NLM_FP16) without double
arithmetic on the device: its colour, chroma, histogram and highlight stages
take their transcendentals from the hardware's float exp2/log2 (kf_exp,
kf_log, kf_pow in k_common.h), its NLM is packed FP16 with FP32
accumulation (test_gpu_nlm.py holds that stage to 2^-11 of clip). Errors are
stated against the FP64 oracle (oracle.f64), in units of the clip level
(1.0 of the balanced camera RGB):

  - the float stages alone (the form without the denoise stage: RCD, chroma
    0.5, ACES AP1 with gamut compression, AWB histogram): output within
    2^-16 of clip of the FP64 rerun; histogram samples that land in another
    bin than the oracle's correctly rounded logarithm puts them: at most one
    in 10^4 (counted and printed);
  - the whole form (FP16 NLM as well) on a neutral profile (unit gains,
    identity matrix, XYZ out, so the output is the developed camera RGB):
    within 2^-11 of clip of the FP64 rerun.
"""
import numpy as np
import pytest

import oracle
from libcraw_amd import capi, scene
from libcraw_amd.pipeline import frame_params

from gpu_util import Runner

pytestmark = pytest.mark.gpu

FP16 = capi.CRAW_PIPELINE_FLAG_NLM_FP16


def moved_samples(got, ref):
    """Samples in a different bin: half the L1 distance of the counts."""
    return int(np.abs(got.astype(np.int64) - ref.astype(np.int64)).sum() // 2)


@pytest.mark.parametrize("size", [(256, 160), (1032, 776), (5760, 3240)])
def test_float_stages_against_fp64(gpu_available, size):
    w, h = size
    flags = FP16 | capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE | capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_RCD, out_space=capi.CRAW_CS_ACES_AP1_LINEAR,
                 flags=flags)
    bayer = scene.make_frame(w, h, seed=17, a=8e-4, b=3e-5)
    fp = frame_params(scene.make_profile(), denoise_chroma_strength=0.5)
    assert run.run(bayer, fp) == 0
    out = run.output().astype(np.float64)
    oracle.f64.set_threads(16)
    oracle.f32d.set_threads(16)
    r64 = oracle.f64.process(bayer, run.pp, fp)
    r32 = oracle.f32d.process(bayer, run.pp, fp, want_hist=True)
    assert r64["ret"] == 0 and r32["ret"] == 0
    err = float(np.abs(out - r64["out"]).max())
    err32 = float(np.abs(r32["out"].astype(np.float64) - r64["out"]).max())
    hist = run.pipe.histogram()
    moved = moved_samples(hist[0], r32["hist"][0])
    total = int(r32["hist"][0].sum())
    print("precision form, float stages, %dx%d: max |gpu - fp64| = %.3g (2^%.1f) of clip, "
          "float oracle %.3g; %d of %d histogram samples in another bin"
          % (w, h, err, np.log2(max(err, 1e-30)), err32, moved, total))
    assert err <= 2.0 ** -16
    assert int(hist[0].sum()) == total
    assert moved <= max(1, total // 10000)


@pytest.mark.parametrize("size", [(1032, 776), (5760, 3240)])
def test_whole_form_against_fp64(gpu_available, size):
    w, h = size
    flags = (FP16 | capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE |
             capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM)
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_RCD, out_space=capi.CRAW_CS_XYZ_D65,
                 flags=flags)
    prof = scene.make_profile(wb=(1.0, 1.0, 1.0), matrix=np.eye(3, dtype=np.float32))
    bayer = scene.make_frame(w, h, seed=23, wb=(1.0, 1.0, 1.0), a=2e-4, b=4e-6)
    fp = frame_params(prof, denoise_strength=1.0, denoise_chroma_strength=0.5,
                      noise_override=(2e-4, 4e-6))
    assert run.run(bayer, fp) == 0
    out = run.output().astype(np.float64)
    oracle.f64.set_threads(16)
    r64 = oracle.f64.process(bayer, run.pp, fp)
    assert r64["ret"] == 0
    err = np.abs(out - r64["out"])[..., :3]
    print("precision form, whole, %dx%d: max |gpu - fp64| = %.3g (2^%.1f) of clip, "
          "99.99th percentile %.3g" % (w, h, float(err.max()), np.log2(float(err.max())),
                                       float(np.percentile(err, 99.99))))
    assert float(err.max()) <= 2.0 ** -11
