"""The denoised mosaic itself, read back through the C API.

With the 12-bit levels (black 512, white 4095), unit gains and an identity
matrix, XYZ output and the bilinear debayer, the output's own channel at
every site is the prep of the denoised sample, (code / 65535 - black) /
range, which inverts to the code exactly (clipped at white, as the prep
clips), so the 16-bit mosaic the denoise kernel wrote is recovered from the
frame libcraw develops.

  - FP32 form: equal to the oracle's deterministic float mode, every code;
  - fast FP32 and packed-FP16 forms: every code within one of the FP64
    rerun of the reference, and within 2^-11 of the clip level (1.75 codes
    of the 3583 between black and white) of the FP64 level before its
    rounding, at low, mid and high noise and with a temporal history.
"""
import ctypes

import numpy as np
import pytest

import oracle
from libcraw_amd import capi, scene
from libcraw_amd.pipeline import frame_params

from gpu_util import Runner

pytestmark = pytest.mark.gpu

FORMS = {"exact": 0, "fast": capi.CRAW_PIPELINE_FLAG_NLM_FAST,
         "fp16": capi.CRAW_PIPELINE_FLAG_NLM_FP16}


BLACK, WHITE = np.float32(512 / 65535.0), np.float32(4095 / 65535.0)


def identity_profile():
    p = capi.CRAWProfile()
    capi.lib().craw_profile_init(ctypes.byref(p))
    for s in range(4):
        p.black_level[s], p.white_level[s] = BLACK, WHITE
    return p


def recover_mosaic(out):
    h, w, _ = out.shape
    s = (np.arange(h)[:, None] & 1) * 2 + (np.arange(w)[None, :] & 1)
    ch = np.choose(s, [0, 1, 1, 2])
    v = np.take_along_axis(out, ch[..., None], axis=2)[..., 0].astype(np.float64)
    raw = (v * np.float64(WHITE - BLACK) + np.float64(BLACK)) * 65535.0
    code = np.rint(raw)
    assert np.all(np.abs(raw - code) < 1e-2)
    return code.astype(np.int64)


def clip_white(codes):
    return np.minimum(codes, 4095)


def gpu_sequence(form, frames, strength, alpha, a, b):
    h, w = frames[0].shape
    flags = capi.CRAW_PIPELINE_FLAG_DENOISE | FORMS[form]
    if alpha > 0:
        flags |= capi.CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_BILINEAR, flags=flags)
    prof = identity_profile()
    outs = []
    for f in frames:
        fp = frame_params(prof, denoise_strength=strength, denoise_temporal_strength=alpha,
                          noise_override=(a, b))
        assert run.run(f, fp) == 0
        outs.append(recover_mosaic(run.output()))
    return outs


def oracle_sequence(mode, frames, strength, alpha, a, b):
    prof = identity_profile()
    prev, outs = None, []
    for f in frames:
        ret, out, nxt, _ = mode.denoise_temporal(f, prof, a, b, strength, alpha_max=alpha,
                                                 prev=prev, want_state=alpha > 0)
        assert ret == 0
        outs.append(clip_white(out.astype(np.int64)))
        prev = nxt
    return outs


NOISE = [(1e-4, 2e-6), (2e-4, 4e-6), (8e-4, 3e-5)]


@pytest.mark.parametrize("a,b", NOISE)
def test_exact_form_equals_oracle(gpu_available, a, b):
    frames = [scene.make_frame(160, 128, seed=12, a=a, b=b)]
    g = gpu_sequence("exact", frames, 1.0, 0.0, a, b)
    o = oracle_sequence(oracle.f32d, frames, 1.0, 0.0, a, b)
    assert np.array_equal(g[0], o[0])


@pytest.mark.parametrize("size", [(912, 272), (1000, 300), (458, 140)])
def test_exact_form_large_tiles(gpu_available, size):
    """Planes with whole tiles away from every border (no mirror in the
    staging), partial tiles at the right and bottom, and sizes that are
    not multiples of the tile."""
    a, b = 2e-4, 4e-6
    w, h = size
    frames = [scene.make_frame(w, h, seed=21, a=a, b=b)]
    g = gpu_sequence("exact", frames, 1.0, 0.0, a, b)
    o = oracle_sequence(oracle.f32d, frames, 1.0, 0.0, a, b)
    d = g[0] != o[0]
    assert not d.any(), "%d samples differ, first at %s" % (int(d.sum()), np.argwhere(d)[0])


CLIP = 4095.0 - 512.0      # the clip level above black, in codes


def fp64_levels(frames, strength, alpha, a, b):
    """The FP64 rerun's codes and the levels they round (clipped at white
    as the recovery clips)."""
    prof = identity_profile()
    prev, outs = None, []
    for f in frames:
        ret, out, nxt, lvl = oracle.f64.denoise_temporal(f, prof, a, b, strength,
                                                         alpha_max=alpha, prev=prev,
                                                         want_state=alpha > 0, want_level=True)
        assert ret == 0
        outs.append((clip_white(out.astype(np.int64)), np.minimum(lvl, 4095.0)))
        prev = nxt
    return outs


def check_bound(g, o, lvl, what):
    """Every code within one of the FP64 rerun's, and within 2^-11 of the
    clip level of the FP64 level before its rounding (the rounding's half
    code included)."""
    d = np.abs(g - o)
    assert d.max() <= 1, "%s: %d samples off by more than one code" % (what, int((d > 1).sum()))
    err = np.abs(g - lvl) / CLIP
    assert err.max() <= 2.0 ** -11, "%s: error %.3g of clip" % (what, err.max())


@pytest.mark.parametrize("form", ["fast", "fp16"])
@pytest.mark.parametrize("a,b", NOISE)
def test_fast_forms_within_bound_of_fp64(gpu_available, form, a, b):
    frames = [scene.make_frame(224, 160, seed=13, a=a, b=b)]
    g = gpu_sequence(form, frames, 1.0, 0.0, a, b)[0]
    o, lvl = fp64_levels(frames, 1.0, 0.0, a, b)[0]
    check_bound(g, o, lvl, "%s a=%g" % (form, a))


@pytest.mark.parametrize("form", ["exact", "fast", "fp16"])
def test_temporal_forms(gpu_available, form):
    a, b = 4e-4, 1.6e-5
    frames = [scene.clip_frame(96, 80, i, square=32, speed=4, a=a, b=b) for i in range(4)]
    g = gpu_sequence(form, frames, 1.0, 0.8, a, b)
    if form == "exact":
        o = oracle_sequence(oracle.f32d, frames, 1.0, 0.8, a, b)
        for i in range(4):
            assert np.array_equal(g[i], o[i]), "frame %d" % i
    else:
        for i, (o, lvl) in enumerate(fp64_levels(frames, 1.0, 0.8, a, b)):
            check_bound(g[i], o, lvl, "%s frame %d" % (form, i))


@pytest.mark.parametrize("size,nframes", [((1000, 300), 3), ((458, 140), 3), ((40, 520), 3),
                                          ((5760, 3240), 2)])
def test_fp16_history_bands_within_bound_of_fp64(gpu_available, size, nframes):
    """The precision form with a temporal history runs on 224x18 bands (a
    shorter band than the exact form's 30 rows, so that three workgroups
    share a CU): many bands, partial last bands across and down, a plane
    narrower than a band's reach, and the benchmarked 5760x3240 frame, each
    frame of a moving clip within the bound."""
    a, b = 4e-4, 1.6e-5
    w, h = size
    oracle.f64.set_threads(16)
    sq = max(8, min(w, h, 400) // 4 * 2)
    frames = [scene.clip_frame(w, h, i, square=sq, speed=4, a=a, b=b) for i in range(nframes)]
    g = gpu_sequence("fp16", frames, 1.0, 0.8, a, b)
    for i, (o, lvl) in enumerate(fp64_levels(frames, 1.0, 0.8, a, b)):
        check_bound(g[i], o, lvl, "fp16 history %dx%d frame %d" % (w, h, i))


@pytest.mark.parametrize("size,strength", [((912, 272), 1.0), ((1000, 300), 0.6), ((458, 140), 1.0),
                                           ((40, 520), 1.0), ((5760, 3240), 1.0)])
def test_fp16_pairwise_bands_within_bound_of_fp64(gpu_available, size, strength):
    """The precision form's spatial modes run on the pairwise band kernel
    (224x48 bands, packed-FP16 distances): several bands, partial last
    bands across and down, a plane narrower than a band's reach, and the
    benchmarked 5760x3240 frame, each within the bound."""
    a, b = 2e-4, 4e-6
    w, h = size
    oracle.f64.set_threads(16)
    frames = [scene.make_frame(w, h, seed=29, a=a, b=b)]
    g = gpu_sequence("fp16", frames, strength, 0.0, a, b)[0]
    o, lvl = fp64_levels(frames, strength, 0.0, a, b)[0]
    check_bound(g, o, lvl, "fp16 %dx%d h %g" % (w, h, strength))
