"""GPU parity: libcraw.so's kernels against the oracle, through the C API.

Bar: the float32 output and the AWB histogram bit for bit equal to the
oracle's deterministic-math float mode (oracle.f32d), which restates
cyanreg/libcraw's lib/ref; inputs are the seeded synthetic frames of
libcraw_amd.scene at sizes the oracle finishes in seconds, including sizes
that are not multiples of any tile, the 8x8 minimum, and clipped and
below-black samples.
"""
import numpy as np
import pytest

import oracle
from libcraw_amd import capi, scene
from libcraw_amd.pipeline import frame_params

from gpu_util import Runner, assert_bit_exact

pytestmark = pytest.mark.gpu

METHODS = [capi.CRAW_DEBAYER_BILINEAR, capi.CRAW_DEBAYER_MHC, capi.CRAW_DEBAYER_RCD]
SPACES = [capi.CRAW_CS_XYZ_D65, capi.CRAW_CS_BT709_LINEAR, capi.CRAW_CS_ACES_AP1_LINEAR]


def _check(run, bayer, fp, what, hist=False):
    assert run.run(bayer, fp) == 0
    ref = oracle.f32d.process(bayer, run.pp, fp, want_hist=hist)
    assert ref["ret"] == 0
    out = run.output()
    if run.half:
        assert_bit_exact(out, ref["out"].astype(np.float16).astype(np.float32), what)
    else:
        assert_bit_exact(out, ref["out"], what)
    if hist:
        g = run.pipe.histogram()
        for name, a, b in zip(("count", "weight", "bright"), g, ref["hist"]):
            assert np.array_equal(a, b), "%s: histogram %s differs (%d bins)" % (
                what, name, int((a != b).sum()))
        assert g[0].sum() > 0
    return out


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("space", SPACES)
def test_debayer_colour_awb(gpu_available, method, space):
    w, h = 200, 136
    run = Runner(w, h, debayer=method, out_space=space,
                 flags=capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM)
    prof = scene.make_profile()
    bayer = scene.make_frame(w, h, seed=11 + method)
    _check(run, bayer, frame_params(prof), "method %d space %d" % (method, space), hist=True)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("size", [(8, 8), (24, 16), (70, 34), (130, 66)])
def test_odd_tilings(gpu_available, method, size):
    w, h = size
    run = Runner(w, h, debayer=method, flags=capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM)
    prof = scene.make_profile()
    bayer = scene.make_frame(w, h, seed=5)
    _check(run, bayer, frame_params(prof), "%dx%d method %d" % (w, h, method))


def test_half_output(gpu_available):
    w, h = 96, 64
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_MHC, out_space=capi.CRAW_CS_BT2020_LINEAR,
                 half=True)
    prof = scene.make_profile()
    _check(run, scene.make_frame(w, h, seed=2), frame_params(prof), "RGBA16F")


def test_frame_flags(gpu_available):
    w, h = 96, 64
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_RCD, out_space=capi.CRAW_CS_ACES_AP1_LINEAR)
    prof = scene.make_profile()
    bayer = scene.make_frame(w, h, seed=3)
    for flags in (capi.CRAW_FRAME_FLAG_CLAMP_NEGATIVE, capi.CRAW_FRAME_FLAG_NO_GAMUT_COMPRESSION):
        _check(run, bayer, frame_params(prof, flags=flags, exposure_stops=0.5),
               "frame flags %d" % flags)


@pytest.mark.parametrize("method", METHODS)
def test_denoise_spatial(gpu_available, method):
    w, h = 128, 96
    run = Runner(w, h, debayer=method, flags=capi.CRAW_PIPELINE_FLAG_DENOISE |
                 capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM)
    prof = scene.make_profile()
    bayer = scene.make_frame(w, h, seed=21)
    for strength in (1.0, 0.5):
        _check(run, bayer, frame_params(prof, denoise_strength=strength,
                                        noise_override=(2e-4, 4e-6)),
               "nlm strength %g" % strength, hist=True)


def test_chroma(gpu_available):
    w, h = 150, 98
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_RCD, out_space=capi.CRAW_CS_ACES_AP1_LINEAR,
                 flags=capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE |
                 capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM)
    prof = scene.make_profile()
    bayer = scene.make_frame(w, h, seed=31, a=8e-4, b=3e-5)
    _check(run, bayer, frame_params(prof, denoise_chroma_strength=0.5), "chroma", hist=True)


@pytest.mark.parametrize("chroma", [True, False])
@pytest.mark.parametrize("size", [(150, 98), (520, 336), (610, 298), (96, 40), (166, 70), (64, 32)])
def test_rcd_interior_and_ring(gpu_available, size, chroma):
    """RCD develops the frame's interior with the streamed kernel and the
    ring of tiles around it with the tiled one, into the chroma stage or to
    the output (AWB samples and colour): interiors of several strips
    (partial strips at the right and bottom), last tiles narrower than the
    filter's reach (the interior then stops a tile earlier), and frames too
    small to have an interior."""
    w, h = size
    flags = capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM
    if chroma:
        flags |= capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_RCD, out_space=capi.CRAW_CS_ACES_AP1_LINEAR,
                 flags=flags)
    prof = scene.make_profile()
    bayer = scene.make_frame(w, h, seed=33 + w)
    _check(run, bayer, frame_params(prof, denoise_chroma_strength=0.7),
           "rcd %dx%d chroma %d" % (w, h, chroma), hist=True)


@pytest.mark.parametrize("size", [(520, 336), (64, 32), (5760, 3240)])
def test_rcd_colour_from_planes(gpu_available, size):
    """RCD to ACES AP1 without the chroma stage: the debayer writes its
    planes and the colour stage with the gamut compression runs from them
    (stage "color"); with the compression vetoed by the frame the colour
    stage stays fused into the debayer. Both bit-exact to the oracle, and
    a frame of each kind on one pipeline in turn."""
    w, h = size
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_RCD, out_space=capi.CRAW_CS_ACES_AP1_LINEAR,
                 flags=capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM | capi.CRAW_PIPELINE_FLAG_PROFILE)
    prof = scene.make_profile()
    bayer = scene.make_frame(w, h, seed=71)
    oracle.f32d.set_threads(16)
    for flags, stage in ((0, True), (capi.CRAW_FRAME_FLAG_NO_GAMUT_COMPRESSION, False), (0, True)):
        _check(run, bayer, frame_params(prof, flags=flags, exposure_stops=0.25),
               "rcd %dx%d frame flags %d" % (w, h, flags), hist=True)
        run.stream.synchronize()
        names = [n for n, _, _ in run.pipe.stage_times()]
        assert ("color" in names) == stage, names


@pytest.mark.parametrize("method", ["bilinear", "mhc"])
@pytest.mark.parametrize("chroma", [True, False])
@pytest.mark.parametrize("size", [(250, 98), (520, 336), (610, 298), (96, 40), (194, 50),
                                  (192, 48), (1000, 200)])
def test_simple_sizes(gpu_available, method, size, chroma):
    """Bilinear and MHC over 64x16 tiles, into the chroma stage or to the
    output (AWB samples and colour): partial last tiles across and down,
    a last tile narrower than the filter's reach, frames a whole number of
    tiles, and frames of a single row of tiles."""
    w, h = size
    flags = capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM
    if chroma:
        flags |= capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE
    debayer = capi.CRAW_DEBAYER_BILINEAR if method == "bilinear" else capi.CRAW_DEBAYER_MHC
    run = Runner(w, h, debayer=debayer, out_space=capi.CRAW_CS_BT2020_LINEAR, flags=flags)
    prof = scene.make_profile()
    bayer = scene.make_frame(w, h, seed=51 + w + h)
    _check(run, bayer, frame_params(prof, denoise_chroma_strength=0.6),
           "%s %dx%d chroma %d" % (method, w, h, chroma), hist=True)


def test_temporal_sequence(gpu_available):
    """Four frames of a moving clip: every frame's output equals the
    oracle run with the state threaded through it."""
    w, h = 96, 80
    flags = (capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE)
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_MHC, flags=flags)
    prof = scene.make_profile()
    state = None
    for i in range(4):
        bayer = scene.clip_frame(w, h, i, square=32, speed=4)
        fp = frame_params(prof, denoise_strength=1.0, denoise_temporal_strength=0.8,
                          noise_override=(4e-4, 1.6e-5))
        assert run.run(bayer, fp) == 0
        ref = oracle.f32d.process(bayer, run.pp, fp, prev=state, want_state=True)
        assert ref["ret"] == 0
        assert_bit_exact(run.output(), ref["out"], "temporal frame %d" % i)
        state = ref["state"]


@pytest.mark.parametrize("size", [(912, 272), (1000, 300), (458, 140), (896, 240), (40, 520)])
def test_temporal_history_bands(gpu_available, size):
    """The exact NLM with a temporal history over several bands: interior
    bands away from every border, partial last bands across and down,
    planes a whole number of bands, and a plane narrower than a band's
    reach. Three frames of a moving clip (the first without history), each
    equal to the oracle with the state threaded through."""
    w, h = size
    flags = (capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE)
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_BILINEAR, flags=flags)
    prof = scene.make_profile()
    state = None
    for i in range(3):
        bayer = scene.clip_frame(w, h, i, square=min(64, w // 2, h // 2), speed=6)
        fp = frame_params(prof, denoise_strength=0.8 + 0.2 * i, denoise_temporal_strength=0.7,
                          noise_override=(4e-4, 1.6e-5))
        assert run.run(bayer, fp) == 0
        ref = oracle.f32d.process(bayer, run.pp, fp, prev=state, want_state=True)
        assert ref["ret"] == 0
        assert_bit_exact(run.output(), ref["out"], "history %dx%d frame %d" % (w, h, i))
        state = ref["state"]


def test_full_path(gpu_available):
    """Config 2's stages at a small size: temporal first frame + NLM, RCD,
    chroma 0.5, AWB, ACEScg with gamut compression."""
    w, h = 256, 160
    flags = (capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE |
             capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM)
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_RCD, out_space=capi.CRAW_CS_ACES_AP1_LINEAR,
                 flags=flags)
    prof = scene.make_profile()
    bayer = scene.make_frame(w, h, seed=41)
    fp = frame_params(prof, denoise_strength=1.0, denoise_chroma_strength=0.5,
                      noise_override=(2e-4, 4e-6))
    _check(run, bayer, fp, "full path", hist=True)


@pytest.mark.parametrize("case", ["black_zero", "black_tiny", "gain_large"])
def test_profiles_either_side_of_the_short_division(gpu_available, case):
    """The streamed RCD divides through kdiv_rn only when the frame's black
    levels, ranges and gains keep its quotients in that sequence's range
    (black 0 or in [2^-16, 1], range in [2^-16, 2], gains in [2^-4, 2^4]);
    otherwise through the full division. Both ways the frame equals the
    oracle's: black 0 inside the range, a black level of 2^-20 and a red
    gain of 24 outside it."""
    w, h = 160, 96
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_RCD, out_space=capi.CRAW_CS_ACES_AP1_LINEAR,
                 flags=capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM)
    if case == "black_zero":
        prof = scene.make_profile(black=0.0)
    elif case == "black_tiny":
        prof = scene.make_profile(black=2.0 ** -20)
    else:
        prof = scene.make_profile(wb=(24.0, 1.0, 1.6))
    bayer = scene.make_frame(w, h, seed=23)
    _check(run, bayer, frame_params(prof), case, hist=True)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("layout", ["odd_pitch", "offset_base", "wide_pitch"])
def test_source_pitch_and_alignment(gpu_available, layout, method):
    """The streamed RCD and the bilinear / MHC tiles load two samples as one
    32-bit word where the mosaic's rows are 4-byte aligned, and as two
    16-bit loads where they are not: a row pitch of an odd number of
    samples, a base address 2 bytes past a 4-byte boundary, and an aligned
    padded pitch all develop the frame the oracle develops."""
    from libcraw_amd import hip
    from libcraw_amd.pipeline import image

    w, h = 200, 112
    pad, off = {"odd_pitch": (1, 0), "offset_base": (2, 2), "wide_pitch": (6, 0)}[layout]
    run = Runner(w, h, debayer=method, out_space=capi.CRAW_CS_ACES_AP1_LINEAR,
                 flags=capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM)
    prof = scene.make_profile()
    bayer = scene.make_frame(w, h, seed=29)
    fp = frame_params(prof)
    padded = np.zeros((h, w + pad), np.uint16)
    padded[:, :w] = bayer
    buf = hip.DeviceBuffer(padded.nbytes + 16)
    raw = np.zeros(padded.nbytes + 16, np.uint8)
    raw[off:off + padded.nbytes] = padded.view(np.uint8).ravel()
    buf.upload(raw)
    src = image(buf.value + off, w, h, capi.CRAW_FORMAT_R16_UNORM, pitch=(w + pad) * 2)
    assert run.pipe.record(run.stream, src, image(run.dst, w, h, run.fmt), fp) == 0
    run.stream.synchronize()
    ref = oracle.f32d.process(bayer, run.pp, fp, want_hist=True)
    assert ref["ret"] == 0
    assert_bit_exact(run.output(), ref["out"], layout)
    for name, a, b in zip(("count", "weight", "bright"), run.pipe.histogram(), ref["hist"]):
        assert np.array_equal(a, b), "%s: histogram %s differs" % (layout, name)
    buf.close()
