"""The oracle's highlight reconstruction (lib/ref/ref_highlight.c:91-335 of
the reference) against what the reference's tests and the arithmetic fix:

  - below the clip ramp every site is its balanced value stored as R16F:
    against numpy's float16 rounding (to nearest, ties to even), normal and
    subnormal halves and values below black included;
  - the reference's definition test on its highlight scene
    (tests/test_pipeline.c:4441-4524): unclipped sites kept within 2^-11,
    at least 200 sites of the G-only clip rebuilt within 2% of the scene's
    own G, the blown core's R/G and B/G within 3% of the blob's colour and
    its G within 3% of the lowest level that clears every clip, no core site
    below its clip, no ramped site lowered;
  - the three modes agree to the R16F rounding;
  - the chain: the rebuilt mosaic replaces the prep (its debayer runs on it
    unclipped) only with both the pipeline flag and the frame flag; a scene
    with nothing clipped develops as the plain chain up to the rounding;
  - argument checks.
"""
import ctypes

import numpy as np
import pytest

import oracle
import hl_scene as sc
from libcraw_amd import capi
from libcraw_amd.pipeline import frame_params, pipeline_params

MODES = [oracle.f32l, oracle.f32d, oracle.f64]
HL = capi.CRAW_PIPELINE_FLAG_HIGHLIGHT_RECONSTRUCTION
FHL = capi.CRAW_FRAME_FLAG_HIGHLIGHT_RECONSTRUCTION


def test_store_below_the_ramp():
    """u < 0.92 everywhere: a = 0, the site is u * wb rounded to half."""
    rng = np.random.default_rng(5)
    h, w = 64, 96
    m = rng.integers(0, 50000, (h, w)).astype(np.uint16)
    m[:8] = rng.integers(200, 300, (8, w))          # at and below black: tiny and negative
    prof = sc.identity_profile()
    for mode in (oracle.f32l, oracle.f32d):
        ret, out = mode.highlights(m, prof, sc.WB)
        assert ret == 0
        raw = m.astype(np.float32) * np.float32(1.0 / 65535.0)
        u = (raw - sc.BLACK) * (np.float32(1.0) / (sc.WHITE - sc.BLACK))
        assert (u < 0.92).all()
        want = (u * sc.WB[sc.channels(h, w)]).astype(np.float16).astype(np.float32)
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32))
        assert (np.abs(want[:8]) < 6.1e-5).any() and (want[:8] < 0).any()


@pytest.mark.parametrize("mode", MODES, ids=lambda m: m.prefix)
def test_reference_definition(mode):
    m = sc.mosaic(1.0)
    ret, out = mode.highlights(m, sc.identity_profile(), sc.WB)
    assert ret == 0
    out = out.astype(np.float64)
    u = sc.u_of(m)
    chan = sc.channels()
    xb = u.astype(np.float64) * sc.WB[chan]
    truth = sc.truth()
    keep = u < 0.92
    worst_keep = (np.abs(out - xb) / np.maximum(xb, 1e-3))[keep].max()
    lowered = ((u >= 0.92) & (out < xb * (1.0 - 1.0 / 1024.0))).sum()

    h, w = m.shape
    pad = np.pad(u, 1, constant_values=np.nan)
    pch = np.pad(chan, 1, constant_values=-1)

    def around(mask_ch, lo, hi):
        ok = np.ones((h, w), bool)
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                uu = pad[1 + dy:1 + dy + h, 1 + dx:1 + dx + w]
                cc = pch[1 + dy:1 + dy + h, 1 + dx:1 + dx + w]
                sel = np.isin(cc, mask_ch)
                ok &= ~sel | ((uu >= lo) & (uu < hi))
        return ok

    x = np.arange(w)[None, :].repeat(h, 0)
    part = (chan == 1) & (x < 128) & (u >= 0.98) & around([0, 2], 0.0, 0.92)
    worst_part = (np.abs(out - truth[..., 1]) / truth[..., 1])[part].max()
    core = (x >= 128) & around([0, 1, 2], 0.999, 2.0)
    r, g, b = (out[core & (chan == c)].mean() for c in range(3))
    want = max(sc.WB[0] / sc.FULL[0], 1.0, sc.WB[2] / sc.FULL[2])
    core_min = (out / sc.WB[chan])[core].min()

    assert worst_keep <= 1.0 / 2048.0, worst_keep
    assert lowered == 0
    assert part.sum() >= 200 and worst_part <= 0.02, (part.sum(), worst_part)
    assert (core & (chan == 1)).sum() >= 100
    assert abs(r / g / sc.FULL[0] - 1.0) <= 0.03 and abs(b / g / sc.FULL[2] - 1.0) <= 0.03
    assert abs(g / want - 1.0) <= 0.03 and core_min >= 0.999, (g, want, core_min)


def test_modes_agree():
    m = sc.mosaic(1.0)
    prof = sc.identity_profile()
    outs = [mode.highlights(m, prof, sc.WB)[1].astype(np.float64) for mode in MODES]
    scale = np.maximum(np.abs(outs[2]), 1e-3)
    assert (np.abs(outs[0] - outs[1]) / scale).max() <= 2.0 ** -10
    assert (np.abs(outs[1] - outs[2]) / scale).max() <= 2.0 ** -10


def test_chain():
    m = sc.mosaic(1.0)
    prof = sc.identity_profile()
    mode = oracle.f32d
    plain = mode.process(m, pipeline_params(sc.W, sc.H, debayer=capi.CRAW_DEBAYER_MHC),
                         frame_params(prof))
    for flags, fflags in ((HL, 0), (0, FHL)):
        r = mode.process(m, pipeline_params(sc.W, sc.H, debayer=capi.CRAW_DEBAYER_MHC,
                                            flags=flags), frame_params(prof, flags=fflags))
        assert r["ret"] == 0 and np.array_equal(r["out"], plain["out"])
    pp = pipeline_params(sc.W, sc.H, debayer=capi.CRAW_DEBAYER_MHC, flags=HL)
    r = mode.process(m, pp, frame_params(prof, flags=FHL), want_inter=True)
    assert r["ret"] == 0
    ret, v = mode.highlights(m, prof, sc.WB)
    ret, want = mode.debayer_mosaic(v, capi.CRAW_DEBAYER_MHC)
    assert np.array_equal(r["inter"], want)
    assert r["inter"][..., 1].max() > 1.5       # unclipped: the rebuilt core is above 1
    # nothing clipped: the plain chain up to the R16F rounding
    dim = sc.mosaic(0.2)
    a = mode.process(dim, pp, frame_params(prof))["out"]
    b = mode.process(dim, pp, frame_params(prof, flags=FHL))["out"]
    assert np.abs(a - b).max() <= 2e-3


def test_errors():
    m = sc.mosaic(1.0)
    prof = sc.identity_profile()
    mode = oracle.f32d
    assert mode.highlights(m[:6, :6], prof, sc.WB)[0] == -22
    assert mode.highlights(m[:, :31], prof, sc.WB)[0] == -22
    other = sc.identity_profile()
    other.cfa = capi.CRAW_CFA_RGGB + 1
    assert mode.highlights(m, other, sc.WB)[0] == -95
    bad = sc.identity_profile()
    bad.white_level[2] = bad.black_level[2]
    assert mode.highlights(m, bad, sc.WB)[0] == -22
