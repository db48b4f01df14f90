"""The oracle's raw-domain denoise, temporal average and chroma denoise
against what the mathematics and the reference's tests fix:

  - a constant mosaic is a fixed point of the NLM, and strength -> 0 is the
    identity (igat(gat(u)) = u through the 16-bit re-encode);
  - flat-field noise falls to at most 60% and a checkerboard keeps at least
    70% of its contrast (tests/test_pipeline.c:699-730 of the reference);
  - on a static scene the temporal state's variance converges to
    (1 - a) / (1 + a), the fixed point of the state recursion of
    lib/ref/ref_denoise.c:82-105, and the noise falls every frame;
  - the temporal pass without history is the spatial filter exactly;
  - chroma denoise leaves G bit for bit, keeps a flat colour and reduces
    chroma blotches (tests/test_pipeline.c:4256-4312);
  - the reference's argument checks and error codes;
  - the float/glibc mode, the deterministic mode and the double mode agree
    to a stated number of 16-bit codes.
"""
import ctypes

import numpy as np
import pytest

import oracle
from libcraw_amd import capi, scene
from libcraw_amd.pipeline import frame_params, pipeline_params

MODES = [oracle.f32l, oracle.f32d, oracle.f64]
A, B = 2e-4, 4e-6


def prof12():
    return scene.make_profile()


def flat_mosaic(h, w, level, a=A, b=B, seed=1):
    rgb = np.full((h, w, 3), level)
    return scene.mosaic_from_scene(rgb, wb=(1.0, 1.0, 1.0), a=a, b=b, seed=seed)


@pytest.mark.parametrize("mode", MODES, ids=lambda m: m.prefix)
def test_constant_is_fixed_point(mode):
    bayer = np.empty((32, 40), np.uint16)
    bayer[0::2, 0::2], bayer[0::2, 1::2], bayer[1::2, 0::2], bayer[1::2, 1::2] = (
        900, 2100, 2100, 3000)
    ret, out, _, _ = mode.denoise_temporal(bayer, prof12(), A, B, 1.0)
    assert ret == 0
    assert np.array_equal(out, bayer)


@pytest.mark.parametrize("mode", MODES, ids=lambda m: m.prefix)
def test_vanishing_strength_is_identity(mode):
    bayer = scene.make_frame(48, 40, seed=2)
    ret, out, _, _ = mode.denoise_temporal(bayer, prof12(), A, B, 1e-4)
    assert ret == 0
    assert np.array_equal(out, bayer)


@pytest.mark.parametrize("mode", [oracle.f32l, oracle.f64], ids=lambda m: m.prefix)
def test_flat_noise_and_checkerboard(mode):
    h, w = 96, 96
    flat = flat_mosaic(h, w, 0.3)
    ret, out, _, _ = mode.denoise_temporal(flat, prof12(), A, B, 1.0)
    assert ret == 0
    inner = (slice(8, -8), slice(8, -8))
    assert out[inner].astype(float).std() <= 0.6 * flat[inner].astype(float).std()

    yy, xx = np.mgrid[0:h, 0:w]
    chk = np.where(((xx >> 2) + (yy >> 2)) & 1, 0.6, 0.2)
    rgb = np.repeat(chk[..., None], 3, axis=2)
    bayer = scene.mosaic_from_scene(rgb, wb=(1.0, 1.0, 1.0), a=A, b=B, seed=3)
    ret, out, _, _ = mode.denoise_temporal(bayer, prof12(), A, B, 1.0)
    hi = chk[inner] > 0.4
    c_in = bayer[inner][hi].mean() - bayer[inner][~hi].mean()
    c_out = out[inner][hi].mean() - out[inner][~hi].mean()
    assert c_out >= 0.7 * c_in


def test_static_scene_temporal_convergence():
    """Variance of the state on a static noise-only scene settles at
    (1 - a)/(1 + a) = 0.111 at alpha_max 0.8 (the fixed point of the
    recursion v = (1 - a)^2 + a^2 v of lib/ref/ref_denoise.c:82-105) where the
    full weight is taken, and the
    output noise falls over the first frames."""
    h, w = 64, 64
    p = prof12()
    prev = None
    stds = []
    for i in range(14):
        f = flat_mosaic(h, w, 0.3, seed=100 + i)
        ret, out, nxt, _ = oracle.f32l.denoise_temporal(f, p, A, B, 0.0, alpha_max=0.8,
                                                        prev=prev, want_state=True)
        assert ret == 0
        stds.append(out[8:-8, 8:-8].astype(float).std())
        prev = nxt
    var = prev[1][8:-8, 8:-8]
    # (1 - a)/(1 + a) is the floor, reached where alpha stays at alpha_max;
    # samples whose 3x3 difference fluctuates above 2 sigma^2 take less.
    assert (1 - 0.8) / (1 + 0.8) - 0.01 <= var.mean() <= 0.2
    assert abs(np.median(var) - (1 - 0.8) / (1 + 0.8)) < 0.03
    assert stds[-1] < 0.6 * stds[0]
    assert all(stds[i + 1] <= stds[i] * 1.05 for i in range(5))


def test_temporal_first_frame_is_spatial_alone():
    bayer = scene.make_frame(64, 48, seed=4)
    p = prof12()
    _, a, _, _ = oracle.f32d.denoise_temporal(bayer, p, A, B, 1.0)
    ret, b, nxt, _ = oracle.f32d.denoise_temporal(bayer, p, A, B, 1.0, alpha_max=0.8,
                                                   want_state=True)
    assert ret == 0 and np.array_equal(a, b)
    assert np.all(nxt[1] == 1.0)


def test_denoise_error_codes():
    p = prof12()
    b = scene.make_frame(32, 32, seed=1)
    f = oracle.f32l.denoise_temporal
    assert f(b, p, A, B, 0.0)[0] == -capi.EINVAL                    # neither filter
    assert f(b, p, 0.0, B, 1.0)[0] == -capi.EINVAL                  # a <= 0
    assert f(b, p, A, B, 1.0, alpha_max=1.5)[0] == -capi.EINVAL
    assert f(scene.make_frame(8, 8), p, A, B, 1.0)[0] == -capi.EINVAL   # below 2*(S+P+2)
    p.cfa = 2
    assert f(b, p, A, B, 1.0)[0] == -capi.ENOTSUP


def test_float_modes_and_fp64_error_statement():
    """Where glibc's expf rounds differently from the correctly rounded one
    a weight moves by an ulp; the 16-bit output of the two float modes then
    differs in a handful of samples by one code. Against the FP64 rerun the
    FP32 output is within one code everywhere (the re-encode's own
    rounding), and its pre-rounding level within 2^-11 of clip."""
    bayer = scene.make_frame(192, 160, seed=6)
    p = prof12()
    _, ol, _, _ = oracle.f32l.denoise_temporal(bayer, p, A, B, 1.0)
    _, od, _, lvl = oracle.f32d.denoise_temporal(bayer, p, A, B, 1.0, want_level=True)
    _, o64, _, lvl64 = oracle.f64.denoise_temporal(bayer, p, A, B, 1.0, want_level=True)
    d = np.abs(ol.astype(int) - od.astype(int))
    assert d.max() <= 1 and (d > 0).mean() < 1e-3
    assert np.abs(od.astype(int) - o64.astype(int)).max() <= 1
    clip = 65535.0 * (4095 - 512) / 65535.0    # the clip level in codes above black
    assert np.abs(lvl.astype(np.float64) - lvl64).max() / clip <= 2.0 ** -11


def _band_noise(rgb, y0, y1, x0, x1):
    """Spread of the 4x4 block means of R - G and B - G: the band the
    chroma stage works on."""
    d = np.stack([rgb[..., 0] - rgb[..., 1], rgb[..., 2] - rgb[..., 1]], -1)[y0:y1, x0:x1]
    h, w = d.shape[0] // 4, d.shape[1] // 4
    m = d[:4 * h, :4 * w].reshape(h, 4, w, 4, 2).mean(axis=(1, 3))
    return float(np.sqrt(((m - m.mean(axis=(0, 1))) ** 2).mean()))


@pytest.mark.parametrize("mode", MODES, ids=lambda m: m.prefix)
def test_chroma_invariants(mode):
    """tests/test_pipeline.c:4256-4312 of the reference: +-0.02 chroma
    blotches per 8x8 block over a neutral patch lose at least 40% of their
    spread in the quarter-resolution band, a flat coloured patch keeps its
    colour within 5% at its centre, G and alpha are untouched bit for bit."""
    rng = np.random.default_rng(5)
    h, w = 96, 128
    rgb = np.ones((h, w, 4))
    rgb[..., :3] = 0.45
    rgb[:40, :40, :3] = [0.75, 0.30, 0.12]                      # warm flat patch
    blot = rng.integers(-20, 21, size=(h // 8, w // 8, 2)).repeat(8, 0).repeat(8, 1) * 0.001
    rgb[48:, 48:, 0] += blot[48:, 48:, 0]
    rgb[48:, 48:, 2] += blot[48:, 48:, 1]
    a = np.ascontiguousarray(rgb, mode.dtype)
    ret, out = mode.chroma_denoise(rgb, 1.0)
    assert ret == 0
    assert np.array_equal(out[..., 1], a[..., 1]) and np.array_equal(out[..., 3], a[..., 3])
    before = _band_noise(a, 56, h - 8, 56, w - 8)
    after = _band_noise(out, 56, h - 8, 56, w - 8)
    assert after <= 0.6 * before
    for c in (0, 2):
        d0 = (a[16:24, 16:24, c] - a[16:24, 16:24, 1]).mean()
        d1 = (out[16:24, 16:24, c] - out[16:24, 16:24, 1]).mean()
        assert abs(d1 - d0) <= 0.05 * abs(d0)
    flat = np.ones((h, w, 4)) * np.array([0.5, 0.25, 0.1, 1.0])
    ret, out = mode.chroma_denoise(flat, 0.5)
    assert np.abs(out - np.ascontiguousarray(flat, mode.dtype)).max() < 1e-6
    assert mode.chroma_denoise(np.ones((16, 40, 4)), 0.5)[0] == -capi.EINVAL
    assert mode.chroma_denoise(flat, 1.5)[0] == -capi.EINVAL


def test_full_path_modes_agree():
    """The whole path of the benchmark's workload at a small size: the two
    float modes within a few 16-bit re-encode flips (one code each), the
    FP32 output against FP64 within 2^-11 of the clip level."""
    w, h = 128, 96
    bayer = scene.make_frame(w, h, seed=8)
    pp = pipeline_params(w, h, debayer=capi.CRAW_DEBAYER_RCD,
                         out_space=capi.CRAW_CS_ACES_AP1_LINEAR,
                         flags=capi.CRAW_PIPELINE_FLAG_DENOISE |
                         capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE)
    fp = frame_params(scene.make_profile(), denoise_strength=1.0, denoise_chroma_strength=0.5,
                      noise_override=(A, B))
    rl = oracle.f32l.process(bayer, pp, fp, want_clean=True)
    rd = oracle.f32d.process(bayer, pp, fp, want_clean=True)
    r64 = oracle.f64.process(bayer, pp, fp, want_clean=True)
    assert rl["ret"] == rd["ret"] == r64["ret"] == 0
    assert (rl["clean"] != rd["clean"]).sum() <= 4
    same = rl["clean"] == rd["clean"]
    assert same.mean() > 0.999
    # where the denoised mosaics agree the rest of the path is float math
    # and must agree closely
    assert np.abs(rd["out"] - r64["out"]).max() < 2.0 ** -11 * 4
