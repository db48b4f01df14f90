"""The oracle's prep and debayer against the known answers of the
reference's tests/test_ref_debayer.c, in all three oracle modes, plus an
independent brute-force restatement of bilinear and MHC as convolutions
with mirrored borders (float64) on small random mosaics."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle
from libcraw_amd import capi
from libcraw_amd.pipeline import frame_params, pipeline_params

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_vectors.json")))
MODES = [oracle.f32l, oracle.f32d, oracle.f64]
METHODS = [capi.CRAW_DEBAYER_BILINEAR, capi.CRAW_DEBAYER_MHC, capi.CRAW_DEBAYER_RCD]


def unit_profile():
    p = capi.CRAWProfile()
    capi.lib().craw_profile_init(ctypes.byref(p))
    return p


def sites(h, w, v):
    m = np.empty((h, w), np.uint16)
    m[0::2, 0::2], m[0::2, 1::2], m[1::2, 0::2], m[1::2, 1::2] = v
    return m


def finite_alpha(rgb):
    return np.isfinite(rgb).all() and np.all(rgb[..., 3] == 1.0)


@pytest.mark.parametrize("mode", MODES, ids=lambda m: m.prefix)
@pytest.mark.parametrize("method", METHODS)
def test_constant_colour(mode, method):
    g = GOLD["debayer_constant"]
    bayer = sites(g["h"], g["w"], g["site"])
    ret, rgb = mode.debayer_wb(bayer, unit_profile(), [1, 1, 1], method)
    assert ret == 0 and finite_alpha(rgb)
    want = np.array([g["site"][0], g["site"][1], g["site"][3]]) / 65535.0
    assert np.abs(rgb[..., :3] - want).max() <= g["tol"]


@pytest.mark.parametrize("mode", MODES, ids=lambda m: m.prefix)
@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("vertical", [False, True])
def test_ramps(mode, method, vertical):
    g = GOLD["debayer_ramp"]
    w, h = g["w"], g["h"]
    yy, xx = np.mgrid[0:h, 0:w]
    bayer = (g["base"] + g["step"] * (yy if vertical else xx)).astype(np.uint16)
    ret, rgb = mode.debayer_wb(bayer, unit_profile(), [1, 1, 1], method)
    assert ret == 0 and finite_alpha(rgb)
    mg = g["margin"][method]
    want = bayer.astype(np.float64) / 65535.0
    err = np.abs(rgb[mg:h - mg, mg:w - mg, :3] - want[mg:h - mg, mg:w - mg, None]).max()
    assert err <= g["tol"][method]


@pytest.mark.parametrize("mode", MODES, ids=lambda m: m.prefix)
def test_directional_step(mode):
    g = GOLD["directional"]
    w, h = g["w"], g["h"]
    bayer = np.where(np.arange(w)[None, :] < w // 2, g["lo"], g["hi"]).astype(np.uint16)
    bayer = np.repeat(bayer, h, axis=0)
    err = {}
    for m in METHODS:
        ret, rgb = mode.debayer_wb(bayer, unit_profile(), [1, 1, 1], m)
        assert ret == 0 and finite_alpha(rgb)
        want = bayer.astype(np.float64) / 65535.0
        err[m] = np.abs(rgb[4:h - 4, 4:w - 4, :3] - want[4:h - 4, 4:w - 4, None]).sum()
    assert err[capi.CRAW_DEBAYER_RCD] < 0.01 * err[capi.CRAW_DEBAYER_BILINEAR]
    assert err[capi.CRAW_DEBAYER_RCD] < 0.01 * err[capi.CRAW_DEBAYER_MHC]


@pytest.mark.parametrize("mode", MODES, ids=lambda m: m.prefix)
@pytest.mark.parametrize("method", METHODS)
def test_hostile_inputs_stay_finite(mode, method):
    w, h = 40, 24
    yy, xx = np.mgrid[0:h, 0:w]
    checker = np.where(((xx >> 1) + (yy >> 1)) & 1, 65535, 0).astype(np.uint16)
    seed = 0x12345678
    noise = np.empty(w * h, np.uint16)
    for i in range(w * h):
        seed ^= (seed << 13) & 0xFFFFFFFF
        seed ^= seed >> 17
        seed ^= (seed << 5) & 0xFFFFFFFF
        noise[i] = seed & 0xFFFF
    for b in (checker, noise.reshape(h, w), np.zeros((h, w), np.uint16)):
        ret, rgb = mode.debayer_wb(b, unit_profile(), [1, 1, 1], method)
        assert ret == 0 and finite_alpha(rgb)


@pytest.mark.parametrize("mode", MODES, ids=lambda m: m.prefix)
@pytest.mark.parametrize("method", METHODS)
def test_tiny_frame(mode, method):
    g = GOLD["debayer_tiny"]
    bayer = sites(4, 4, g["site"])
    ret, rgb = mode.debayer_wb(bayer, unit_profile(), [1, 1, 1], method)
    assert ret == 0 and finite_alpha(rgb)
    want = np.array([g["site"][0], g["site"][1], g["site"][3]]) / 65535.0
    assert np.abs(rgb[..., :3] - want).max() <= g["tol"]


@pytest.mark.parametrize("mode", MODES, ids=lambda m: m.prefix)
def test_prep_arithmetic(mode):
    g = GOLD["prep"]
    p = unit_profile()
    black, white = np.float32(g["black"] / 65535.0), np.float32(g["white"] / 65535.0)
    for s in range(4):
        p.black_level[s], p.white_level[s] = black, white
    wb = np.array(g["wb"], np.float32)
    for c in range(3):
        p.as_shot_wb[c] = wb[c]
    bayer = sites(8, 8, g["site"])
    ret, rgb = mode.debayer_wb(bayer, p, wb, capi.CRAW_DEBAYER_BILINEAR)
    assert ret == 0
    f = np.float32
    lo = (f(100) / f(65535) - black) / (white - black) * wb[0]
    mid = (f(32000) / f(65535) - black) / (white - black)
    hi = (f(60000) / f(65535) - black) / (white - black) * wb[2]
    assert lo < 0 and hi > 1
    assert np.abs(rgb[..., 0] - lo).max() <= 1e-7      # below black is kept
    assert np.abs(rgb[..., 1] - mid).max() <= 1e-6
    assert np.abs(rgb[..., 2] - 1.0).max() <= 1e-7     # clipped after the gain


@pytest.mark.parametrize("mode", MODES, ids=lambda m: m.prefix)
def test_error_codes(mode):
    b = np.zeros((8, 8), np.uint16)
    p = unit_profile()
    p.cfa = 1
    assert mode.debayer_wb(b, p, [1, 1, 1], capi.CRAW_DEBAYER_RCD)[0] == -capi.ENOTSUP
    p.cfa = 0
    assert mode.debayer_wb(b, p, [1, 1, 1], 3)[0] == -capi.EINVAL
    assert mode.debayer_wb(np.zeros((8, 7), np.uint16), p, [1, 1, 1], 2)[0] == -capi.EINVAL
    assert mode.debayer_wb(b, p, [1, 1, 1], 2, stride=4)[0] == -capi.EINVAL


@pytest.mark.parametrize("mode", MODES, ids=lambda m: m.prefix)
def test_process_neutral(mode):
    """tests/test_ref_debayer.c:345-413: a neutral stays neutral through the
    chain at the luminance the gain says; an extra stop doubles it."""
    g = GOLD["process_neutral"]
    w, h = g["w"], g["h"]
    p = unit_profile()
    p.as_shot_wb[0], p.as_shot_wb[2] = g["wb"][0], g["wb"][2]
    cam = np.array([[0.9505, 0.9505, 0.9505], [1.0, 1.0, 1.0], [1.0890, 1.0890, 1.0890]]) / 3
    for r in range(3):
        for c in range(3):
            p.as_shot_matrix[r][c] = np.float32(cam[r][c])
    p.gain = g["gain"]
    chan = np.array([2.0, 1.0, 1.0, 1.6], np.float32)
    s = (np.arange(h)[:, None] & 1) * 2 + (np.arange(w)[None, :] & 1)
    bayer = (np.float32(0.25) * np.float32(65535.0) / chan[s] + np.float32(0.5)).astype(np.uint16)
    pp = pipeline_params(w, h, debayer=capi.CRAW_DEBAYER_MHC, out_space=capi.CRAW_CS_BT709_LINEAR)
    for stops, want, tol in ((0.0, 0.5, g["tol"][0]), (1.0, 1.0, g["tol"][1])):
        fp = frame_params(p, exposure_stops=stops)
        r = mode.process(bayer, pp, fp)
        assert r["ret"] == 0
        assert np.abs(r["out"][..., :3] - want).max() <= tol
        assert np.all(r["out"][..., 3] == 1.0)


def _conv_mirror(v, k):
    """Brute force: sum_k k[dy,dx] * v[mirror(y+dy), mirror(x+dx)], float64."""
    h, w = v.shape
    r = k.shape[0] // 2

    def mir(i, n):
        while i < 0 or i >= n:
            i = -i if i < 0 else 2 * n - 2 - i
        return i
    out = np.zeros((h, w))
    for y in range(h):
        for x in range(w):
            acc = 0.0
            for dy in range(-r, r + 1):
                for dx in range(-r, r + 1):
                    if k[dy + r, dx + r] != 0:
                        acc += k[dy + r, dx + r] * v[mir(y + dy, h), mir(x + dx, w)]
            out[y, x] = acc
    return out


def test_bilinear_and_mhc_against_brute_force_convolutions():
    """Independent restatement of the reference's bilinear and MHC filters
    (lib/ref/ref_debayer.c:65-176) as kernels applied by site,
    float64, on a random prepped mosaic: bilinear and MHC of every oracle
    mode within float rounding (1e-6)."""
    rng = np.random.default_rng(4)
    h, w = 14, 18
    v = rng.uniform(-0.05, 1.0, size=(h, w))
    z = np.zeros((5, 5))
    K = {}
    e = z.copy(); e[1, 2] = e[3, 2] = e[2, 1] = e[2, 3] = 0.25; K["edge"] = e
    c = z.copy(); c[1, 1] = c[1, 3] = c[3, 1] = c[3, 3] = 0.25; K["corner"] = c
    hz = z.copy(); hz[2, 1] = hz[2, 3] = 0.5; K["horiz"] = hz
    vt = z.copy(); vt[1, 2] = vt[3, 2] = 0.5; K["vert"] = vt
    ident = z.copy(); ident[2, 2] = 1.0; K["id"] = ident
    # Malvar-He-Cutler 2004, Fig. 2, over 8
    g = z.copy(); g[2, 2] = 4; g[1, 2] = g[3, 2] = g[2, 1] = g[2, 3] = 2
    g[0, 2] = g[4, 2] = g[2, 0] = g[2, 4] = -1; K["g"] = g / 8
    ch = z.copy(); ch[2, 2] = 5; ch[2, 1] = ch[2, 3] = 4
    ch[1, 1] = ch[1, 3] = ch[3, 1] = ch[3, 3] = -1; ch[2, 0] = ch[2, 4] = -1
    ch[0, 2] = ch[4, 2] = 0.5; K["ch"] = ch / 8
    K["cv"] = K["ch"].T.copy()
    d = z.copy(); d[2, 2] = 6; d[1, 1] = d[1, 3] = d[3, 1] = d[3, 3] = 2
    d[0, 2] = d[4, 2] = d[2, 0] = d[2, 4] = -1.5; K["d"] = d / 8
    conv = {k: _conv_mirror(v, m) for k, m in K.items()}
    s = (np.arange(h)[:, None] & 1) * 2 + (np.arange(w)[None, :] & 1)
    pick = lambda table: np.stack([np.choose(s, [conv[t[i]] for t in table]) for i in range(3)],
                                  -1)
    want_bil = pick([("id", "edge", "corner"), ("horiz", "id", "vert"),
                     ("vert", "id", "horiz"), ("corner", "edge", "id")])
    want_mhc = pick([("id", "g", "d"), ("ch", "id", "cv"), ("cv", "id", "ch"),
                     ("d", "g", "id")])
    for mode in MODES:
        for m, want in ((capi.CRAW_DEBAYER_BILINEAR, want_bil), (capi.CRAW_DEBAYER_MHC, want_mhc)):
            ret, rgb = mode.debayer_mosaic(v.astype(mode.dtype), m)
            assert ret == 0
            assert np.abs(rgb[..., :3] - want).max() < 1e-6, (mode.prefix, m)


def test_float_modes_agree_on_debayer():
    """Debayer and colour use no transcendental outside the gamut
    compression: the two float modes must agree bit for bit there."""
    from libcraw_amd import scene
    bayer = scene.make_frame(64, 48, seed=9)
    prof = scene.make_profile()
    for m in METHODS:
        pp = pipeline_params(64, 48, debayer=m, out_space=capi.CRAW_CS_BT709_LINEAR)
        a = oracle.f32l.process(bayer, pp, frame_params(prof))["out"]
        b = oracle.f32d.process(bayer, pp, frame_params(prof))["out"]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
