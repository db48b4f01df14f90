"""The AWB loop closed on the GPU: each frame's histogram slot, read back
through craw_pipeline_get_awb_histogram, goes to craw_awb_estimate and the
smoother, and the next frame is recorded with CRAW_WB_REBALANCE at the
illuminant they give, as cineraw does (tools/cineraw.c:1336-1380 of
cyanreg/libcraw: the track seeded at the as-shot gains, the first reading a
cut). Every frame's output and histogram equal the oracle's development of
the same mosaic at the same illuminant, bit for bit, and the track settles
on the scene's illuminant.
"""
import ctypes

import numpy as np
import pytest

import oracle
from libcraw_amd import capi, scene
from libcraw_amd.pipeline import frame_params

from gpu_util import Runner, assert_bit_exact

pytestmark = pytest.mark.gpu

KNOTS = [(2000.0, (0.8454, 1.0, 3.9482)), (3200.0, (1.3763, 1.0, 2.4707)),
         (5500.0, (2.02539, 1.0, 1.66797)), (7500.0, (2.2, 1.0, 1.45))]


def tabled_profile():
    p = scene.make_profile(wb=KNOTS[2][1])
    for i, (cct, wb) in enumerate(KNOTS):
        p.wb_knots[i].cct = cct
        for c in range(3):
            p.wb_knots[i].wb[c] = wb[c]
    p.nb_wb_knots = len(KNOTS)
    # two matrix knots: REBALANCE interpolates between them in mired order
    for i, (cct, scale) in enumerate(((2856.0, 0.97), (6504.0, 1.0))):
        p.matrix_knots[i].cct = cct
        for r in range(3):
            for c in range(3):
                p.matrix_knots[i].m[r][c] = float(scene.CAMERA_MATRIX[r][c]) * (
                    scale if r == c else 1.0)
    p.nb_matrix_knots = 2
    return p


def grey_clip(p, cct, tint, w, h, n, seed=5):
    """Grey patches of several levels lit by (cct, tint) plus a coloured
    square moving over them, mosaiced at the as-shot gains, fresh noise each
    frame."""
    lib = capi.lib()
    rg, bg = ctypes.c_float(), ctypes.c_float()
    assert lib.craw_profile_illuminant_to_neutral(ctypes.byref(p),
                                                  ctypes.byref(capi.CRAWIlluminant(cct, tint)),
                                                  ctypes.byref(rg), ctypes.byref(bg)) == 0
    rng = np.random.Generator(np.random.Philox(seed))
    g = np.repeat(np.repeat(rng.uniform(0.05, 0.45, size=(h // 16, w // 16)), 16, 0), 16, 1)
    wb = np.array(p.as_shot_wb[:])
    base = np.stack([g * rg.value, g, g * bg.value], axis=-1) * wb
    out = []
    for i in range(n):
        f = base.copy()
        x0 = 8 + 12 * i
        f[40:72, x0:x0 + 32] = np.array([0.5, 0.2, 0.1]) * wb
        out.append(scene.mosaic_from_scene(f, wb=tuple(wb), a=2e-4, b=4e-6, seed=seed * 31 + i))
    return out


@pytest.mark.parametrize("cct, tint", [(3400.0, 0.02), (6200.0, -0.03)])
def test_awb_loop(gpu_available, cct, tint):
    lib = capi.lib()
    w, h = 256, 192
    p = tabled_profile()
    flags = (capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM | capi.CRAW_PIPELINE_FLAG_DENOISE |
             capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE)
    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_RCD, out_space=capi.CRAW_CS_ACES_AP1_LINEAR,
                 flags=flags, nb_awb_slots=3)
    sp = capi.CRAWAwbSmootherParams()
    lib.craw_awb_smoother_params_init(ctypes.byref(sp))
    sm = ctypes.c_void_p()
    assert lib.craw_awb_smoother_create(ctypes.byref(sp), ctypes.byref(sm)) == 0
    seed = capi.CRAWIlluminant()
    assert lib.craw_profile_neutral_to_illuminant(ctypes.byref(p), 1.0 / p.as_shot_wb[0],
                                                  1.0 / p.as_shot_wb[2],
                                                  ctypes.byref(seed)) == 0
    lib.craw_awb_smoother_set(sm, 0.0, ctypes.byref(seed))
    illum = None
    track = []
    try:
        for i, bayer in enumerate(grey_clip(p, cct, tint, w, h, 6)):
            kw = dict(denoise_strength=1.0, denoise_chroma_strength=0.5,
                      noise_override=(2e-4, 4e-6))
            if illum is None:
                fp = frame_params(p, **kw)
            else:
                fp = frame_params(p, wb_mode=capi.CRAW_WB_REBALANCE,
                                  illuminant=(illum.cct, illum.tint), **kw)
            assert run.run(bayer, fp) == 0
            ref = oracle.f32d.process(bayer, run.pp, fp, want_hist=True)
            assert ref["ret"] == 0
            assert_bit_exact(run.output(), ref["out"], "frame %d" % i)
            # the slot read back through the C API, as the caller reads it
            hs = capi.CRAWAwbHistogram()
            assert lib.craw_pipeline_get_awb_histogram(run.pipe.handle, ctypes.byref(hs)) == 0
            n = hs.bins * hs.bins
            got = [np.ctypeslib.as_array(a, shape=(n,)).copy()
                   for a in (hs.count, hs.weight, hs.weight_bright)]
            for name, a, b in zip(("count", "weight", "bright"), got, ref["hist"]):
                assert np.array_equal(a, b), "frame %d histogram %s" % (i, name)
            est = capi.CRAWAwbEstimate()
            assert lib.craw_awb_estimate(ctypes.byref(p), ctypes.byref(hs), None,
                                         ctypes.byref(est)) == 0
            assert est.valid
            nxt = capi.CRAWIlluminant()
            ret = lib.craw_awb_smoother_push(sm, i / 24.0, ctypes.byref(est), i == 0,
                                             ctypes.byref(nxt))
            assert ret == 0
            illum = nxt
            track.append((est.illuminant.cct, nxt.cct, nxt.tint))
    finally:
        lib.craw_awb_smoother_free(ctypes.byref(sm))
    # the estimates read the scene's light, the track holds it
    assert all(abs(1e6 / e - 1e6 / cct) < 10.0 for e, _, _ in track), track
    assert abs(1e6 / track[-1][1] - 1e6 / cct) < 10.0 and abs(track[-1][2] - tint) < 0.03
