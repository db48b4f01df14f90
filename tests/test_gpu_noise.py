"""The noise fit with the frame measured on the device
(craw_noise_estimate_image) against the host measurement
(craw_noise_estimate, itself bit-exact against the reference's noise.c):
the same bins, counts and medians (the selection is exact), sums of means
exact on the device, so the model agrees to a few ulp. Frames of several
scenes, noise levels, sizes (whole and partial 64x60 tiles, the smallest
frame), the most bins, a flat region (a tile's samples all on one key) and
options, on one context whose scratch the calls share; error paths."""
import ctypes

import numpy as np
import pytest

from libcraw_amd import capi, hip, scene
from libcraw_amd.pipeline import image
from gpu_util import context

pytestmark = pytest.mark.gpu
P = ctypes.POINTER


def host_estimate(mosaic, prof, params):
    h, w = mosaic.shape
    m = np.ascontiguousarray(mosaic, np.uint16)
    out = capi.CRAWNoiseEstimate()
    ret = capi.lib().craw_noise_estimate(m.ctypes.data_as(P(ctypes.c_uint16)), w, h, w,
                                         ctypes.byref(prof), ctypes.byref(params),
                                         ctypes.byref(out))
    return ret, out


def device_estimate(mosaic, prof, params, pitch_pad=0):
    h, w = mosaic.shape
    pitch = w + pitch_pad
    padded = np.zeros((h, pitch), np.uint16)
    padded[:, :w] = mosaic
    buf = hip.DeviceBuffer(h * pitch * 2)
    buf.upload(padded)
    stream = hip.Stream()
    out = capi.CRAWNoiseEstimate()
    img = image(buf, w, h, capi.CRAW_FORMAT_R16_UNORM, pitch=pitch * 2)
    ret = capi.lib().craw_noise_estimate_image(context().handle, stream.handle, ctypes.byref(img),
                                               ctypes.byref(prof), ctypes.byref(params),
                                               ctypes.byref(out))
    buf.close()
    return ret, out


def params(**kw):
    p = capi.CRAWNoiseEstimateParams()
    capi.lib().craw_noise_estimate_params_init(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def ulps(a, b):
    ia = np.array([a], np.float32).view(np.int32)[0]
    ib = np.array([b], np.float32).view(np.int32)[0]
    return abs(int(ia) - int(ib))


@pytest.mark.parametrize("case", range(7))
def test_device_measurement_matches_host(gpu_available, case):
    w, h, a, b, kw, pad = [
        (1024, 768, 2e-4, 4e-6, {}, 0),
        (1000, 750, 8e-4, 3e-5, {}, 24),
        (2048, 1080, 4e-4, 1.6e-5, {"nb_bins": 256}, 0),
        (640, 480, 1e-4, 2e-6, {"nb_bins": 64, "min_bin_samples": 500, "max_level": 0.6}, 8),
        (96, 64, 4e-4, 1.6e-5, {"min_bin_samples": 50}, 0),
        (1536, 1024, 4e-4, 1.6e-5, {"nb_bins": 512, "min_bin_samples": 200}, 0),
        (1280, 720, 2e-4, 4e-6, {"min_bin_samples": 2000}, 0),
    ][case]
    prof = scene.make_profile()
    mosaic = scene.make_frame(w, h, seed=70 + case, a=a, b=b)
    if case == 6:
        mosaic[:, :512] = 1500     # flat: every residual 0, one key a tile
    pr = params(**kw)
    rh, eh = host_estimate(mosaic, prof, pr)
    rd, ed = device_estimate(mosaic, prof, pr, pitch_pad=pad)
    assert rh == rd, (rh, rd)
    if rh == 0:
        assert ed.nb_bins == eh.nb_bins and ed.nb_inliers == eh.nb_inliers
        for f in ("a", "b"):
            assert ulps(getattr(ed.model, f), getattr(eh.model, f)) <= 4, f
        assert ulps(ed.fit_rms, eh.fit_rms) <= 64
        assert ulps(ed.read_noise_lsb, eh.read_noise_lsb) <= 4


def test_device_errors(gpu_available):
    prof = scene.make_profile()
    m = scene.make_frame(64, 64, seed=1)
    assert device_estimate(m[:, :30], prof, params())[0] == -22
    assert device_estimate(m, prof, params(nb_bins=4))[0] == -22
    other = scene.make_profile()
    other.cfa = 1
    assert device_estimate(m, other, params())[0] == -95
    assert device_estimate(m, prof, params())[0] == -61      # too few samples a bin
