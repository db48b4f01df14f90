"""The noise-model fit of a frame (craw_noise_estimate).

Pins: the reference's own lib/noise.c built from its source under
oracle/_ref, every output field bit for bit on frames of several scenes,
noise levels and options (skipped where the reference tree is absent); and
on their own, a smooth field carrying Gaussian noise of a known model,
which the fit recovers, and the error paths."""
import ctypes

import numpy as np
import pytest

import oracle
from libcraw_amd import capi, scene

REF = oracle.LazyRef()
HAVE_REF = oracle.ref_available()
needs_ref = pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (no reference tree)")
P = ctypes.POINTER
EINVAL, ENOTSUP, ENODATA = 22, 95, 61


def bind_ref():
    f = REF.craw_noise_estimate
    f.restype = ctypes.c_int
    f.argtypes = [P(ctypes.c_uint16), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_size_t,
                  P(capi.CRAWProfile), P(capi.CRAWNoiseEstimateParams),
                  P(capi.CRAWNoiseEstimate)]
    REF.craw_noise_estimate_params_init.argtypes = [P(capi.CRAWNoiseEstimateParams)]
    REF.craw_noise_estimate_params_init.restype = None
    return REF


def estimate(l, mosaic, prof, params=None):
    h, w = mosaic.shape
    m = np.ascontiguousarray(mosaic, dtype=np.uint16)
    out = capi.CRAWNoiseEstimate()
    ret = l.craw_noise_estimate(m.ctypes.data_as(P(ctypes.c_uint16)), w, h, w,
                                ctypes.byref(prof),
                                ctypes.byref(params) if params is not None else None,
                                ctypes.byref(out))
    return ret, out


def fields(e):
    f32 = lambda x: int(np.array([x], np.float32).view(np.uint32)[0])
    return (f32(e.model.a), f32(e.model.b), e.nb_bins, e.nb_inliers, f32(e.fit_rms),
            f32(e.read_noise_lsb))


def params(l, **kw):
    p = capi.CRAWNoiseEstimateParams()
    l.craw_noise_estimate_params_init(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def smooth_field(w, h, a, b, seed):
    """A smooth ramp in every channel under Gaussian noise of variance
    a u + b, encoded as the 12-bit mosaic the scene generator makes."""
    rng = np.random.Generator(np.random.Philox(seed))
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    u = 0.02 + 0.9 * (x / w) * (0.3 + 0.7 * y / h)
    u = u + rng.standard_normal(u.shape) * np.sqrt(a * u + b)
    black, white = scene.BLACK_12, scene.WHITE_12
    raw = np.rint((black + u * (white - black)) * 65535.0)
    return np.clip(raw, 0, 4095).astype(np.uint16)


@needs_ref
@pytest.mark.parametrize("case", range(6))
def test_estimate_agrees_bitwise(case):
    ref = bind_ref()
    lib = capi.lib()
    prof = scene.make_profile()
    w, h, a, b, kw = [
        (512, 384, 2e-4, 4e-6, {}),
        (640, 480, 8e-4, 3e-5, {}),
        (1024, 768, 4e-4, 1.6e-5, {"nb_bins": 64}),
        (768, 512, 1e-4, 2e-6, {"nb_bins": 256, "min_bin_samples": 500}),
        (512, 512, 4e-4, 1.6e-5, {"max_level": 0.5, "min_bin_samples": 1000}),
        (96, 64, 4e-4, 1.6e-5, {"min_bin_samples": 50}),
    ][case]
    mosaic = scene.make_frame(w, h, seed=90 + case, a=a, b=b)
    ra, ea = estimate(lib, mosaic, prof, params(lib, **kw))
    rb, eb = estimate(ref, mosaic, prof, params(ref, **kw))
    assert ra == rb
    if ra == 0:
        assert fields(ea) == fields(eb)
    # the smooth field too
    f = smooth_field(w, h, a, b, seed=case)
    ra, ea = estimate(lib, f, prof, params(lib, **kw))
    rb, eb = estimate(ref, f, prof, params(ref, **kw))
    assert ra == rb and (ra != 0 or fields(ea) == fields(eb))


@pytest.mark.parametrize("a,b", [(2e-4, 4e-6), (8e-4, 3e-5)])
def test_smooth_field_model_recovered(a, b):
    lib = capi.lib()
    prof = scene.make_profile()
    ret, e = estimate(lib, smooth_field(1024, 768, a, b, seed=5), prof,
                      params(lib, min_bin_samples=2000))
    assert ret == 0
    assert abs(e.model.a / a - 1.0) < 0.05, e.model.a
    assert abs(e.model.b - b) < 0.25 * b + 2e-6, e.model.b
    assert e.fit_rms < 0.1 and e.nb_inliers >= 0.8 * e.nb_bins


def test_estimate_errors():
    lib = capi.lib()
    prof = scene.make_profile()
    m = smooth_field(64, 64, 2e-4, 4e-6, seed=1)
    assert estimate(lib, m[:, :31], prof)[0] == -EINVAL        # odd / too narrow
    assert estimate(lib, m[:30, :30], prof)[0] == -EINVAL
    assert estimate(lib, m, prof, params(lib, nb_bins=4))[0] == -EINVAL
    other = scene.make_profile()
    other.cfa = 1
    assert estimate(lib, m, other)[0] == -ENOTSUP
    # a 64x64 frame has too few samples for the default bins
    assert estimate(lib, m, prof)[0] == -ENODATA
