"""The colour stage's gamut compression (kgamut) on the GPU against the oracle.

kgamut (libcraw_amd/csrc/k_common.h) is lib/gamut.c:66-92 with a screen in
front: a triple whose three distances are, by an estimate from the
hardware reciprocal, clearly below their thresholds returns unchanged
without the three correctly rounded divisions. These tests run kgamut over
triples whose distances sit within a few ulps either side of a threshold,
at magnitudes inside and outside the screen's range, over scene-like
triples and over special values, and check
- the exact form bit for bit against the oracle's gamut compression
  (oracle f32d, lib/gamut.c restated);
- both forms (exact and the precision form's float pow) with the screen
  against the same kernel with the screen switched off, bit for bit, also
  with a scale that makes every compression past a threshold visible.
The screen's bound itself is checked on the CPU: thr_lo / (1 - 2^-22) lies
below the midpoint between each threshold and its predecessor.
"""
import ctypes
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
from libcraw_amd import hip

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "hip", "libgamut_check.so")


def _lib():
    if not os.path.exists(LIB):
        raise RuntimeError("%s is not built (make -C tests/hip)" % LIB)
    lib = ctypes.CDLL(LIB)
    fp3 = ctypes.POINTER(ctypes.c_float * 3)
    lib.gamut_check.restype = ctypes.c_int
    lib.gamut_check.argtypes = [ctypes.c_void_p, ctypes.c_size_t, fp3, fp3, ctypes.c_float,
                                ctypes.c_float, ctypes.c_int, ctypes.c_int]
    lib.gamut_screen_bound.restype = ctypes.c_float
    lib.gamut_screen_bound.argtypes = [ctypes.c_float]
    return lib


def _params():
    thr, scl = (ctypes.c_float * 3)(), (ctypes.c_float * 3)()
    pwr, inv = ctypes.c_float(), ctypes.c_float()
    f = oracle.host_fn("oracle_gamut_params", None)
    f(thr, scl, ctypes.byref(pwr), ctypes.byref(inv))
    return thr, scl, pwr.value, inv.value


def _f32(bits):
    return np.asarray(bits, dtype=np.uint32).view(np.float32)


def _near_threshold(rng, thr):
    """Triples with one channel's distance swept over 64 ulps of its
    float value either side of the one at its threshold, the maximum in
    another channel, at magnitudes 2^-80 .. 2^127 of either sign."""
    out = []
    exps = [-80, -70, -65, -64, -63, -40, -10, -1, 0, 1, 10, 40, 62, 63, 64, 65, 70, 100,
            120, 126, 127]
    steps = np.arange(-64, 65, dtype=np.int64)
    for ch in range(3):
        k, j = (ch + 1) % 3, (ch + 2) % 3
        for e in exps:
            for sign in (1.0, -1.0):
                for _ in range(8):
                    m = np.float32(2.0 ** e * (1.0 + rng.random()))
                    ach = np.float32(sign * m)
                    mag = abs(float(ach))
                    xc = np.float32(float(ach) - float(thr[ch]) * mag)
                    xs = _f32((np.int64(xc.view(np.uint32)) + steps).astype(np.uint32))
                    t = np.empty((xs.size, 3), np.float32)
                    t[:, ch] = xs
                    t[:, k] = ach
                    t[:, j] = np.float32(float(ach) - 0.5 * float(thr[j]) * mag)
                    out.append(t)
    return np.concatenate(out)


def _scene_like(rng, n=1 << 20):
    t = rng.lognormal(-1.0, 1.2, size=(n, 3)).astype(np.float32)
    t[: n // 8] *= rng.choice(np.float32([-1.0, 1.0]), size=(n // 8, 3))
    t[n // 8: n // 4] = rng.uniform(-1.0, 2.0, size=(n // 8, 3)).astype(np.float32)
    return t


def _specials():
    s = np.array([0.0, -0.0, 1.0, -1.0, 0.5, 2.0 ** -149, -(2.0 ** -149), 2.0 ** -126, 1e-40,
                  3.4028235e38, -3.4028235e38, np.inf, -np.inf, np.nan, 2.0 ** 64, 2.0 ** -64,
                  2.0 ** 65, 2.0 ** -65], np.float32)
    a, b, c = np.meshgrid(s, s, s, indexing="ij")
    return np.stack([a.ravel(), b.ravel(), c.ravel()], axis=1).astype(np.float32)


def _run(lib, rgb, params, screen, float_math):
    thr, scl, pwr, inv = params
    buf = hip.DeviceBuffer(rgb.nbytes)
    buf.upload(rgb)
    assert lib.gamut_check(buf.value, rgb.shape[0], ctypes.byref(thr), ctypes.byref(scl),
                           pwr, inv, int(screen), int(float_math)) == 0
    out = buf.download(np.float32, rgb.shape)
    buf.close()
    return out


def _same(a, b):
    """Bit for bit, any NaN equal to any NaN."""
    na, nb = np.isnan(a), np.isnan(b)
    return (na == nb) & (na | (a.view(np.uint32) == b.view(np.uint32)))


def test_screen_bound_margin():
    """thr_lo / (1 - 2^-22) < (pred(thr) + thr) / 2 for the ACES thresholds
    and for random thresholds across [2^-100, 2^100]; -inf outside."""
    lib = _lib()
    thr, _, _, _ = _params()
    rng = np.random.default_rng(5)
    e = rng.integers(-100, 100, size=2000)
    m = rng.integers(0, 1 << 23, size=2000, dtype=np.uint32)
    cands = list(thr) + list(_f32(((e + 127).astype(np.uint32) << 23) | m))
    for t in cands:
        t = np.float32(t)
        lo = lib.gamut_screen_bound(float(t))
        pred = np.nextafter(t, np.float32(0))
        mid = (Fraction(float(pred)) + Fraction(float(t))) / 2
        assert 0 < lo < t
        assert Fraction(float(lo)) / (1 - Fraction(1, 1 << 22)) < mid, float(t)
    for t in (0.0, -0.5, 2.0 ** -101, 2.0 ** 101, float("nan"), float("inf")):
        assert lib.gamut_screen_bound(t) == -np.inf


@pytest.mark.gpu
def test_gamut_matches_oracle_and_screen_off(gpu_available):
    lib = _lib()
    params = _params()
    rng = np.random.default_rng(23)
    near = _near_threshold(rng, params[0])
    sets = {"near": near, "scene": _scene_like(rng), "special": _specials()}
    for name, rgb in sets.items():
        ref = oracle.f32d.gamut_compress(rgb)
        got = _run(lib, rgb, params, screen=True, float_math=False)
        off = _run(lib, rgb, params, screen=False, float_math=False)
        bad = ~np.all(_same(got, ref), axis=1)
        print("%s: %d triples, %d changed by the compression, %d differ from the oracle"
              % (name, rgb.shape[0], int((~np.all(_same(ref, rgb), axis=1)).sum()),
                 int(bad.sum())))
        if bad.any():
            i = int(np.flatnonzero(bad)[0])
            raise AssertionError("%s: kgamut(%r) = %r, oracle %r"
                                 % (name, rgb[i].tolist(), got[i].tolist(), ref[i].tolist()))
        assert np.all(_same(got, off))
        fm_on = _run(lib, rgb, params, screen=True, float_math=True)
        fm_off = _run(lib, rgb, params, screen=False, float_math=True)
        assert np.all(_same(fm_on, fm_off)), name

    # Just past a threshold the ACES curve is the identity to within an
    # ulp, so a channel the screen wrongly skipped there would look
    # unchanged either way. With a scale of 2^-40 every distance at or
    # past its threshold is pulled to the threshold itself, which moves
    # the channel: the same sweep, screen on against screen off.
    thr, scl, pwr, inv = params
    tiny = (ctypes.c_float * 3)(*([2.0 ** -40] * 3))
    for float_math in (False, True):
        on = _run(lib, near, (thr, tiny, pwr, inv), screen=True, float_math=float_math)
        off = _run(lib, near, (thr, tiny, pwr, inv), screen=False, float_math=float_math)
        changed = ~np.all(_same(off, near), axis=1)
        print("near, scale 2^-40 (float pow %d): %d of %d triples compressed"
              % (float_math, int(changed.sum()), near.shape[0]))
        assert 0.3 < changed.mean() < 0.7
        assert np.all(_same(on, off))
