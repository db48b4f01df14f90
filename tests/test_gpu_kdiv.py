"""kdiv_rn, the short division the streamed RCD uses, against IEEE division.

kdiv_rn (libcraw_amd/csrc/k_common.h) is the compiler's correctly rounded
float division without its scaling and fix-up steps, for divisors in
[2^-40, 2^40] and dividends +0 or in magnitude [2^-80, 2^40]. The planner
selects it only for frames whose constants keep the debayer's quotients
there. This test runs it on the GPU over operands across that range, over
quotients next to rounding midpoints, over the ratios the debayer forms,
and checks every result bit for bit against numpy's IEEE float32 division
(and the device's own division against the same).
"""
import ctypes
import os

import numpy as np
import pytest

from libcraw_amd import hip

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "hip", "libkdiv_check.so")

pytestmark = pytest.mark.gpu


def _lib():
    if not os.path.exists(LIB):
        raise RuntimeError("%s is not built (make -C tests/hip)" % LIB)
    lib = ctypes.CDLL(LIB)
    lib.kdiv_check.restype = ctypes.c_int
    lib.kdiv_check.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_size_t]
    return lib


def _f32(bits):
    return np.asarray(bits, dtype=np.uint32).view(np.float32)


def _log_uniform(rng, n, lo, hi):
    """Floats with exponents uniform in [lo, hi) and random significands."""
    e = rng.integers(lo, hi, size=n)
    m = rng.integers(0, 1 << 23, size=n, dtype=np.uint32)
    return _f32(((e + 127).astype(np.uint32) << 23) | m)


def _operands(seed=11):
    rng = np.random.default_rng(seed)
    ns, ds = [], []

    # across the range: divisors in [2^-40, 2^40), dividends in [2^-80, 2^40)
    n = 1 << 22
    d = _log_uniform(rng, n, -40, 40)
    x = _log_uniform(rng, n, -80, 40) * rng.choice(np.float32([-1, 1]), size=n)
    ns.append(x)
    ds.append(d)
    # +0 dividends (a -0 dividend would give +0: outside the range, and the
    # debayer forms none, its dividends being differences and sums of
    # non-negative values)
    ns.append(np.zeros(4096, np.float32))
    ds.append(_log_uniform(rng, 4096, -40, 40))

    # quotients next to a midpoint of the float grid: n = fl(m d) for a
    # midpoint m between a float and its successor
    n = 1 << 21
    q = _log_uniform(rng, n, -30, 30)
    qb = q.view(np.uint32)
    mid = (q.astype(np.float64) + _f32(qb + 1).astype(np.float64)) / 2.0
    d = _log_uniform(rng, n, -20, 20)
    x = (mid * d.astype(np.float64)).astype(np.float32)
    ok = np.isfinite(x) & (np.abs(x) >= 2.0 ** -80) & (np.abs(x) < 2.0 ** 40)
    ns.append(x[ok])
    ds.append(d[ok])

    # significands of all ones and all zeros, and n a few ulps from d
    sig = np.array([0, 1, 0x400000, 0x7ffffe, 0x7fffff], dtype=np.uint32)
    e = np.arange(-40, 40)
    dd = _f32(((e[:, None] + 127).astype(np.uint32) << 23) | sig[None, :]).ravel()
    for k in (-3, -1, 0, 1, 3):
        nb = dd.view(np.uint32).astype(np.int64) + k
        ns.append(_f32(nb.astype(np.uint32)))
        ds.append(dd)
    ns.append(np.repeat(dd, dd.size))
    ds.append(np.tile(dd, dd.size))

    # the debayer's ratios: (a - b) / (a + b + 1e-5) and w / (w + v + 1e-5)
    # for a, b >= 0 of prepped samples (multiples of 2^-28 up to 1, and
    # small values down to 2^-60)
    n = 1 << 21
    a = (rng.integers(0, 1 << 28, size=n) * 2.0 ** -28).astype(np.float32)
    b = (rng.integers(0, 1 << 28, size=n) * 2.0 ** -28).astype(np.float32)
    a[: n // 4] = _log_uniform(rng, n // 4, -60, 0)
    eps = np.float32(1e-5)
    ns.append(a - b)
    ds.append((a + b) + eps)
    ns.append(a)
    ds.append((a + b) + eps)
    return np.concatenate(ns).astype(np.float32), np.concatenate(ds).astype(np.float32)


def test_kdiv_rn_equals_ieee_division(gpu_available):
    lib = _lib()
    n, d = _operands()
    assert np.all(np.isfinite(n)) and np.all(d > 0)
    count = n.size
    bufs = [hip.DeviceBuffer(count * 4) for _ in range(4)]
    bufs[0].upload(n)
    bufs[1].upload(d)
    assert lib.kdiv_check(bufs[0].value, bufs[1].value, bufs[2].value, bufs[3].value,
                          count) == 0
    q_rn = bufs[2].download(np.float32, (count,))
    q_div = bufs[3].download(np.float32, (count,))
    with np.errstate(all="ignore"):
        ref = np.divide(n, d, dtype=np.float32)
    bad_div = int(np.count_nonzero(q_div.view(np.uint32) != ref.view(np.uint32)))
    bad_rn = q_rn.view(np.uint32) != ref.view(np.uint32)
    print("kdiv_rn: %d operand pairs, %d differ from IEEE division (device division: %d)"
          % (count, int(bad_rn.sum()), bad_div))
    assert bad_div == 0
    if bad_rn.any():
        i = int(np.flatnonzero(bad_rn)[0])
        raise AssertionError("kdiv_rn(%r, %r) = %r, IEEE %r (%d pairs differ)"
                             % (float(n[i]), float(d[i]), float(q_rn[i]), float(ref[i]),
                                int(bad_rn.sum())))
    for b in bufs:
        b.close()
