"""Python binding of the oracle (liboracle.so) and of the reference's own
host sources built under oracle/_ref.

TEST INFRASTRUCTURE: imported by tests/, by __graft_entry__.smoke() and by
bench.py's CPU-baseline leg only, always as the checker or as the timed CPU
baseline, never by the library.

    f32l  float with glibc expf/logf/powf: lib/ref of cyanreg/libcraw as
          written;
    f32d  float with the deterministic transcendentals the kernels use:
          what the GPU reproduces bit for bit;
    f64   double throughout: the FP64 rerun for error statements.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CRAW_ORACLE_LIB names another build of the oracle (the Makefile's
# `sanitize` target)
LIB = os.environ.get("CRAW_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libcraw_ref.so")

_lib = None


def build(ref=True):
    subprocess.check_call(["make", "-C", HERE, "-j4"])
    if ref and os.path.isdir("/root/reference"):
        subprocess.check_call(["make", "-C", HERE, "ref"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("liboracle.so is not built; run make -C oracle")
        _lib = ctypes.CDLL(LIB)
    return _lib


_ref = None


def ref_available():
    """Whether oracle/_ref was built (the reference tree was present at
    build time). Loads nothing."""
    return os.path.exists(REF_LIB)


def ref_lib():
    """The reference's own lib/awb.c, colorspace.c, profile.c, gamut.c and
    noise.c, or None where they were not built (no reference tree). Loaded
    on the first call only, so that merely importing a test module maps
    nothing of it into the process."""
    global _ref
    if _ref is None and ref_available():
        _ref = ctypes.CDLL(REF_LIB)
    return _ref


class LazyRef:
    """Attribute access on the reference's host library, loading it on
    first use (tests that need it are skipped where it is absent)."""

    def __getattr__(self, name):
        lib_ = ref_lib()
        if lib_ is None:
            raise RuntimeError("oracle/_ref is not built (no reference tree)")
        return getattr(lib_, name)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Mode:
    def __init__(self, prefix, dtype, path=None):
        """One mode's set of functions; `path` names another build of the
        oracle (a planted-mistake variant of a test) instead of liboracle.so."""
        self.prefix = prefix
        self.dtype = np.dtype(dtype)
        self.path = path
        self._handle = None

    def lib(self):
        if self.path is None:
            return lib()
        if self._handle is None:
            self._handle = ctypes.CDLL(self.path)
        return self._handle

    def _f(self, name):
        f = getattr(self.lib(), self.prefix + name)
        f.restype = ctypes.c_int
        return f

    def _real(self):
        return ctypes.c_float if self.dtype == np.float32 else ctypes.c_double

    def set_threads(self, n):
        f = getattr(self.lib(), self.prefix + "set_threads")
        f.restype = None
        f(ctypes.c_int(n))

    # -- test hooks: the denoise and chroma weights one value at a time
    def gat(self, a, b, u, t_in=None):
        """T(u) for the array u, I(t) for the array t_in (default T(u)), and
        the constants (a, c, sc, fwd, inv, off, t0)."""
        u = np.array(u, self.dtype).ravel()
        t = np.empty_like(u)
        consts = np.empty(7, self.dtype)
        f = getattr(self.lib(), self.prefix + "gat_hook")
        f.restype = None
        first = np.empty_like(u)
        f(ctypes.c_float(a), ctypes.c_float(b), _ptr(u), _ptr(t), _ptr(u), _ptr(first),
          ctypes.c_size_t(u.size), _ptr(consts))
        # the inverse, of T(u) or of the given levels (one array length each)
        t_in = np.array(t if t_in is None else t_in, self.dtype).ravel()
        scratch = np.zeros_like(t_in)
        back = np.empty_like(t_in)
        f(ctypes.c_float(a), ctypes.c_float(b), _ptr(scratch), _ptr(np.empty_like(t_in)),
          _ptr(t_in), _ptr(back), ctypes.c_size_t(t_in.size), _ptr(consts))
        return t, back, consts

    def nlm_weight(self, d2, inv_h2, vc=-1.0, vq=-1.0):
        f = getattr(self.lib(), self.prefix + "nlm_weight_hook")
        r = self._real()
        f.restype, f.argtypes = r, [r, r, r, r]
        return f(d2, inv_h2, vc, vq)

    def alpha_weight(self, d2, alpha_max, inv_h2):
        f = getattr(self.lib(), self.prefix + "alpha_weight_hook")
        r = self._real()
        f.restype, f.argtypes = r, [r, r, r]
        return f(d2, alpha_max, inv_h2)

    def blend(self, tc, tp, a, v_s):
        """(the blended level, its variance, the variance as stored)"""
        f = getattr(self.lib(), self.prefix + "blend_hook")
        r = self._real()
        out = (r * 3)()
        f.restype, f.argtypes = None, [r, r, r, r, ctypes.c_void_p]
        f(tc, tp, a, v_s, out)
        return tuple(out)

    def cd_weight(self, dx, dy, g, dg, dd0, dd1):
        f = getattr(self.lib(), self.prefix + "cd_weight_hook")
        r = self._real()
        f.restype, f.argtypes = r, [ctypes.c_int, ctypes.c_int, r, r, r, r]
        return f(dx, dy, g, dg, dd0, dd1)

    def prep(self, bayer, profile, wb):
        h, w = bayer.shape
        v = np.empty((h, w), self.dtype)
        wb = np.asarray(wb, np.float32)
        f = getattr(self.lib(), self.prefix + "prep")
        f.restype = None
        f(_ptr(bayer), ctypes.c_uint32(w), ctypes.c_uint32(h), ctypes.c_size_t(w),
          ctypes.byref(profile), _ptr(wb), _ptr(v))
        return v

    def debayer_mosaic(self, v, method):
        h, w = v.shape
        v = np.ascontiguousarray(v, self.dtype)
        rgb = np.empty((h, w, 4), self.dtype)
        ret = self._f("debayer_mosaic")(_ptr(v), ctypes.c_uint32(w), ctypes.c_uint32(h),
                                        ctypes.c_int(method), _ptr(rgb))
        return ret, rgb

    def debayer_wb(self, bayer, profile, wb, method, stride=None):
        h, w = bayer.shape
        rgb = np.empty((h, w, 4), self.dtype)
        wb = np.asarray(wb, np.float32)
        ret = self._f("debayer_wb")(_ptr(bayer), ctypes.c_uint32(w), ctypes.c_uint32(h),
                                    ctypes.c_size_t(stride if stride is not None else w),
                                    ctypes.byref(profile), _ptr(wb), ctypes.c_int(method),
                                    _ptr(rgb))
        return ret, rgb

    def denoise_temporal(self, bayer, profile, a, b, strength, alpha_max=0.0, prev=None,
                         want_state=False, want_level=False):
        from libcraw_amd import capi
        h, w = bayer.shape
        nm = capi.CRAWNoiseModel(a, b)
        out = np.empty((h, w), np.uint16)
        nxt = np.empty((2, h, w), self.dtype) if (want_state or alpha_max > 0) else None
        lvl = np.empty((h, w), self.dtype) if want_level else None
        prev = None if prev is None else np.ascontiguousarray(prev, self.dtype)
        ret = self._f("denoise_temporal")(
            _ptr(bayer), ctypes.c_uint32(w), ctypes.c_uint32(h), ctypes.c_size_t(w),
            ctypes.byref(profile), ctypes.byref(nm), ctypes.c_float(strength),
            ctypes.c_float(alpha_max), _ptr(prev), _ptr(nxt), _ptr(out), ctypes.c_size_t(w),
            _ptr(lvl))
        return ret, out, nxt, lvl

    def chroma_denoise(self, rgb, strength):
        rgb = np.ascontiguousarray(rgb, self.dtype).copy()
        h, w, _ = rgb.shape
        ret = self._f("chroma_denoise")(_ptr(rgb), ctypes.c_uint32(w), ctypes.c_uint32(h),
                                        ctypes.c_float(strength))
        return ret, rgb

    def false_color(self, rgb, passes):
        rgb = np.ascontiguousarray(rgb, self.dtype).copy()
        h, w, _ = rgb.shape
        ret = self._f("false_color")(_ptr(rgb), ctypes.c_uint32(w), ctypes.c_uint32(h),
                                     ctypes.c_uint32(passes))
        return ret, rgb

    def highlights(self, bayer, profile, wb, stride=None):
        """The rebuilt, balanced mosaic (R16F-rounded in the float modes)."""
        h, w = bayer.shape
        out = np.empty((h, w), self.dtype)
        wb = np.asarray(wb, np.float32)
        ret = self._f("highlights")(_ptr(np.ascontiguousarray(bayer, np.uint16)),
                                    ctypes.c_uint32(w), ctypes.c_uint32(h),
                                    ctypes.c_size_t(stride if stride else w),
                                    ctypes.byref(profile), _ptr(wb), _ptr(out))
        return ret, out

    def color(self, rgb, fp, out_space):
        rgb = np.ascontiguousarray(rgb, self.dtype).copy()
        h, w, _ = rgb.shape
        ret = self._f("color")(_ptr(rgb), ctypes.c_uint32(w), ctypes.c_uint32(h),
                               ctypes.byref(fp), ctypes.c_int(out_space))
        return ret, rgb

    def gamut_compress(self, rgb3):
        rgb3 = np.ascontiguousarray(rgb3, self.dtype).copy()
        f = getattr(self.lib(), self.prefix + "gamut_compress")
        f.restype = None
        f(_ptr(rgb3), ctypes.c_size_t(rgb3.size // 3))
        return rgb3

    def awb_histogram(self, rgb, wb, channels=4, stride=None):
        rgb = np.ascontiguousarray(rgb, self.dtype)
        h, w = rgb.shape[0], rgb.shape[1]
        n = 128 * 128
        cnt, wt, br = (np.empty(n, np.uint32) for _ in range(3))
        wb = np.asarray(wb, np.float32)
        ret = self._f("awb_histogram")(_ptr(rgb), ctypes.c_uint32(w), ctypes.c_uint32(h),
                                       ctypes.c_size_t(stride if stride else w * channels),
                                       ctypes.c_int(channels), _ptr(wb), _ptr(cnt), _ptr(wt),
                                       _ptr(br))
        return ret, cnt, wt, br

    def process(self, bayer, pp, fp, prev=None, want_state=False, want_hist=False,
                want_inter=False, want_clean=False):
        """The whole path: returns dict(ret, out, [state, hist, inter, clean])."""
        h, w = bayer.shape
        out = np.empty((h, w, 4), self.dtype)
        nxt = np.empty((2, h, w), self.dtype) if want_state else None
        inter = np.empty((h, w, 4), self.dtype) if want_inter else None
        clean = np.empty((h, w), np.uint16) if want_clean else None
        hist = [np.empty(128 * 128, np.uint32) for _ in range(3)] if want_hist else [None] * 3
        prev = None if prev is None else np.ascontiguousarray(prev, self.dtype)
        ret = self._f("process")(_ptr(bayer), ctypes.c_uint32(w), ctypes.c_uint32(h),
                                 ctypes.c_size_t(w), ctypes.byref(pp), ctypes.byref(fp),
                                 _ptr(prev), _ptr(nxt), _ptr(clean), _ptr(inter),
                                 _ptr(hist[0]), _ptr(hist[1]), _ptr(hist[2]), _ptr(out))
        return dict(ret=ret, out=out, state=nxt, inter=inter, clean=clean,
                    hist=tuple(hist) if want_hist else None)


f32l = Mode("oracle_f32l_", np.float32)
f32d = Mode("oracle_f32d_", np.float32)
f64 = Mode("oracle_f64_", np.float64)


def host_fn(name, restype=ctypes.c_int):
    f = getattr(lib(), name)
    f.restype = restype
    return f
