"""Python host mirror of libcraw's pipeline interface (include/craw/pipeline.h).

Same operations, argument meaning and error codes as the C API, which it
calls through ctypes: craw_context_create, craw_pipeline_create, record,
get_awb_histogram(_n), get_stage_times(_n), and the batched form of
include/craw/batch.h. Errors come back as the C API's
negative errno; `CrawError` carries it.
"""
import ctypes

import numpy as np

from . import capi, hip


class CrawError(RuntimeError):
    def __init__(self, what, code):
        super().__init__("%s failed: %d" % (what, code))
        self.code = code


def _ok(ret, what):
    if ret < 0:
        raise CrawError(what, ret)
    return ret


class Context:
    def __init__(self, device=0):
        self._lib = capi.lib()
        self.handle = ctypes.c_void_p()
        init = capi.CRAWHipInitParams()
        init.device = device
        _ok(self._lib.craw_context_create(ctypes.byref(self.handle), ctypes.byref(init), None),
            "craw_context_create")

    def close(self):
        if self.handle.value:
            self._lib.craw_context_free(ctypes.byref(self.handle))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pipeline_params(width, height, debayer=capi.CRAW_DEBAYER_RCD,
                    out_space=capi.CRAW_CS_XYZ_D65,
                    out_format=capi.CRAW_FORMAT_R32G32B32A32_SFLOAT, flags=0, nb_awb_slots=0):
    pp = capi.CRAWPipelineParams()
    capi.lib().craw_pipeline_params_init(ctypes.byref(pp))
    pp.width, pp.height = width, height
    pp.debayer, pp.out_space, pp.out_format = debayer, out_space, out_format
    pp.flags, pp.nb_awb_slots = flags, nb_awb_slots
    return pp


def frame_params(profile, **kw):
    """CRAWFrameParams; keyword arguments set fields by name
    (noise_override=(a, b), illuminant=(cct, tint))."""
    fp = capi.CRAWFrameParams()
    capi.lib().craw_frame_params_init(ctypes.byref(fp))
    fp.profile = ctypes.pointer(profile)
    fp._profile_ref = profile
    for k, v in kw.items():
        if k == "noise_override":
            fp.noise_override.a, fp.noise_override.b = v
        elif k == "illuminant":
            fp.illuminant.cct, fp.illuminant.tint = v
        else:
            setattr(fp, k, v)
    return fp


def image(buf, width, height, fmt, pitch=None):
    bpp = {capi.CRAW_FORMAT_R16_UNORM: 2, capi.CRAW_FORMAT_R16G16B16A16_SFLOAT: 8,
           capi.CRAW_FORMAT_R32G32B32A32_SFLOAT: 16}[fmt]
    img = capi.CRAWImage()
    img.data = buf.value if isinstance(buf, hip.DeviceBuffer) else buf
    img.pitch = pitch if pitch is not None else width * bpp
    img.format = fmt
    img.width, img.height = width, height
    return img


class Pipeline:
    def __init__(self, ctx, params):
        self._lib = capi.lib()
        self.ctx = ctx
        self.params = params
        self.handle = ctypes.c_void_p()
        _ok(self._lib.craw_pipeline_create(ctx.handle, ctypes.byref(params),
                                           ctypes.byref(self.handle)), "craw_pipeline_create")

    def record(self, stream, src, dst, fp):
        """Returns the C API's return code (0 or a negative errno)."""
        return self._lib.craw_pipeline_record(self.handle, stream.handle if stream else None,
                                              ctypes.byref(src), ctypes.byref(dst),
                                              ctypes.byref(fp))

    def histogram(self, frames_back=0):
        h = capi.CRAWAwbHistogram()
        _ok(self._lib.craw_pipeline_get_awb_histogram_n(self.handle, frames_back,
                                                        ctypes.byref(h)), "get_awb_histogram")
        n = h.bins * h.bins
        return tuple(np.ctypeslib.as_array(p, shape=(n,)).copy()
                     for p in (h.count, h.weight, h.weight_bright))

    def stage_times(self, frames_back=0):
        out = (capi.CRAWStageTime * 8)()
        n = self._lib.craw_pipeline_get_stage_times_n(self.handle, frames_back, out, 8)
        if n < 0:
            raise CrawError("get_stage_times", n)
        return [(out[i].name.decode(), out[i].count, out[i].ms) for i in range(min(n, 8))]

    def close(self):
        if self.handle.value:
            self._lib.craw_pipeline_free(ctypes.byref(self.handle))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_params(max_frames, max_width, max_height, max_pixels, debayer=capi.CRAW_DEBAYER_RCD,
                 out_space=capi.CRAW_CS_XYZ_D65,
                 out_format=capi.CRAW_FORMAT_R32G32B32A32_SFLOAT, flags=0, use_graphs=False):
    bp = capi.CRAWBatchParams()
    capi.lib().craw_batch_params_init(ctypes.byref(bp))
    bp.max_frames, bp.max_width, bp.max_height = max_frames, max_width, max_height
    bp.max_pixels = max_pixels
    bp.debayer, bp.out_space, bp.out_format = debayer, out_space, out_format
    bp.flags, bp.use_graphs = flags, 1 if use_graphs else 0
    return bp


class TemporalState:
    """One clip stream's temporal denoise state (craw_temporal_state_*)."""

    def __init__(self, ctx, width, height):
        self._lib = capi.lib()
        self.ctx = ctx
        self.handle = ctypes.c_void_p()
        _ok(self._lib.craw_temporal_state_create(ctx.handle, width, height,
                                                 ctypes.byref(self.handle)),
            "craw_temporal_state_create")

    def reset(self):
        self._lib.craw_temporal_state_reset(self.handle)

    def close(self):
        if self.handle.value:
            self._lib.craw_temporal_state_free(ctypes.byref(self.handle))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_frame(src, dst, fp, state=None):
    f = capi.CRAWBatchFrame()
    f.src, f.dst = src, dst
    f.params = ctypes.pointer(fp)
    f.state = state.handle.value if state is not None else None
    f._refs = (fp, state)
    return f


class Batch:
    """Many frames of any sizes per recording (include/craw/batch.h)."""

    def __init__(self, ctx, params):
        self._lib = capi.lib()
        self.ctx = ctx
        self.params = params
        self.handle = ctypes.c_void_p()
        _ok(self._lib.craw_batch_create(ctx.handle, ctypes.byref(params),
                                        ctypes.byref(self.handle)), "craw_batch_create")

    @staticmethod
    def _array(frames):
        arr = (capi.CRAWBatchFrame * max(len(frames), 1))()
        for i, f in enumerate(frames):
            arr[i] = f
        return arr

    def record(self, stream, frames):
        """Returns the C API's return code (0 or a negative errno)."""
        return self._lib.craw_batch_record(self.handle, stream.handle if stream else None,
                                           self._array(frames), len(frames))

    def autotune(self, stream, frames, iterations=3):
        return _ok(self._lib.craw_batch_autotune(self.handle, stream.handle if stream else None,
                                                 self._array(frames), len(frames), iterations),
                   "craw_batch_autotune")

    def histogram(self, index):
        h = capi.CRAWAwbHistogram()
        _ok(self._lib.craw_batch_get_awb_histogram(self.handle, index, ctypes.byref(h)),
            "craw_batch_get_awb_histogram")
        n = h.bins * h.bins
        return tuple(np.ctypeslib.as_array(p, shape=(n,)).copy()
                     for p in (h.count, h.weight, h.weight_bright))

    def stage_times(self):
        out = (capi.CRAWStageTime * 8)()
        n = self._lib.craw_batch_get_stage_times(self.handle, out, 8)
        if n < 0:
            raise CrawError("craw_batch_get_stage_times", n)
        return [(out[i].name.decode(), out[i].count, out[i].ms) for i in range(min(n, 8))]

    def close(self):
        if self.handle.value:
            self._lib.craw_batch_free(ctypes.byref(self.handle))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
