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
        out = (capi.CRAWStageTime * 32)()
        n = self._lib.craw_pipeline_get_stage_times_n(self.handle, frames_back, out, 32)
        if n < 0:
            raise CrawError("get_stage_times", n)
        return [(out[i].name.decode(), out[i].count, out[i].ms) for i in range(min(n, 32))]

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

    def set_launch_shape(self, wg_per_cu, denoise_options=0):
        arr = (ctypes.c_int32 * 4)(*wg_per_cu)
        _ok(self._lib.craw_batch_set_launch_shape(self.handle, arr, denoise_options),
            "craw_batch_set_launch_shape")

    def launch_shape(self):
        arr = (ctypes.c_int32 * 4)()
        opt = ctypes.c_int32()
        _ok(self._lib.craw_batch_get_launch_shape(self.handle, arr, ctypes.byref(opt)),
            "craw_batch_get_launch_shape")
        return list(arr), opt.value

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


def server_params(max_width, max_height, max_batch_pixels, max_batch_frames=16, nb_slots=2,
                  hbm_budget=64 << 30, debayer=capi.CRAW_DEBAYER_RCD,
                  out_space=capi.CRAW_CS_XYZ_D65,
                  out_format=capi.CRAW_FORMAT_R32G32B32A32_SFLOAT, flags=0, use_graphs=False,
                  max_queue=1024):
    sp = capi.CRAWServerParams()
    capi.lib().craw_server_params_init(ctypes.byref(sp))
    sp.hbm_budget, sp.max_batch_pixels = hbm_budget, max_batch_pixels
    sp.nb_slots, sp.max_batch_frames = nb_slots, max_batch_frames
    sp.max_width, sp.max_height = max_width, max_height
    sp.debayer, sp.out_space, sp.out_format = debayer, out_space, out_format
    sp.flags, sp.use_graphs, sp.max_queue = flags, 1 if use_graphs else 0, max_queue
    return sp


class Server:
    """Continuous-batching request loop (include/craw/serve.h)."""

    def __init__(self, ctx, params):
        self._lib = capi.lib()
        self.ctx = ctx
        self.params = params
        self.handle = ctypes.c_void_p()
        self._keep = {}
        _ok(self._lib.craw_server_create(ctx.handle, ctypes.byref(params),
                                         ctypes.byref(self.handle)), "craw_server_create")

    def open_stream(self, width, height):
        sid = ctypes.c_int32()
        _ok(self._lib.craw_server_open_stream(self.handle, width, height, ctypes.byref(sid)),
            "craw_server_open_stream")
        return sid.value

    def close_stream(self, sid):
        _ok(self._lib.craw_server_close_stream(self.handle, sid), "craw_server_close_stream")

    @staticmethod
    def request(src, dst, fp, stream=-1, tag=0, histogram=None):
        """A CRAWServeRequest. `histogram`: None or a uint32 array of
        3 x 128 x 128 words, filled before the completion is posted."""
        r = capi.CRAWServeRequest()
        r.src, r.dst, r.params = src, dst, fp
        r.stream, r.tag = stream, tag
        if histogram is not None:
            assert histogram.dtype == np.uint32 and histogram.size == 3 * 128 * 128
            r.awb_histogram = histogram.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        r._refs = (fp, histogram)
        return r

    def submit(self, src, dst, fp, stream=-1, tag=0, histogram=None):
        """Returns the C API's return code."""
        return self.submit_n([self.request(src, dst, fp, stream, tag, histogram)])

    def submit_n(self, reqs):
        """Queues requests (from request()) together; the C API's return code."""
        arr = (capi.CRAWServeRequest * max(len(reqs), 1))()
        for i, r in enumerate(reqs):
            arr[i] = r
        ret = self._lib.craw_server_submit_n(self.handle, arr, len(reqs))
        if ret == 0:
            # the profiles and the histograms must outlive the requests
            for r in reqs:
                self._keep[r.tag] = r._refs
        return ret

    def poll(self, max_n=256, timeout_ms=-1):
        out = (capi.CRAWServeCompletion * max_n)()
        n = _ok(self._lib.craw_server_poll(self.handle, out, max_n, timeout_ms),
                "craw_server_poll")
        res = []
        for i in range(n):
            c = out[i]
            self._keep.pop(c.tag, None)
            res.append({"tag": c.tag, "status": c.status, "stream": c.stream,
                        "submit_ms": c.submit_ms, "launch_ms": c.launch_ms,
                        "done_ms": c.done_ms, "batch_frames": c.batch_frames, "slot": c.slot})
        return res

    def drain(self):
        _ok(self._lib.craw_server_drain(self.handle), "craw_server_drain")

    def memory_usage(self):
        return int(self._lib.craw_server_get_memory_usage(self.handle))

    def close(self):
        if self.handle.value:
            self._lib.craw_server_free(ctypes.byref(self.handle))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
