"""Shared GPU test plumbing: run a frame through libcraw's C API."""
import ctypes

import numpy as np

from libcraw_amd import capi, hip
from libcraw_amd.pipeline import Context, Pipeline, frame_params, image, pipeline_params

_ctx = None


def context():
    global _ctx
    if _ctx is None:
        _ctx = Context(0)
    return _ctx


def half_to_float(u16):
    return u16.view(np.float16).astype(np.float32)


class Runner:
    """One pipeline plus device buffers for frames of one size."""

    def __init__(self, width, height, debayer=capi.CRAW_DEBAYER_RCD,
                 out_space=capi.CRAW_CS_XYZ_D65, flags=0, half=False, nb_awb_slots=0):
        self.w, self.h = width, height
        self.half = half
        self.fmt = (capi.CRAW_FORMAT_R16G16B16A16_SFLOAT if half
                    else capi.CRAW_FORMAT_R32G32B32A32_SFLOAT)
        self.pp = pipeline_params(width, height, debayer=debayer, out_space=out_space,
                                  out_format=self.fmt, flags=flags, nb_awb_slots=nb_awb_slots)
        self.pipe = Pipeline(context(), self.pp)
        self.src = hip.DeviceBuffer(width * height * 2)
        self.dst = hip.DeviceBuffer(width * height * (8 if half else 16))
        self.stream = hip.Stream()

    def run(self, bayer, fp, sync=True):
        assert bayer.shape == (self.h, self.w) and bayer.dtype == np.uint16
        self.src.upload(bayer)
        ret = self.pipe.record(self.stream, image(self.src, self.w, self.h,
                                                  capi.CRAW_FORMAT_R16_UNORM),
                               image(self.dst, self.w, self.h, self.fmt), fp)
        if ret < 0 or not sync:
            return ret
        self.stream.synchronize()
        return 0

    def output(self):
        if self.half:
            raw = self.dst.download(np.uint16, (self.h, self.w, 4))
            return half_to_float(raw)
        return self.dst.download(np.float32, (self.h, self.w, 4))


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bit_exact(gpu, ref, what):
    gb, rb = bits(gpu), bits(ref)
    diff = gb != rb
    if diff.any():
        idx = np.argwhere(diff)
        i = tuple(idx[0])
        raise AssertionError("%s: %d of %d values differ; first at %s: gpu %r ref %r "
                             "(max abs %.3g)" % (what, int(diff.sum()), diff.size, i,
                                                 float(gpu[i]), float(ref[i]),
                                                 float(np.nanmax(np.abs(gpu.astype(np.float64)
                                                                        - ref)))))
