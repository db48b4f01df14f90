"""The serving form (include/craw/serve.h): continuous batching over clip
streams and stills.

Bar: each request's output and AWB histogram equal, bit for bit, to the
plain pipeline's for the same frame and parameters, whether it ran alone or
among many others in the mix; a clip stream's frames, all submitted at
once, developed in order on the stream's own temporal state, equal to a
plain pipeline running that clip alone; the HBM budget refuses what does
not fit; a request that cannot be planned fails alone.
"""
import numpy as np
import pytest

from libcraw_amd import capi, hip, scene
from libcraw_amd.pipeline import Server, frame_params, image, server_params

from gpu_util import Runner, assert_bit_exact, context

pytestmark = pytest.mark.gpu

F32 = capi.CRAW_FORMAT_R32G32B32A32_SFLOAT
FULL = (capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE |
        capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM | capi.CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE)


class Job:
    """A frame on the device with its own destination."""

    def __init__(self, mosaic, fp, stream=-1):
        self.h, self.w = mosaic.shape
        self.mosaic, self.fp, self.stream = mosaic, fp, stream
        self.src = hip.DeviceBuffer(self.w * self.h * 2)
        self.src.upload(mosaic)
        self.dst = hip.DeviceBuffer(self.w * self.h * 16)
        self.hist = np.zeros(3 * 128 * 128, np.uint32)

    def request(self, tag):
        return Server.request(image(self.src, self.w, self.h, capi.CRAW_FORMAT_R16_UNORM),
                              image(self.dst, self.w, self.h, F32), self.fp,
                              stream=self.stream, tag=tag, histogram=self.hist)

    def submit(self, srv, tag):
        return srv.submit_n([self.request(tag)])

    def output(self):
        return self.dst.download(np.float32, (self.h, self.w, 4))


def new_server(flags=FULL, nb_slots=2, max_frames=8, max_px=600_000, budget=8 << 30):
    return Server(context(), server_params(320, 240, max_px, max_batch_frames=max_frames,
                                           nb_slots=nb_slots, hbm_budget=budget, flags=flags))


def wait_all(srv, n):
    got = []
    while len(got) < n:
        got += srv.poll(timeout_ms=-1)
    return {c["tag"]: c for c in got}


def plain_of(job, flags):
    run = Runner(job.w, job.h, flags=flags)
    assert run.run(job.mosaic, job.fp) == 0
    return run.output(), np.concatenate(run.pipe.histogram())


def still(w, h, seed, prof):
    return Job(scene.make_frame(w, h, seed=seed),
               frame_params(prof, denoise_strength=1.0, denoise_chroma_strength=0.5,
                            noise_override=(2e-4, 4e-6)))


def test_still_alone_and_in_the_mix(gpu_available):
    prof = scene.make_profile()
    flags = FULL & ~capi.CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE
    probe = still(202, 122, 77, prof)
    ref, ref_hist = plain_of(probe, flags)

    srv = new_server(flags=flags)
    assert probe.submit(srv, 1) == 0
    c = wait_all(srv, 1)[1]
    assert c["status"] == 0 and c["batch_frames"] == 1
    assert_bit_exact(probe.output(), ref, "alone")
    assert np.array_equal(probe.hist, ref_hist)

    # the same request again among frames of other sizes, all queued at once
    others = [still(w, h, 100 + i, prof) for i, (w, h) in
              enumerate([(320, 240), (96, 64), (256, 160), (64, 48), (130, 66), (200, 136)])]
    probe.dst.zero()
    probe.hist[:] = 0
    mix = ([j.request(10 + i) for i, j in enumerate(others[:3])] + [probe.request(2)] +
           [j.request(20 + i) for i, j in enumerate(others[3:])])
    assert srv.submit_n(mix) == 0
    done = wait_all(srv, 7)
    assert all(c["status"] == 0 for c in done.values())
    assert done[2]["batch_frames"] == 7
    assert_bit_exact(probe.output(), ref, "in the mix")
    assert np.array_equal(probe.hist, ref_hist)
    for i, j in enumerate(others):
        out, hist = plain_of(j, flags)
        assert_bit_exact(j.output(), out, "other %d" % i)
        assert np.array_equal(j.hist, hist)
    srv.close()


def test_highlights_and_false_colour_served(gpu_available):
    """Requests with highlight reconstruction and false colour, mixed with
    ones without, in one server's batches: each frame as the plain form
    develops it."""
    prof = scene.make_profile()
    flags = (capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE | capi.CRAW_PIPELINE_FLAG_AWB_HISTOGRAM |
             capi.CRAW_PIPELINE_FLAG_HIGHLIGHT_RECONSTRUCTION |
             capi.CRAW_PIPELINE_FLAG_FALSE_COLOR)
    jobs = []
    for i, (w, h) in enumerate([(202, 122), (96, 64), (256, 160), (130, 66)]):
        m = scene.make_frame(w, h, seed=600 + i)
        m[h // 4:h // 2, w // 3:w // 2] = 4095
        fp = frame_params(prof, flags=capi.CRAW_FRAME_FLAG_HIGHLIGHT_RECONSTRUCTION
                          if i % 2 == 0 else 0, false_color_passes=(i + 1) % 3,
                          denoise_chroma_strength=0.5)
        jobs.append(Job(m, fp))
    srv = new_server(flags=flags)
    assert srv.submit_n([j.request(i) for i, j in enumerate(jobs)]) == 0
    done = wait_all(srv, len(jobs))
    assert all(c["status"] == 0 for c in done.values())
    for i, j in enumerate(jobs):
        out, hist = plain_of(j, flags)
        assert_bit_exact(j.output(), out, "request %d" % i)
        assert np.array_equal(j.hist, hist)
    srv.close()


def test_clip_streams_in_order(gpu_available):
    """Three clips, five frames each, all submitted before any completes;
    one clip resets at its third frame. Each clip equals a plain pipeline
    developing it alone."""
    prof = scene.make_profile()
    sizes = [(160, 96), (128, 80), (96, 64)]
    srv = new_server(max_frames=4)
    sids = [srv.open_stream(w, h) for w, h in sizes]
    jobs = {}
    for t in range(5):
        for c, (w, h) in enumerate(sizes):
            fp = frame_params(prof, denoise_strength=1.0, denoise_chroma_strength=0.5,
                              denoise_temporal_strength=0.8, noise_override=(4e-4, 1.6e-5))
            if c == 1 and t == 2:
                fp.flags |= capi.CRAW_FRAME_FLAG_TEMPORAL_RESET
            jobs[(c, t)] = Job(scene.clip_frame(w, h, t, seed=30 + c, square=40, speed=4), fp,
                               stream=sids[c])
    # stills in between, queued behind and among the clip frames
    stills = [still(200, 136, 300 + i, prof) for i in range(3)]
    reqs = []
    for t in range(5):
        for c in range(3):
            reqs.append(jobs[(c, t)].request(100 * c + t))
        if t < 3:
            reqs.append(stills[t].request(1000 + t))
    assert srv.submit_n(reqs) == 0
    done = wait_all(srv, 18)
    assert all(c["status"] == 0 for c in done.values())
    # a stream's frames completed in order
    for c in range(3):
        ends = [done[100 * c + t]["launch_ms"] for t in range(5)]
        assert ends == sorted(ends)
    for c, (w, h) in enumerate(sizes):
        run = Runner(w, h, flags=FULL)
        for t in range(5):
            j = jobs[(c, t)]
            assert run.run(j.mosaic, j.fp) == 0
            assert_bit_exact(j.output(), run.output(), "clip %d frame %d" % (c, t))
            assert np.array_equal(j.hist, np.concatenate(run.pipe.histogram()))
    for i, s in enumerate(stills):
        out, hist = plain_of(s, FULL)
        assert_bit_exact(s.output(), out, "still %d" % i)
    for sid in sids:
        srv.close_stream(sid)
    srv.close()


def test_budget_and_errors(gpu_available):
    prof = scene.make_profile()
    # the slots alone do not fit
    with pytest.raises(Exception):
        new_server(budget=1 << 20)
    srv = new_server(budget=256 << 20)
    base = srv.memory_usage()
    assert 0 < base <= 256 << 20
    # each 320x240 stream holds 10 bytes a sample
    sids = []
    with pytest.raises(Exception):
        while True:
            sids.append(srv.open_stream(320, 240))
    assert srv.memory_usage() == base + len(sids) * 10 * 320 * 240
    srv.close_stream(sids.pop())
    assert srv.memory_usage() == base + len(sids) * 10 * 320 * 240
    sids.append(srv.open_stream(320, 240))

    good = still(96, 64, 5, prof)
    bad = still(96, 64, 6, prof)
    bad.fp.denoise_chroma_strength = 2.0      # planned as invalid
    big = still(96, 64, 7, prof)
    assert srv.submit(image(big.src, 400, 64, capi.CRAW_FORMAT_R16_UNORM),
                      image(big.dst, 400, 64, F32), big.fp) == -capi.EINVAL
    assert good.submit(srv, 1) == 0
    assert bad.submit(srv, 2) == 0
    # an unknown stream; a stream of another size
    assert srv.submit(image(good.src, 96, 64, capi.CRAW_FORMAT_R16_UNORM),
                      image(good.dst, 96, 64, F32), good.fp, stream=999) == -capi.EINVAL
    assert srv.submit(image(good.src, 96, 64, capi.CRAW_FORMAT_R16_UNORM),
                      image(good.dst, 96, 64, F32), good.fp, stream=sids[0]) == -capi.EINVAL
    done = wait_all(srv, 2)
    assert done[1]["status"] == 0 and done[2]["status"] == -capi.EINVAL
    out, _ = plain_of(good, FULL)
    assert_bit_exact(good.output(), out, "good beside bad")
    srv.drain()
    for sid in sids:
        srv.close_stream(sid)
    srv.close()


def test_fp16_temporal_streams_equal_plain_fp16(gpu_available):
    """The served workload's own form: packed-FP16 NLM with the temporal
    pass on clip streams (and stills beside them), each frame equal to a
    plain FP16 pipeline developing that clip alone."""
    prof = scene.make_profile()
    flags = FULL | capi.CRAW_PIPELINE_FLAG_NLM_FP16
    sizes = [(160, 96), (200, 136)]
    srv = new_server(flags=flags, max_frames=3)
    sids = [srv.open_stream(w, h) for w, h in sizes]
    jobs, reqs = {}, []
    for t in range(4):
        for c, (w, h) in enumerate(sizes):
            fp = frame_params(prof, denoise_strength=1.0, denoise_chroma_strength=0.5,
                              denoise_temporal_strength=0.8,
                              noise_override=(2e-4 * (1 + c), 4e-6 * (1 + c)))
            jobs[(c, t)] = Job(scene.clip_frame(w, h, t, seed=70 + c, square=40, speed=4), fp,
                               stream=sids[c])
            reqs.append(jobs[(c, t)].request(100 * c + t))
    st = [still(128, 96, 900 + i, prof) for i in range(2)]
    reqs += [s.request(1000 + i) for i, s in enumerate(st)]
    assert srv.submit_n(reqs) == 0
    done = wait_all(srv, len(reqs))
    assert all(c["status"] == 0 for c in done.values())
    for c, (w, h) in enumerate(sizes):
        run = Runner(w, h, flags=flags)
        for t in range(4):
            j = jobs[(c, t)]
            assert run.run(j.mosaic, j.fp) == 0
            assert_bit_exact(j.output(), run.output(), "fp16 clip %d frame %d" % (c, t))
            assert np.array_equal(j.hist, np.concatenate(run.pipe.histogram()))
    for i, s in enumerate(st):
        out, hist = plain_of(s, flags)
        assert_bit_exact(s.output(), out, "fp16 still %d" % i)
        assert np.array_equal(s.hist, hist)
    for sid in sids:
        srv.close_stream(sid)
    srv.close()
