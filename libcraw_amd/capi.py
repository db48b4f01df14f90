"""ctypes mirror of libcraw's C API (include/craw/*.h).

The structs are laid out field for field as the C headers declare them, so
that CRAWPipelineParams is 256 bytes and CRAWFrameParams 264, the sizes the
reference's own ABI checks hold (lib/pipeline.c:252-253 of cyanreg/libcraw).
`lib()` loads libcraw.so from this package directory and fails loudly if it
is not built: there is no fallback path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# CRAW_LIB names another build of the library (the sanitizer build of the
# csrc Makefile's `sanitize` target, for a host-only test run)
LIB_PATH = os.environ.get("CRAW_LIB") or os.path.join(HERE, "libcraw.so")

c_float = ctypes.c_float
c_u32 = ctypes.c_uint32
c_i32 = ctypes.c_int32

# enums (include/craw/*.h)
CRAW_CFA_RGGB = 0
CRAW_DEBAYER_BILINEAR, CRAW_DEBAYER_MHC, CRAW_DEBAYER_RCD = 0, 1, 2
(CRAW_CS_XYZ_D65, CRAW_CS_BT709_LINEAR, CRAW_CS_BT2020_LINEAR, CRAW_CS_P3D65_LINEAR,
 CRAW_CS_V_GAMUT_LINEAR, CRAW_CS_ACES_AP0_LINEAR, CRAW_CS_ACES_AP1_LINEAR) = range(7)
CRAW_CS_NB = 7
CRAW_WB_AS_SHOT, CRAW_WB_REBALANCE = 0, 1
CRAW_FORMAT_R16_UNORM = 1
CRAW_FORMAT_R16G16B16A16_SFLOAT = 2
CRAW_FORMAT_R32G32B32A32_SFLOAT = 3

CRAW_PIPELINE_FLAG_AWB_HISTOGRAM = 1 << 0
CRAW_PIPELINE_FLAG_DENOISE = 1 << 1
CRAW_PIPELINE_FLAG_LENS = 1 << 2
CRAW_PIPELINE_FLAG_FALSE_COLOR = 1 << 3
CRAW_PIPELINE_FLAG_CHROMA_DENOISE = 1 << 4
CRAW_PIPELINE_FLAG_HIGHLIGHT_RECONSTRUCTION = 1 << 5
CRAW_PIPELINE_FLAG_DEFRINGE = 1 << 6
CRAW_PIPELINE_FLAG_TEMPORAL_DENOISE = 1 << 7
CRAW_PIPELINE_FLAG_SHARPEN = 1 << 8
CRAW_PIPELINE_FLAG_PROFILE = 1 << 9
CRAW_PIPELINE_FLAG_NLM_FP16 = 1 << 16
CRAW_PIPELINE_FLAG_NLM_FAST = 1 << 17

CRAW_BATCH_DENOISE_TALL = 1 << 0
CRAW_BATCH_DENOISE_UNROLL = 1 << 1

CRAW_FRAME_FLAG_CLAMP_NEGATIVE = 1 << 0
CRAW_FRAME_FLAG_NO_GAMUT_COMPRESSION = 1 << 2
CRAW_FRAME_FLAG_HIGHLIGHT_RECONSTRUCTION = 1 << 3
CRAW_FRAME_FLAG_TEMPORAL_RESET = 1 << 4

CRAW_AWB_HIST_BINS = 128

EINVAL, ENOTSUP, EAGAIN, ENOMEM, ENODEV, EIO, ENOSPC = 22, 95, 11, 12, 19, 5, 28


class CRAWIlluminant(ctypes.Structure):
    _fields_ = [("cct", c_float), ("tint", c_float)]


class CRAWWbKnot(ctypes.Structure):
    _fields_ = [("cct", c_float), ("wb", c_float * 3)]


class CRAWMatrixKnot(ctypes.Structure):
    _fields_ = [("cct", c_float), ("m", (c_float * 3) * 3)]


class CRAWNoiseKnot(ctypes.Structure):
    _fields_ = [("iso", c_float), ("a", c_float), ("b", c_float)]


class CRAWNoiseModel(ctypes.Structure):
    _fields_ = [("a", c_float), ("b", c_float)]


class CRAWProfile(ctypes.Structure):
    _fields_ = [
        ("cfa", c_i32),
        ("black_level", c_float * 4),
        ("white_level", c_float * 4),
        ("as_shot_wb", c_float * 3),
        ("as_shot_matrix", (c_float * 3) * 3),
        ("as_shot_cct", c_float),
        ("gain", c_float),
        ("as_shot_ei", c_float),
        ("exposure_index", c_float),
        ("wb_knots", CRAWWbKnot * 16),
        ("nb_wb_knots", c_u32),
        ("matrix_knots", CRAWMatrixKnot * 16),
        ("nb_matrix_knots", c_u32),
        ("name", ctypes.c_char * 64),
        ("noise_knots", CRAWNoiseKnot * 16),
        ("nb_noise_knots", c_u32),
    ]


class CRAWColorParams(ctypes.Structure):
    _fields_ = [("wb", c_float * 3), ("cam_to_xyz", (c_float * 3) * 3), ("gain", c_float)]


class CRAWLensCorrection(ctypes.Structure):
    _fields_ = [("k", (c_float * 4) * 3), ("centre", c_float * 2), ("valid", ctypes.c_bool),
                ("reserved", ctypes.c_uint8 * (64 - 3 * 4 * 4 - 2 * 4 - 1))]


class CRAWPipelineParams(ctypes.Structure):
    _fields_ = [
        ("width", c_u32), ("height", c_u32),
        ("cfa", c_i32), ("debayer", c_i32), ("out_space", c_i32), ("out_format", c_i32),
        ("flags", c_u32), ("nb_awb_slots", c_u32),
        ("padding", ctypes.c_uint8 * (256 - 8 * 4)),
    ]


class CRAWFrameParams(ctypes.Structure):
    _fields_ = [
        ("profile", ctypes.POINTER(CRAWProfile)),
        ("wb_mode", c_i32),
        ("illuminant", CRAWIlluminant),
        ("exposure_stops", c_float),
        ("flags", c_u32),
        ("out_space", c_i32),
        ("out_space_set", ctypes.c_bool),
        ("denoise_strength", c_float),
        ("noise_override", CRAWNoiseModel),
        ("false_color_passes", c_u32),
        ("denoise_chroma_strength", c_float),
        ("lens", CRAWLensCorrection),
        ("defringe_strength", c_float),
        ("denoise_temporal_strength", c_float),
        ("denoise_temporal_replay", ctypes.c_bool),
        ("sharpen_radius", c_float),
        ("sharpen_iterations", c_u32),
        ("sharpen_contrast", c_float),
        ("rotation", c_float),
        ("padding", ctypes.c_uint8 * (256 - 4 * 4 - 8 - 12 - 64 - 8 - 4 - 4 - 4 - 4 - 4 - 12 - 4)),
    ]


class CRAWImage(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("pitch", ctypes.c_size_t), ("format", c_i32),
                ("width", c_u32), ("height", c_u32)]


class CRAWHipInitParams(ctypes.Structure):
    _fields_ = [("device", c_i32), ("padding", ctypes.c_uint8 * 252)]


class CRAWAwbHistogram(ctypes.Structure):
    _fields_ = [("bins", c_u32), ("log_min", c_float), ("log_max", c_float),
                ("count", ctypes.POINTER(c_u32)), ("weight", ctypes.POINTER(c_u32)),
                ("weight_bright", ctypes.POINTER(c_u32))]


class CRAWAwbParams(ctypes.Structure):
    _fields_ = [("band", c_float), ("sigma", c_float), ("min_zone_fraction", c_float),
                ("ambience", c_float), ("ambience_knee_cct", c_float), ("cct_min", c_float),
                ("cct_max", c_float), ("prefer_bright", ctypes.c_bool),
                ("bright_max_spread_mired", c_float)]


class CRAWAwbCandidate(ctypes.Structure):
    _fields_ = [("valid", ctypes.c_bool), ("illuminant", CRAWIlluminant),
                ("in_zone_fraction", c_float), ("spread_mired", c_float)]


CRAW_AWB_SOURCE_NONE, CRAW_AWB_SOURCE_ZONE, CRAW_AWB_SOURCE_BRIGHT = range(3)


class CRAWAwbEstimate(ctypes.Structure):
    _fields_ = [("valid", ctypes.c_bool), ("illuminant", CRAWIlluminant),
                ("confidence", c_float), ("in_zone_fraction", c_float),
                ("spread_mired", c_float), ("source", c_i32), ("zone", CRAWAwbCandidate),
                ("bright", CRAWAwbCandidate)]


class CRAWAwbSmootherParams(ctypes.Structure):
    _fields_ = [("time_constant", c_float), ("hysteresis_mired", c_float),
                ("max_slew_mired_per_s", c_float), ("min_confidence", c_float)]


class CRAWNoiseEstimateParams(ctypes.Structure):
    _fields_ = [("nb_bins", c_u32), ("min_bin_samples", c_u32), ("max_level", c_float),
                ("padding", ctypes.c_uint8 * (64 - 12))]


class CRAWNoiseEstimate(ctypes.Structure):
    _fields_ = [("model", CRAWNoiseModel), ("nb_bins", c_u32), ("nb_inliers", c_u32),
                ("fit_rms", c_float), ("read_noise_lsb", c_float),
                ("padding", ctypes.c_uint8 * (64 - 24))]


class CRAWStageTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("count", c_u32), ("ms", ctypes.c_double)]


class CRAWBatchParams(ctypes.Structure):
    _fields_ = [("max_frames", c_u32), ("max_width", c_u32), ("max_height", c_u32),
                ("flags", c_u32), ("max_pixels", ctypes.c_uint64), ("debayer", c_i32),
                ("out_space", c_i32), ("out_format", c_i32), ("use_graphs", c_u32),
                ("padding", ctypes.c_uint8 * (256 - 4 * 4 - 8 - 4 * 4))]


class CRAWBatchFrame(ctypes.Structure):
    _fields_ = [("src", CRAWImage), ("dst", CRAWImage),
                ("params", ctypes.POINTER(CRAWFrameParams)), ("state", ctypes.c_void_p)]


class CRAWServerParams(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("max_batch_pixels", ctypes.c_uint64),
                ("nb_slots", c_u32), ("max_batch_frames", c_u32), ("max_width", c_u32),
                ("max_height", c_u32), ("debayer", c_i32), ("out_space", c_i32),
                ("out_format", c_i32), ("flags", c_u32), ("use_graphs", c_u32),
                ("max_queue", c_u32), ("padding", ctypes.c_uint8 * (256 - 8 * 2 - 4 * 10))]


class CRAWServeRequest(ctypes.Structure):
    _fields_ = [("src", CRAWImage), ("dst", CRAWImage), ("params", CRAWFrameParams),
                ("stream", c_i32), ("reserved", c_u32), ("tag", ctypes.c_uint64),
                ("awb_histogram", ctypes.POINTER(c_u32))]


class CRAWServeCompletion(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("status", c_i32), ("stream", c_i32),
                ("submit_ms", ctypes.c_double), ("launch_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double), ("batch_frames", c_u32), ("slot", c_u32)]


P = ctypes.POINTER
vp = ctypes.c_void_p

# (name, restype, argtypes): every symbol include/craw/*.h declares.
SIGNATURES = [
    ("craw_context_create", c_i32, [P(vp), P(CRAWHipInitParams), vp]),
    ("craw_context_free", None, [P(vp)]),
    ("craw_log", None, [vp, c_i32, ctypes.c_char_p]),
    ("craw_version_int", c_u32, []),
    ("craw_pipeline_params_init", None, [P(CRAWPipelineParams)]),
    ("craw_frame_params_init", None, [P(CRAWFrameParams)]),
    ("craw_pipeline_create", c_i32, [vp, P(CRAWPipelineParams), P(vp)]),
    ("craw_pipeline_free", None, [P(vp)]),
    ("craw_pipeline_record", c_i32, [vp, vp, P(CRAWImage), P(CRAWImage), P(CRAWFrameParams)]),
    ("craw_develop_noalloc", c_i32, [vp, vp, P(CRAWImage), P(CRAWImage), P(CRAWFrameParams),
                                     c_i32, c_i32, vp]),
    ("craw_pipeline_get_awb_histogram", c_i32, [vp, P(CRAWAwbHistogram)]),
    ("craw_pipeline_get_awb_histogram_n", c_i32, [vp, c_u32, P(CRAWAwbHistogram)]),
    ("craw_pipeline_get_stage_times", c_i32, [vp, P(CRAWStageTime), c_u32]),
    ("craw_pipeline_get_stage_times_n", c_i32, [vp, c_u32, P(CRAWStageTime), c_u32]),
    ("craw_profile_init", None, [P(CRAWProfile)]),
    ("craw_profile_validate", c_i32, [P(CRAWProfile)]),
    ("craw_profile_eval", c_i32, [P(CRAWProfile), c_i32, P(CRAWIlluminant), P(CRAWColorParams)]),
    ("craw_profile_noise_eval", c_i32, [P(CRAWProfile), c_float, P(CRAWNoiseModel)]),
    ("craw_noise_model_eval", c_i32, [P(CRAWNoiseKnot), c_u32, c_float, P(CRAWNoiseModel)]),
    ("craw_colorspace_from_xyz_d65", c_i32, [c_i32, P(c_float)]),
    ("craw_colorspace_to_xyz_d65", c_i32, [c_i32, P(c_float)]),
    ("craw_planckian_xy", c_i32, [c_float, P(c_float)]),
    ("craw_bradford_adaptation", c_i32, [P(c_float), P(c_float), P(c_float)]),
    ("craw_awb_histogram_from_rgb", c_i32, [P(c_float), c_u32, c_u32, ctypes.c_size_t, c_i32,
                                             P(c_float), P(c_u32), P(c_u32), P(c_u32)]),
    ("craw_profile_neutral_to_illuminant", c_i32, [P(CRAWProfile), c_float, c_float,
                                                    P(CRAWIlluminant)]),
    ("craw_profile_illuminant_to_neutral", c_i32, [P(CRAWProfile), P(CRAWIlluminant),
                                                    P(c_float), P(c_float)]),
    ("craw_noise_estimate_params_init", None, [P(CRAWNoiseEstimateParams)]),
    ("craw_noise_estimate", c_i32, [P(ctypes.c_uint16), c_u32, c_u32, ctypes.c_size_t,
                                     P(CRAWProfile), P(CRAWNoiseEstimateParams),
                                     P(CRAWNoiseEstimate)]),
    ("craw_noise_estimate_image", c_i32, [vp, vp, P(CRAWImage), P(CRAWProfile),
                                           P(CRAWNoiseEstimateParams), P(CRAWNoiseEstimate)]),
    ("craw_awb_params_init", None, [P(CRAWAwbParams)]),
    ("craw_awb_estimate", c_i32, [P(CRAWProfile), P(CRAWAwbHistogram), P(CRAWAwbParams),
                                   P(CRAWAwbEstimate)]),
    ("craw_awb_smoother_params_init", None, [P(CRAWAwbSmootherParams)]),
    ("craw_awb_smoother_create", c_i32, [P(CRAWAwbSmootherParams), P(vp)]),
    ("craw_awb_smoother_free", None, [P(vp)]),
    ("craw_awb_smoother_reset", None, [vp]),
    ("craw_awb_smoother_set", None, [vp, ctypes.c_double, P(CRAWIlluminant)]),
    ("craw_awb_smoother_push", c_i32, [vp, ctypes.c_double, P(CRAWAwbEstimate), ctypes.c_bool,
                                        P(CRAWIlluminant)]),
    ("craw_batch_params_init", None, [P(CRAWBatchParams)]),
    ("craw_batch_create", c_i32, [vp, P(CRAWBatchParams), P(vp)]),
    ("craw_batch_free", None, [P(vp)]),
    ("craw_batch_record", c_i32, [vp, vp, P(CRAWBatchFrame), c_u32]),
    ("craw_batch_get_awb_histogram", c_i32, [vp, c_u32, P(CRAWAwbHistogram)]),
    ("craw_batch_get_stage_times", c_i32, [vp, P(CRAWStageTime), c_u32]),
    ("craw_batch_autotune", c_i32, [vp, vp, P(CRAWBatchFrame), c_u32, c_u32]),
    ("craw_batch_set_launch_shape", c_i32, [vp, P(c_i32), c_i32]),
    ("craw_batch_get_launch_shape", c_i32, [vp, P(c_i32), P(c_i32)]),
    ("craw_temporal_state_create", c_i32, [vp, c_u32, c_u32, P(vp)]),
    ("craw_temporal_state_free", None, [P(vp)]),
    ("craw_temporal_state_reset", None, [vp]),
    ("craw_server_params_init", None, [P(CRAWServerParams)]),
    ("craw_server_create", c_i32, [vp, P(CRAWServerParams), P(vp)]),
    ("craw_server_free", None, [P(vp)]),
    ("craw_server_open_stream", c_i32, [vp, c_u32, c_u32, P(c_i32)]),
    ("craw_server_close_stream", c_i32, [vp, c_i32]),
    ("craw_server_submit", c_i32, [vp, P(CRAWServeRequest)]),
    ("craw_server_submit_n", c_i32, [vp, P(CRAWServeRequest), c_u32]),
    ("craw_server_poll", c_i32, [vp, P(CRAWServeCompletion), c_u32, c_i32]),
    ("craw_server_drain", c_i32, [vp]),
    ("craw_server_get_memory_usage", ctypes.c_uint64, [vp]),
]

_lib = None


def lib():
    """libcraw.so with every declared signature bound. Raises if unbuilt."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcraw.so is not built (%s); run __graft_entry__.build()"
                               % LIB_PATH)
        l = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def declared_symbols():
    """The names include/craw/*.h declares with CRAW_API."""
    import re
    inc = os.path.join(os.path.dirname(HERE), "include", "craw")
    names = set()
    for fn in sorted(os.listdir(inc)):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            names.update(re.findall(r"CRAW_API\s+[\w\s\*]+?\b(craw_\w+)\s*\(", text))
    return names
