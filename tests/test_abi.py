"""The C ABI boundary, without a GPU: libcraw.so loads, exports every
symbol include/craw/*.h declares, its structs have the reference's sizes,
and its host-only entry points and error paths behave."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from libcraw_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    declared = capi.declared_symbols()
    assert len(declared) >= 20
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, "not exported: %s" % missing
    bound = {n for n, _, _ in capi.SIGNATURES}
    assert declared == bound, "ctypes mirror out of date: %s" % (declared ^ bound)


def test_exports_nothing_undeclared():
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    craw = {s for s in exported if s.startswith("craw_")}
    assert craw == capi.declared_symbols()


def test_struct_sizes_match_the_reference(tmp_path):
    # lib/pipeline.c:252-253 of the reference asserts 256 and 264.
    assert ctypes.sizeof(capi.CRAWPipelineParams) == 256
    assert ctypes.sizeof(capi.CRAWFrameParams) == 264
    assert ctypes.sizeof(capi.CRAWProfile) == 1264
    assert ctypes.sizeof(capi.CRAWLensCorrection) == 64
    c = str(tmp_path / "abi_sizes.c")
    exe = str(tmp_path / "abi_sizes")
    with open(c, "w") as f:
        f.write('#include <craw/serve.h>\n#include <stdio.h>\n#include <stddef.h>\n'
                'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", '
                'sizeof(CRAWPipelineParams), sizeof(CRAWFrameParams), sizeof(CRAWProfile), '
                'sizeof(CRAWImage), sizeof(CRAWContextOptions), '
                'offsetof(CRAWFrameParams, lens), offsetof(CRAWFrameParams, rotation), '
                'sizeof(CRAWBatchParams), sizeof(CRAWBatchFrame), sizeof(CRAWServerParams), '
                'sizeof(CRAWServeRequest), sizeof(CRAWServeCompletion));}\n')
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
    sizes = [int(x) for x in subprocess.run([exe], capture_output=True, text=True,
                                            check=True).stdout.split()]
    assert sizes == [256, 264, 1264, ctypes.sizeof(capi.CRAWImage), 1024,
                     capi.CRAWFrameParams.lens.offset, capi.CRAWFrameParams.rotation.offset,
                     256, ctypes.sizeof(capi.CRAWBatchFrame), 256,
                     ctypes.sizeof(capi.CRAWServeRequest), ctypes.sizeof(capi.CRAWServeCompletion)]


def test_headers_compile_as_c_and_cpp(tmp_path):
    for compiler, ext in (("gcc", "c"), ("g++", "cpp")):
        path = str(tmp_path / ("craw_hdr_test.%s" % ext))
        with open(path, "w") as f:
            f.write("#include <craw/pipeline.h>\n#include <craw/batch.h>\n"
                    "#include <craw/serve.h>\n"
                    "int main(void){return 0;}\n")
        subprocess.run([compiler, "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-c", path, "-o", str(tmp_path / "craw_hdr_test.o")], check=True)


def test_defaults_and_host_entry_points():
    lib = capi.lib()
    pp = capi.CRAWPipelineParams()
    lib.craw_pipeline_params_init(ctypes.byref(pp))
    assert (pp.cfa, pp.debayer, pp.out_space, pp.out_format) == (
        capi.CRAW_CFA_RGGB, capi.CRAW_DEBAYER_RCD, capi.CRAW_CS_XYZ_D65,
        capi.CRAW_FORMAT_R16G16B16A16_SFLOAT)
    fp = capi.CRAWFrameParams()
    lib.craw_frame_params_init(ctypes.byref(fp))
    assert fp.wb_mode == capi.CRAW_WB_AS_SHOT and not fp.profile
    prof = capi.CRAWProfile()
    lib.craw_profile_init(ctypes.byref(prof))
    assert lib.craw_profile_validate(ctypes.byref(prof)) == 0
    prof.white_level[2] = prof.black_level[2]
    assert lib.craw_profile_validate(ctypes.byref(prof)) == -capi.EINVAL
    assert lib.craw_version_int() == (0 << 16) | (1 << 8)


def test_context_without_gpu_fails_cleanly():
    from libcraw_amd import hip
    if hip.device_count() > 0:
        pytest.skip("a GPU is present")
    lib = capi.lib()
    h = ctypes.c_void_p()
    init = capi.CRAWHipInitParams()
    assert lib.craw_context_create(ctypes.byref(h), ctypes.byref(init), None) == -capi.ENODEV
    assert not h.value


def test_null_arguments():
    lib = capi.lib()
    assert lib.craw_pipeline_create(None, None, None) == -capi.EINVAL
    assert lib.craw_pipeline_record(None, None, None, None, None) == -capi.EINVAL
    lib.craw_pipeline_free(None)
    lib.craw_context_free(None)
    h = capi.CRAWAwbHistogram()
    assert lib.craw_pipeline_get_awb_histogram(None, ctypes.byref(h)) == -capi.EINVAL
    assert lib.craw_pipeline_get_stage_times(None, None, 0) == -capi.EINVAL


def test_bench_driver_built_and_parses_arguments():
    """The command-line bench driver is built beside the library and refuses
    bad arguments with its usage before it touches a device."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "libcraw_amd", "bench", "craw_bench")
    assert os.path.exists(exe), "craw_bench is not built (make -C libcraw_amd/csrc bench)"
    for args in (["--no-such-flag"], ["--width", "7"], ["--frames", "0"]):
        r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=30)
        assert r.returncode == 2 and "usage: craw_bench" in r.stderr, (args, r.stderr)
