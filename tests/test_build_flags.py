"""The library's HIP code is built without packed FP32 instructions.

libcraw_amd/csrc/Makefile passes `-Xclang -target-feature -Xclang
-packed-fp32-ops` to hipcc. The gfx950 device pass honours the feature; only
hipcc's host pass, which does not know it, prints that it ignores it. These
tests show both halves on the CPU: a probe kernel of float2 arithmetic
compiles to v_pk_*_f32 instructions without the flag and to none with it, and
the gfx950 code objects of the library's own built objects hold none.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "libcraw_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
LLVM = "/opt/rocm/llvm/bin"
FLAG = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
PACKED = re.compile(r"\bv_pk_(fma|mul|add)_f32\b")

PROBE = r"""
#include <hip/hip_runtime.h>
extern "C" __global__ void probe(const float2 *a, const float2 *b, float2 *c, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float2 x = a[i], y = b[i], r;
        r.x = x.x * y.x + x.y;
        r.y = x.y * y.y + x.x;
        r.x = r.x * 3.0f + y.y;
        r.y = r.y * 3.0f + y.x;
        c[i] = r;
    }
}
"""

needs_hipcc = pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")


def _device_asm(tmp_path, extra):
    src = tmp_path / "probe.hip"
    src.write_text(PROBE)
    out = tmp_path / ("probe%d.s" % len(extra))
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-ffp-contract=fast",
                    "--cuda-device-only", "-S", str(src), "-o", str(out)] + extra,
                   check=True, capture_output=True)
    return out.read_text()


def test_makefile_passes_the_flag():
    mk = open(os.path.join(CSRC, "Makefile")).read()
    hipflags = mk[mk.index("HIPFLAGS :="):mk.index("HIPFLAGS +=")]
    assert " ".join(FLAG) in " ".join(hipflags.split())


@needs_hipcc
def test_flag_removes_packed_fp32_from_the_device_pass(tmp_path):
    assert PACKED.search(_device_asm(tmp_path, [])), \
        "the probe no longer produces packed FP32 instructions; pick another probe"
    assert not PACKED.search(_device_asm(tmp_path, FLAG))


@needs_hipcc
def test_library_code_objects_hold_no_packed_fp32(tmp_path):
    build = os.path.join(CSRC, "_build")
    objs = sorted(f for f in os.listdir(build) if f.startswith("k_") or f == "noalloc.o") \
        if os.path.isdir(build) else []
    objs = [o for o in objs if o.endswith(".o")]
    if not objs:
        pytest.skip("the library is not built here")
    coded = 0
    for obj in objs:
        fb = tmp_path / (obj + ".fatbin")
        co = tmp_path / (obj + ".co")
        subprocess.run([os.path.join(LLVM, "llvm-objcopy"), "--dump-section",
                        ".hip_fatbin=%s" % fb, os.path.join(build, obj)],
                       check=True, capture_output=True)
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o",
                        "--input=%s" % fb, "--targets=hipv4-amdgcn-amd-amdhsa--gfx950",
                        "--output=%s" % co], check=True, capture_output=True)
        dis = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", str(co)],
                             check=True, capture_output=True, text=True).stdout
        coded += "v_" in dis
        hits = PACKED.findall(dis)
        assert not hits, "%s: %d packed FP32 instructions" % (obj, len(hits))
    assert coded >= len(objs) - 1, "code objects without instructions"
