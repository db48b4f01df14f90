"""libcraw for MI355X.

A from-scratch MI355X (gfx950) implementation of cyanreg/libcraw's raw
development path behind libcraw's own C API (include/craw/*.h), built as
libcraw.so in this directory. This package holds the library's sources
(csrc/) and a thin Python mirror of the API:

    capi      ctypes structs and signatures of include/craw/*.h
    pipeline  Context / Pipeline / frame parameters over the C API
    hip       device buffers, streams and events for handing the C API
              device pointers and a hipStream_t
    scene     seeded synthetic RGGB frames and the synthetic profile
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def build(jobs=8):
    """Compile libcraw.so for gfx950 and the bench driver
    (bench/craw_bench) linked against it; recompiles what changed."""
    subprocess.check_call(["make", "-C", os.path.join(HERE, "csrc"), "-j%d" % jobs, "all",
                           "bench"])
    return os.path.join(HERE, "libcraw.so")
