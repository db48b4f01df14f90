"""The command-line bench driver (libcraw_amd/bench/craw_bench, after
tools/cineraw.c's --bench, cineraw.c:1156-1290 of the reference): a raw
little-endian uint16 frame dump read from a file and developed through the
C API gives the same image, bit for bit, as the same frame recorded through
the library from Python with the same parameters; its pass table names
every pass with its dispatches and positive times; its synthetic frame and
its device noise fit run."""
import os
import re
import subprocess

import numpy as np
import pytest

from libcraw_amd import capi, scene
from libcraw_amd.pipeline import frame_params
from gpu_util import Runner, assert_bit_exact

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "libcraw_amd", "bench", "craw_bench")


def run_bench(*args):
    if not os.path.exists(BENCH):
        pytest.fail("craw_bench is not built (make -C libcraw_amd/csrc bench)")
    r = subprocess.run([BENCH, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def pass_table(text):
    """{pass: (dispatches, avg ms, median ms)} from the driver's table."""
    rows = {}
    for line in text.splitlines():
        m = re.match(r"^  (\w+)\s+(\d+)\s+([\d.]+)\s+([\d.]+)$", line)
        if m:
            rows[m.group(1)] = (int(m.group(2)), float(m.group(3)), float(m.group(4)))
    return rows


@pytest.mark.parametrize("debayer,half", [("rcd", False), ("mhc", True)])
def test_raw_dump_matches_library(gpu_available, tmp_path, debayer, half):
    w, h = 512, 384
    mosaic = scene.make_frame(w, h, seed=5, a=2e-4, b=4e-6)
    raw, out = tmp_path / "frame.raw", tmp_path / "frame.rgba"
    mosaic.astype("<u2").tofile(raw)
    args = ["--raw", raw, "--width", w, "--height", h, "--frames", 6, "--warmup", 2,
            "--debayer", debayer, "--space", "709", "--out", out]
    text = run_bench(*(args + (["--half"] if half else [])))

    run = Runner(w, h, debayer=capi.CRAW_DEBAYER_RCD if debayer == "rcd"
                 else capi.CRAW_DEBAYER_MHC, out_space=capi.CRAW_CS_BT709_LINEAR,
                 flags=capi.CRAW_PIPELINE_FLAG_DENOISE | capi.CRAW_PIPELINE_FLAG_CHROMA_DENOISE,
                 half=half)
    fp = frame_params(scene.make_profile(), denoise_strength=1.0,
                      denoise_chroma_strength=0.5, noise_override=(2e-4, 4e-6))
    assert run.run(mosaic, fp) == 0
    ref = run.output()
    if half:
        got = np.fromfile(out, np.float16).astype(np.float32).reshape(h, w, 4)
    else:
        got = np.fromfile(out, np.float32).reshape(h, w, 4)
    assert_bit_exact(got, ref, "craw_bench --out")

    rows = pass_table(text)
    assert "frames back to back" in text
    assert rows, text
    for name, (count, avg, med) in rows.items():
        assert count >= 1 and avg > 0 and med > 0, (name, rows[name])
    assert any(k.startswith("denoise") for k in rows), rows


def test_synthetic_frame_and_noise_fit(gpu_available):
    text = run_bench("--width", 2048, "--height", 1536, "--frames", 4, "--warmup", 1,
                     "--estimate-noise", "--awb", "--temporal", 0.5, "--cold")
    m = re.search(r"noise model: a ([\d.e+-]+) b ([\d.e+-]+)", text)
    assert m, text
    # the synthetic frame's noise is 2e-4 u + 4e-6 in its own units
    assert 1e-4 < float(m.group(1)) < 4e-4 and 0 < float(m.group(2)) < 2e-5, text
    rows = pass_table(text)
    assert "awb_hist" in rows or any("awb" in k for k in rows), rows
    assert any("temporal" in k or "denoise" in k for k in rows), rows
