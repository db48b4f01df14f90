"""The kernels' table-driven exp equals the oracle's exp on every argument.

The exact filters take their weights through craw_dm_expf_negmax0_t1024 /
_t256 and craw_dm_expf_nonpos_t1024 (libcraw_amd/csrc/detmath.h), which
read 2^(j/N) from a table; the oracle keeps craw_dm_exp_d, whose float
rounding is the correctly rounded exp. These tests compile a C check
against the same header and compare the two on every float argument of
the clamped range [-200, 0] (1,128,792,065 values), and show that the check
fails on a planted mistake: the table computed in single precision, or the
polynomial one degree short. (A single entry one ulp off moves no float
result: that is what test_table_entries_are_correctly_rounded checks.)
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "libcraw_amd", "csrc")

CHECK = r"""
#include <stdio.h>
#include <stdint.h>
#include <string.h>
#include "detmath.h"
#include "detmath_tab.h"

#ifdef PLANT_FLOAT_TABLE
static double tab[N_TAB];
#endif

#ifdef PLANT_DEGREE
/* the N = 1024 form with exp(r) - 1 cut to degree 2 */
static double exp_short(double x, const double *t)
{
    const double shift = 0x1.8p52;
    double tt = __builtin_fma(x, 0x1.71547652b82fep+10, shift);
    double m = tt - shift;
    double r = __builtin_fma(-m, 0x1.62e42fee00000p-11, x);
    uint32_t lo;
    double T, c;

    r = __builtin_fma(-m, 0x1.a39ef35793c76p-43, r);
    lo = (uint32_t)craw_dm_bits(tt);
    T = t[lo & 1023u];
    c = __builtin_fma(r, 0.5, 1.0) * r;
    T = __builtin_fma(T, c, T);
    return T * craw_dm_from_bits((uint64_t)(uint32_t)(((int32_t)lo >> 10) + 1023) << 52);
}
#endif

int main(void)
{
    const float f200 = 200.0f;
    uint32_t hi;
    long bad = 0, nonpos_bad = 0;
#if N_TAB == 256
    const double *t = craw_dm_exp2_tab256;
#else
    const double *t = craw_dm_exp2_tab1024;
#endif
#ifdef PLANT_FLOAT_TABLE
    for (int j = 0; j < N_TAB; j++)
        tab[j] = (double)(float)t[j];
    t = tab;
#endif
    memcpy(&hi, &f200, 4);
#pragma omp parallel for reduction(+:bad, nonpos_bad) schedule(static, 1 << 16)
    for (long u = 0; u <= (long)hi; u++) {
        uint32_t b = (uint32_t)u;
        float c, ref, got;

        memcpy(&c, &b, 4);
        ref = craw_dm_expf_negmax0(c);
#if defined(PLANT_DEGREE)
        got = (float)exp_short(-(double)c, t);
#elif N_TAB == 256
        got = craw_dm_expf_negmax0_t256(c, t);
#else
        got = craw_dm_expf_negmax0_t1024(c, t);
        {
            float g2 = craw_dm_expf_nonpos_t1024(-c, t), r2 = craw_dm_expf_nonpos(-c);
            nonpos_bad += memcmp(&g2, &r2, 4) != 0;
        }
#endif
        bad += memcmp(&ref, &got, 4) != 0;
    }
    printf("%ld %ld %u\n", bad, nonpos_bad, hi + 1);
    return 0;
}
"""


def _run(tmp_path, n, *defs):
    src = tmp_path / "check.c"
    src.write_text(CHECK)
    exe = tmp_path / ("check_%d_%d" % (n, len(defs)))
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-mfma", "-ffp-contract=off", "-fopenmp",
                    "-I" + CSRC, "-DN_TAB=%d" % n] + ["-D" + d for d in defs] +
                   [str(src), "-o", str(exe), "-lm"], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    return int(out[0]), int(out[1]), int(out[2])


@pytest.mark.parametrize("n", [256, 1024])
def test_table_exp_equals_the_oracle_exp_everywhere(tmp_path, n):
    bad, nonpos_bad, count = _run(tmp_path, n)
    assert count == 1128792065
    assert bad == 0
    assert nonpos_bad == 0


@pytest.mark.parametrize("n", [256, 1024])
def test_check_catches_a_single_precision_table(tmp_path, n):
    bad, _, _ = _run(tmp_path, n, "PLANT_FLOAT_TABLE")
    assert bad > 0


def test_check_catches_a_short_polynomial(tmp_path):
    bad, _, _ = _run(tmp_path, 1024, "PLANT_DEGREE")
    assert bad > 0


def test_table_entries_are_correctly_rounded():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.prec = 200
    text = open(os.path.join(CSRC, "detmath_tab.h")).read()
    for name, n in (("craw_dm_exp2_tab256", 256), ("craw_dm_exp2_tab1024", 1024)):
        body = text[text.index(name + "[%d] = {" % n):]
        body = body[body.index("{") + 1:body.index("}")]
        vals = [float.fromhex(v) for v in body.replace("\n", " ").split(",") if v.strip()]
        assert len(vals) == n
        for j, v in enumerate(vals):
            exact = mpmath.mpf(2) ** (mpmath.mpf(j) / n)
            ulp = mpmath.mpf(2) ** (mpmath.floor(mpmath.log(v, 2)) - 52)
            assert abs(mpmath.mpf(v) - exact) <= ulp / 2, (name, j)
