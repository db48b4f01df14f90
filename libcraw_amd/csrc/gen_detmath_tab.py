"""Writes detmath_tab.h: 2^(j/N) rounded to the nearest double, for the
table-driven exp of detmath.h (N = 256 and 1024).

    python libcraw_amd/csrc/gen_detmath_tab.py

mpmath evaluates each entry at 200 bits; converting an mpf to a Python float
rounds to nearest, so every entry is the correctly rounded double.
"""
import os

import mpmath

mpmath.mp.prec = 200
HERE = os.path.dirname(os.path.abspath(__file__))


def table(n):
    return [float(mpmath.mpf(2) ** (mpmath.mpf(j) / n)) for j in range(n)]


def emit(f, name, n):
    f.write("CRAW_DM_TABLE %s[%d] = {\n" % (name, n))
    vals = table(n)
    rows = [", ".join(v.hex() for v in vals[i:i + 3]) for i in range(0, n, 3)]
    f.write(",\n".join("    " + r for r in rows) + "\n")
    f.write("};\n\n")


def main():
    path = os.path.join(HERE, "detmath_tab.h")
    with open(path, "w") as f:
        f.write("/* 2^(j/N) rounded to the nearest double, j = 0 .. N - 1: the tables of\n"
                " * detmath.h's table-driven exp. Written by gen_detmath_tab.py; do not edit. */\n"
                "#ifndef CRAW_DETMATH_TAB_H\n#define CRAW_DETMATH_TAB_H\n\n"
                "#if defined(__HIPCC__)\n#define CRAW_DM_TABLE static __device__ const double\n"
                "#else\n#define CRAW_DM_TABLE static const double\n#endif\n\n")
        emit(f, "craw_dm_exp2_tab256", 256)
        emit(f, "craw_dm_exp2_tab1024", 1024)
        f.write("#endif /* CRAW_DETMATH_TAB_H */\n")


if __name__ == "__main__":
    main()
