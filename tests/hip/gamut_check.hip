/*
 * Test support: kgamut (libcraw_amd/csrc/k_common.h), the colour stage's
 * gamut compression, over arrays of AP1 triples, for
 * tests/test_gpu_gamut.py. Built with the library's own flags by
 * tests/hip/Makefile; not part of the library.
 */
#include "k_common.h"

__global__ void k_gamut_check(float *__restrict__ rgb, size_t count, KColor col)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;

    if (i < count) {
        float r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];

        kgamut(col, r, g, b);
        rgb[3 * i] = r;
        rgb[3 * i + 1] = g;
        rgb[3 * i + 2] = b;
    }
}

/* `rgb`: device pointer to `count` triples, compressed in place with the
 * constants given; `screen` 0 turns kgamut's screen off (its bounds -inf),
 * `float_math` selects the precision form's float pow. 0 or a hipError_t. */
extern "C" int gamut_check(float *rgb, size_t count, const float thr[3], const float scl[3],
                           float pwr, float inv_pwr, int screen, int float_math)
{
    const unsigned blocks = (unsigned)((count + 255) / 256);
    KColor col = {};

    if (!count)
        return 0;
    for (int i = 0; i < 3; i++) {
        col.thr[i] = thr[i];
        col.scl[i] = scl[i];
        col.thr_lo[i] = screen ? kgamut_screen_bound(thr[i]) : -__builtin_inff();
    }
    col.pwr = pwr;
    col.inv_pwr = inv_pwr;
    col.flags = KCOLOR_GAMUT_COMPRESS | (float_math ? KCOLOR_FLOAT_MATH : 0u);
    hipLaunchKernelGGL(k_gamut_check, dim3(blocks), dim3(256), 0, 0, rgb, count, col);
    if (hipGetLastError() != hipSuccess)
        return 1;
    return (int)hipDeviceSynchronize();
}

/* The screen's bound on the host, for the test to check its margin. */
extern "C" float gamut_screen_bound(float thr)
{
    return kgamut_screen_bound(thr);
}
