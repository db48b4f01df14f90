/*
 * Test support: kdiv_rn (libcraw_amd/csrc/k_common.h) and the compiler's
 * correctly rounded division over arrays of operands, for
 * tests/test_gpu_kdiv.py. Built with the library's own flags by
 * tests/hip/Makefile; not part of the library.
 */
#include "k_common.h"

__global__ void k_kdiv_check(const float *__restrict__ n, const float *__restrict__ d,
                             float *__restrict__ q_rn, float *__restrict__ q_div, size_t count)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;

    if (i < count) {
        q_rn[i] = kdiv_rn(n[i], d[i]);
        q_div[i] = n[i] / d[i];
    }
}

/* Device pointers of `count` floats each; 0 or a hipError_t. */
extern "C" int kdiv_check(const float *n, const float *d, float *q_rn, float *q_div,
                          size_t count)
{
    const unsigned blocks = (unsigned)((count + 255) / 256);

    if (!count)
        return 0;
    hipLaunchKernelGGL(k_kdiv_check, dim3(blocks), dim3(256), 0, 0, n, d, q_rn, q_div, count);
    if (hipGetLastError() != hipSuccess)
        return 1;
    return (int)hipDeviceSynchronize();
}
