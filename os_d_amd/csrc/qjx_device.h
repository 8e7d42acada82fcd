// qjx_device.h -- device helpers shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qjx_internal.h"

namespace qjx {

// MCG59 (MKL VSL_BRNG_MCG59): x <- a x mod 2^59, a = 13^13; output x * 2^-59.
// The per-trajectory leap-frog stream uses the multiplier a^mns.
__host__ __device__ inline uint64_t mcg_mul(uint64_t x, uint64_t y)
{
    return (x * y) & ((1ULL << 59) - 1ULL);
}

__host__ __device__ inline uint64_t mcg_pow(uint64_t a, uint64_t e)
{
    uint64_t r = 1;
    while (e) {
        if (e & 1ULL) r = mcg_mul(r, a);
        a = mcg_mul(a, a);
        e >>= 1;
    }
    return r;
}

__device__ inline double mcg_next(uint64_t &x, uint64_t stride)
{
    const double u = (double)x * 0x1p-59;
    x = mcg_mul(x, stride);
    return u;
}

__device__ inline double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace qjx
