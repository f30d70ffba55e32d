// ag_fixed_dev.h -- device implementation of the fixed-point formats of fixedPoint.h.
//
// Same definitions as the host header, written with explicitly rounded HIP intrinsics so
// the result does not depend on contraction or fast-math settings:
//   encode: round-half-up of x * 2^frac formed in double, saturating at the code width;
//           NaN, zero and negatives encode to 0
//   decode: round-to-nearest-even conversion of the u64 sum, then an exact power-of-two scale
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

__device__ __forceinline__ uint64_t ag_dev_encode_scaled(double sph, double limit, uint64_t maxcode)
{
    if (!(sph >= 1.0))
        return 0ull;
    if (sph >= limit)
        return maxcode;
    return __double2ull_rz(sph);
}

__device__ __forceinline__ uint64_t ag_dev_double_to_fixed64(double x)
{
    double s = __dmul_rn(x, 4294967296.0);
    return ag_dev_encode_scaled(__dadd_rn(s, 0.5), 18446744073709551616.0, 0xFFFFFFFFFFFFFFFFull);
}

__device__ __forceinline__ double ag_dev_fixed64_to_double(uint64_t q)
{
    return __dmul_rn(__ull2double_rn(q), 1.0 / 4294967296.0);
}

// MULFixed64V1 of fixedPoint.h: bits [32, 96) of the exact 128-bit product
__device__ __forceinline__ uint64_t ag_dev_mul_fixed64(uint64_t a, uint64_t b)
{
    return (__umul64hi(a, b) << 32) | ((a * b) >> 32);
}

__device__ __forceinline__ uint32_t ag_dev_float_to_fixed32(float x)
{
    double s = __dmul_rn((double)x, 65536.0);
    return (uint32_t)ag_dev_encode_scaled(__dadd_rn(s, 0.5), 4294967296.0, 0xFFFFFFFFull);
}

__device__ __forceinline__ float ag_dev_fixed32_to_float(uint64_t q)
{
    return __fmul_rn(__ull2float_rn(q), 1.0f / 65536.0f);
}

__device__ __forceinline__ uint16_t ag_dev_float_to_fixed16(float x)
{
    double s = __dmul_rn((double)x, 256.0);
    return (uint16_t)ag_dev_encode_scaled(__dadd_rn(s, 0.5), 65536.0, 0xFFFFull);
}

__device__ __forceinline__ float ag_dev_fixed16_to_float(uint64_t q)
{
    return __fmul_rn(__ull2float_rn(q), 1.0f / 256.0f);
}

__device__ __forceinline__ uint8_t ag_dev_float_to_fixed8(float x)
{
    double s = __dmul_rn((double)x, 16.0);
    return (uint8_t)ag_dev_encode_scaled(__dadd_rn(s, 0.5), 256.0, 0xFFull);
}

__device__ __forceinline__ float ag_dev_fixed8_to_float(uint64_t q)
{
    return __fmul_rn(__ull2float_rn(q), 1.0f / 16.0f);
}
