/*
 * fixedPoint.h -- fixed-point formats of the integer PageRank / SpMV pull paths.
 *
 * The reference calls FloatToFixed32 / Fixed32ToFloat / DoubleToFixed64 / Fixed64ToDouble /
 * FloatToFixed16 / Fixed16ToFloat / FloatToFixed8 / Fixed8ToFloat
 * (02_capi_graph/src/algorithms/capi/pageRank.c:1373,1400,1553,1581,1730,1758,1911,1939;
 * SPMV.c:924-926,953,982) from OpenGraph's fixedPoint.h, which is absent from the
 * reference tree. The only scale the tree fixes is SCALEF = 16 for the 32-bit SpMV
 * datapath (03_capi_integration/accelerator_rtl/cu_control/cu_SPMV/CSR/PULL/FixedPoint/
 * pkg/globals_cu_pkg.sv:37), i.e. Q16.16. The formats here are this library's own:
 *
 *   width  fraction bits  encode (value -> code)                  decode (u64 sum -> value)
 *   64     32             round-half-up of x*2^32, saturating      (double)q * 2^-32
 *   32     16             round-half-up of x*2^16, saturating      (float)q  * 2^-16
 *   16      8             round-half-up of x*2^8,  saturating      (float)q  * 2^-8
 *    8      4             round-half-up of x*2^4,  saturating      (float)q  * 2^-4
 *
 * Encoding maps NaN, zero and negative inputs to 0. The scaled value is formed in double
 * (exact for a float input), so "round-half-up" is exact. Decoding converts the (possibly
 * wide) integer sum with round-to-nearest-even and then scales by a power of two, which
 * is exact. The device kernels implement the same functions independently and the GPU
 * parity tests compare the two bit for bit.
 */
#ifndef FIXEDPOINT_H
#define FIXEDPOINT_H

#include <stdint.h>

#define AG_FIX64_FRAC 32
#define AG_FIX32_FRAC 16
#define AG_FIX16_FRAC 8
#define AG_FIX8_FRAC 4

static inline uint64_t ag_encode_scaled_u(double scaled_plus_half, double limit, uint64_t maxcode)
{
    if (!(scaled_plus_half >= 1.0))
        return 0;
    if (scaled_plus_half >= limit)
        return maxcode;
    return (uint64_t)scaled_plus_half;
}

static inline uint64_t DoubleToFixed64(double x)
{
    double s = x * 4294967296.0;
    return ag_encode_scaled_u(s + 0.5, 18446744073709551616.0, UINT64_MAX);
}

static inline double Fixed64ToDouble(uint64_t q)
{
    return (double)q * (1.0 / 4294967296.0);
}

/* Q32.32 product used by the fixed-point SpMV push (SPMV.c:1092); OpenGraph's definition is
 * absent, so here it is the exact 128-bit product shifted right by 32, truncated to 64 bits. */
static inline uint64_t MULFixed64V1(uint64_t a, uint64_t b)
{
    return (uint64_t)(((unsigned __int128)a * b) >> 32);
}

static inline uint32_t FloatToFixed32(float x)
{
    double s = (double)x * 65536.0;
    return (uint32_t)ag_encode_scaled_u(s + 0.5, 4294967296.0, UINT32_MAX);
}

static inline float Fixed32ToFloat(uint64_t q)
{
    return (float)q * (1.0f / 65536.0f);
}

static inline uint16_t FloatToFixed16(float x)
{
    double s = (double)x * 256.0;
    return (uint16_t)ag_encode_scaled_u(s + 0.5, 65536.0, UINT16_MAX);
}

static inline float Fixed16ToFloat(uint64_t q)
{
    return (float)q * (1.0f / 256.0f);
}

static inline uint8_t FloatToFixed8(float x)
{
    double s = (double)x * 16.0;
    return (uint8_t)ag_encode_scaled_u(s + 0.5, 256.0, UINT8_MAX);
}

static inline float Fixed8ToFloat(uint64_t q)
{
    return (float)q * (1.0f / 16.0f);
}

#endif
