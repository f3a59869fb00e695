// Device-side helpers shared by the HIP kernels.  Every kernel is compiled with
// -ffp-contract=off and f32 denormal flushing, so the float/double operation
// sequence below is exactly the one written (no implicit contraction), and
// f32 denormals flush like the reference process's FTZ/DAZ mode.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

#include "internal.hpp"

namespace asz {
namespace dev {

constexpr int kWave = 64;

__device__ __forceinline__ float daz(float x) { return fabsf(x) < FLT_MIN ? copysignf(0.0f, x) : x; }

// GCC -Ofast expands llround(double) inline: trunc(y + copysign(pred(0.5), y)),
// then cvttsd2si (out of range -> INT64_MIN, whose low 32 bits are 0).
__device__ __forceinline__ int32_t lround_q(double y) {
    double s = y + copysign(0.49999999999999994, y);
    if (!(fabs(s) < 9.223372036854775808e18)) return 0;
    return (int32_t)(int64_t)s;
}
__device__ __forceinline__ int64_t lround_f64(float y) {
    float s = y + copysignf(0.49999997f, y);
    if (!(fabsf(s) < 9.223372036854775808e18f)) return INT64_MIN;
    return (int64_t)s;
}

// The reference predictor with GCC's contraction of compress.tpp:9-34
// (identical in decompress.tpp:8-33).  r points at the target in the
// reconstruction; dist is the element distance of one half stride.
template <bool QUIN>
__device__ __forceinline__ float predict(const float* __restrict__ r, uint32_t i, uint32_t h, uint32_t right,
                                         int64_t dist, float t) {
    if (QUIN && i >= 5 * h && i + 5 * h < right) {
        float s1 = r[-dist] + r[dist];
        float s3 = r[-3 * dist] + r[3 * dist];
        float s5 = r[-5 * dist] + r[5 * dist];
        double acc = (double)s3 * (-25.0 / 256);
        acc = __fma_rn((double)s1, 150.0 / 256, acc);
        acc = __fma_rn((double)s5, 3.0 / 256, acc);
        return (float)acc;
    }
    float a = r[-dist];
    if (i <= h) {
        if (i + h >= right) return a;
        float b = r[dist];
        if (i + 3 * h >= right) return (a + b) * 0.5f;
        float e = r[3 * dist];
        return (__fmaf_rn(a, 3.0f, b * 6.0f) - e) * 0.125f;
    }
    float c = r[-3 * dist];
    if (i + h >= right) return __fmaf_rn(a, 3.0f, -c) * 0.5f;
    float b = r[dist];
    if (i + 3 * h >= right) return __fmaf_rn(b, 3.0f, __fmaf_rn(a, 6.0f, -c)) * 0.125f;
    float e = r[3 * dist];
    double td = (double)t;
    float p0 = (float)((0.25 - td) * (double)(e + c));
    return (float)__fma_rn(td + 0.25, (double)(a + b), (double)p0);
}

// Quantizer (compress.tpp:35).  The reference's hot loop inlines the axis-0
// call with a hoisted reciprocal, q = lround((0.5/eb)*(p-x)); the axis-1/2
// loops call the out-of-line copy, q = lround(((p-x)*0.5)/eb).
template <int AX>
__device__ __forceinline__ int32_t quantize(float p, float x, const RoundDesc& R) {
    double diff = (double)p - (double)daz(x);
    if (AX == 0) return lround_q(R.inv_half_eb * diff);
    return lround_q((diff * 0.5) / R.eb_d);
}
__device__ __forceinline__ float reconstruct(float p, int32_t q, double neg_two_eb) {
    return (float)__fma_rn(neg_two_eb, (double)q, (double)p);
}

// ---------------------------------------------------------------- reductions

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { T w = __shfl_xor(v, o, kWave); v = w > v ? w : v; }
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { T w = __shfl_xor(v, o, kWave); v = w < v ? w : v; }
    return v;
}

__device__ __forceinline__ void atomic_max_pos_double(double* p, double v) {
    atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

using asz::ArrayAcc;
using asz::ErrAcc;

}  // namespace dev
}  // namespace asz
