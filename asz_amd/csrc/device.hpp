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

// Global-memory view of a pointer whose address space the compiler cannot
// infer (one loaded from a descriptor in memory, or merged with such a load):
// accesses through it compile to global_* instead of flat_* instructions,
// which the LDS wait counter would also track.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T* p) {
    return (gptr<T>)p;
}

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
// (identical in decompress.tpp:8-33), over a register window of coarse values
// w[m] = value at i + (2m - 5) h, m = 0..5 (w[0] = i-5h ... w[5] = i+5h).  i is
// the target's coordinate along the round axis and right the end of its
// block's extent (min((b+1) B + 1, D)); entries outside what the branch taken
// reads may hold anything.
template <bool QUIN>
__device__ __forceinline__ float predict_w(uint32_t i, uint32_t h, uint32_t right, const float w[6], float t) {
    if (QUIN && i >= 5 * h && i + 5 * h < right) {
        float s1 = w[2] + w[3];
        float s3 = w[1] + w[4];
        float s5 = w[0] + w[5];
        double acc = (double)s3 * (-25.0 / 256);
        acc = __fma_rn((double)s1, 150.0 / 256, acc);
        acc = __fma_rn((double)s5, 3.0 / 256, acc);
        return (float)acc;
    }
    const float a = w[2];
    if (i <= h) {
        if (i + h >= right) return a;
        const float b = w[3];
        if (i + 3 * h >= right) return (a + b) * 0.5f;
        return (__fmaf_rn(a, 3.0f, b * 6.0f) - w[4]) * 0.125f;
    }
    const float c = w[1];
    if (i + h >= right) return __fmaf_rn(a, 3.0f, -c) * 0.5f;
    const float b = w[3];
    if (i + 3 * h >= right) return __fmaf_rn(b, 3.0f, __fmaf_rn(a, 6.0f, -c)) * 0.125f;
    const float e = w[4];
    const double td = (double)t;
    const float p0 = (float)((0.25 - td) * (double)(e + c));
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
// The same with the bound's two forms given (a field's own bounds).
template <int AX>
__device__ __forceinline__ int32_t quantize_eb(float p, float x, double inv_half_eb, double eb_d) {
    double diff = (double)p - (double)daz(x);
    if (AX == 0) return lround_q(inv_half_eb * diff);
    return lround_q((diff * 0.5) / eb_d);
}
__device__ __forceinline__ float reconstruct(float p, int32_t q, double neg_two_eb) {
    return (float)__fma_rn(neg_two_eb, (double)q, (double)p);
}

// Workgroup barrier for LDS traffic only: waits for this wave's LDS
// operations, then s_barrier.  Unlike __syncthreads() it does not wait for
// outstanding global loads, so loads issued for the next piece of work stay
// in flight across it.  Only for barriers that order LDS accesses.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// ---------------------------------------------------------------- reductions

// Wave-wide reductions on DPP row operations (VALU data movement) instead of
// LDS-path lane shuffles: an all-reduce inside each quad (quad_perm
// [1,0,3,2], [2,3,0,1]) and each row of 16 lanes (row_ror 4, 8), then the
// four row results read with readlane.  Every lane gets the result; the wave
// must be fully active.
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppRor4 = 0x124, kDppRor8 = 0x128;

template <int CTRL>
__device__ __forceinline__ uint32_t dpp32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
template <int CTRL, typename T>
__device__ __forceinline__ T dpp_move(T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit lanes");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, dpp32<CTRL>(__builtin_bit_cast(uint32_t, v)));
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint64_t lo = dpp32<CTRL>((uint32_t)u), hi = dpp32<CTRL>((uint32_t)(u >> 32));
        return __builtin_bit_cast(T, lo | (hi << 32));
    }
}
template <typename T>
__device__ __forceinline__ T read_lane(T v, int lane) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, (uint32_t)__builtin_amdgcn_readlane((int)__builtin_bit_cast(uint32_t, v), lane));
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint64_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
        const uint64_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
        return __builtin_bit_cast(T, lo | (hi << 32));
    }
}
template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce(T v, Op op) {
    v = op(v, dpp_move<kDppXor1>(v));
    v = op(v, dpp_move<kDppXor2>(v));
    v = op(v, dpp_move<kDppRor4>(v));
    v = op(v, dpp_move<kDppRor8>(v));
    return op(op(read_lane(v, 0), read_lane(v, 16)), op(read_lane(v, 32), read_lane(v, 48)));
}
// Inclusive prefix sum of 32-bit lanes over the wave on DPP moves: shifts
// right by 1, 2, 4, 8 inside each row of 16 (lanes without a source add 0),
// then row 0's total into row 1 and row 2's into row 3 (row_bcast:15), and
// the total of rows 0-1 into rows 2 and 3 (row_bcast:31).  The wave must be
// fully active.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_rows(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, true);
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    x += dpp_rows<0x111, 0xF>(x);  // row_shr:1
    x += dpp_rows<0x112, 0xF>(x);  // row_shr:2
    x += dpp_rows<0x114, 0xF>(x);  // row_shr:4
    x += dpp_rows<0x118, 0xF>(x);  // row_shr:8
    x += dpp_rows<0x142, 0xA>(x);  // row_bcast:15 into rows 1 and 3
    x += dpp_rows<0x143, 0xC>(x);  // row_bcast:31 into rows 2 and 3
    return x;
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    return wave_reduce(v, [](T a, T b) { return a + b; });
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
    return wave_reduce(v, [](T a, T b) { return b > a ? b : a; });
}
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
    return wave_reduce(v, [](T a, T b) { return b < a ? b : a; });
}

__device__ __forceinline__ void atomic_max_pos_double(double* p, double v) {
    atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
    return (uint32_t)(((uint64_t)n * f.m) >> f.sh);
}

// One QB into the workgroup's LDS window of win bins (symbols [-win/2, win/2)),
// or into the array's global window with an atomic when it falls outside.
// Symbols outside the global window are left to the dedicated histogram pass.
__device__ __forceinline__ void hist_add(uint32_t* lh, uint32_t win, uint32_t* gwin, int32_t q) {
    const uint32_t w = (uint32_t)q + (win >> 1);
    if (w < win) {
        atomicAdd(&lh[w], 1u);
    } else {
        const uint32_t g = (uint32_t)q + (uint32_t)kHistCentre;
        if (g < kHistGlobal) atomicAdd(&gwin[g], 1u);
    }
}

// Per-workgroup results of a compress round kernel (NW waves), at the
// workgroup's linear index wgl within the round: one Partial (max |x - x'|,
// sum (x - x')^2, QB range) stored, and the LDS histogram window's nonzero
// bins added into copy wgl % hist_copies of the round's window.  Called by
// every thread after its last QB; the barrier inside also completes the LDS
// histogram.
template <int NW = 4>
__device__ __forceinline__ void finish_wg(double err, double sq, int32_t qmin, int32_t qmax, const RoundDesc& R,
                                          uint64_t wgl, Partial* part, const uint32_t* lh, uint32_t* phist) {
    __shared__ double s_err[NW], s_sq[NW];
    __shared__ int32_t s_min[NW], s_max[NW];
    err = wave_max(err);
    sq = wave_sum(sq);
    qmin = wave_min(qmin);
    qmax = wave_max(qmax);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) { s_err[w] = err; s_sq[w] = sq; s_min[w] = qmin; s_max[w] = qmax; }
    __syncthreads();
    const uint64_t wg = R.part_off + wgl;
    if (threadIdx.x == 0) {
        for (int k = 1; k < NW; k++) {
            err = fmax(err, s_err[k]);
            sq += s_sq[k];
            qmin = min(qmin, s_min[k]);
            qmax = max(qmax, s_max[k]);
        }
        part[wg] = Partial{err, sq, qmin, qmax};
    }
    uint32_t* cw = phist + R.hist_off + (uint64_t)((uint32_t)wgl % R.hist_copies) * R.lds_win;
    for (uint32_t i = threadIdx.x; i < R.lds_win; i += blockDim.x) {
        const uint32_t v = lh[i];
        if (v) atomicAdd(&cw[i], v);
    }
}

// Multi-field kernels: the layer scalars of one field replace the plan's.
__device__ __forceinline__ void field_round(const FieldDesc& F, RoundDesc& R) {
    const FieldRound& q = F.r[R.layer];
    R.inv_half_eb = q.inv_half_eb;
    R.eb_d = q.eb_d;
    R.neg_two_eb = q.neg_two_eb;
    R.eb = q.eb;
    R.layer_slope = q.layer_slope;
}

using asz::ArrayAcc;
using asz::ErrAcc;

}  // namespace dev
}  // namespace asz
