// Residual energy of one round and its adjoint (gradient-tuned form).
//
// Within a round every coarse neighbour is fixed (written in earlier rounds),
// so the pre-quantization residual of a target is linear in the block's
// predictor parameters:
//   1D4R (interior cubic, compress.tpp:30-32):  r = A t + B,
//     A = (x[-h] + x[+h]) - (x[-3h] + x[+3h]),  B = (x[-h] + x[+h] + x[-3h] + x[+3h]) / 4 - x0;
//   1D6R with free weights (a, b) (compress.tpp:10-13 has a = 3/256, b = -25/256):
//     r = a (S5 - S1) + b (S3 - S1) + S1 / 2 - x0,  Sk = x[-kh] + x[+kh].
// E_b = sum over the block's targets that use that predictor of r^2;
// dE_b/dtheta = sum 2 r dr/dtheta.  The differences are formed pairwise,
// (x[-h] - x[-3h]) + (x[+h] - x[+3h]) and so on: neighbouring values of a
// smooth field are close, so each pairwise difference is exact and A, B and
// the features carry one rounding each instead of a cancellation error.
//
// One workgroup per block.  Target set per block (the ownership of
// compress.tpp:57-59 with the block extents of compress.tpp:183-188): along
// the round axis the odd multiples i of h in [b B, min((b+1) B, D)) with
// i >= 3h (5h) and i + 3h (5h) < right, right = min((b+1) B + 1, D); along
// another axis the multiples of its stride in [0, min(B, D-1)] for block 0
// and in [b B + s, min((b+1) B, D-1)] otherwise.
#include "device.hpp"
#include "launch.hpp"

namespace asz {
namespace dev {

struct BlockTargets {
    uint32_t lo[3], step[3], cnt[3];
    uint64_t n;
};

__device__ __forceinline__ BlockTargets block_targets(const RoundDesc& R, uint32_t bid, uint32_t reach) {
    const uint32_t bb[3] = {bid / (R.nb[2] * R.nb[1]), (bid / R.nb[2]) % R.nb[1], bid % R.nb[2]};
    BlockTargets T{};
    T.n = 1;
    for (int a = 0; a < 3; a++) {
        const uint32_t B0 = bb[a] * R.B[a];
        if (a == (int)R.d) {
            const uint32_t h = R.h;
            const uint32_t right = min(B0 + R.B[a] + 1, R.D[a]);
            const uint32_t end = min(B0 + R.B[a], R.D[a]);   // exclusive
            // first odd multiple of h >= max(B0, reach h)
            uint32_t first = max(B0, reach * h);
            first = first <= h ? h : h + ((first - h + 2 * h - 1) / (2 * h)) * 2 * h;
            // last target: i + reach h < right and i < end
            const int64_t lim = min<int64_t>((int64_t)end - 1, (int64_t)right - 1 - (int64_t)reach * h);
            T.lo[a] = first;
            T.step[a] = 2 * h;
            T.cnt[a] = lim < (int64_t)first ? 0 : (uint32_t)((lim - first) / (2 * h) + 1);
        } else {
            const uint32_t s = R.s[a];
            const uint32_t lo = bb[a] == 0 ? 0 : B0 + s;
            const uint32_t hi = min(B0 + R.B[a], R.D[a] - 1);  // inclusive
            T.lo[a] = lo;
            T.step[a] = s;
            T.cnt[a] = hi < lo ? 0 : (hi - lo) / s + 1;
        }
        T.n *= T.cnt[a];
    }
    return T;
}

// Linear features of one target: cubic (f0 = B, f1 = A), quintic (f0 = S1/2 - x0,
// f1 = S5 - S1, f2 = S3 - S1).
template <bool QUIN>
__device__ __forceinline__ void features(const float* __restrict__ p, int64_t dist, float f[3]) {
    const float x0 = daz(p[0]);
    const float a = p[-dist], b = p[dist], c = p[-3 * dist], e = p[3 * dist];
    if (!QUIN) {
        f[1] = (a - c) + (b - e);
        f[0] = (((a - x0) + (b - x0)) + ((c - x0) + (e - x0))) * 0.25f;
        f[2] = 0.0f;
    } else {
        const float g = p[-5 * dist], k = p[5 * dist];
        f[0] = ((a - x0) + (b - x0)) * 0.5f;
        f[1] = (g - a) + (k - b);
        f[2] = (c - a) + (e - b);
    }
}

// FWD: E_out[b] = sum r^2 (FP32 residuals, FP64 sum).
// ADJ: g_out[b * np + j] = gE[b] * sum 2 r dr/dtheta_j (FP32 throughout).
template <bool QUIN, bool ADJ>
__global__ __launch_bounds__(256) void k_energy(RoundDesc R, const float* __restrict__ work,
                                                const float* __restrict__ theta, const float* __restrict__ gE,
                                                double* __restrict__ E_out, float* __restrict__ g_out) {
    constexpr int NP = QUIN ? 2 : 1;
    __shared__ double s_e[4];
    __shared__ float s_g[4][NP];
    const uint32_t bid = blockIdx.x;
    const BlockTargets T = block_targets(R, bid, QUIN ? 5 : 3);
    const int64_t dist = (int64_t)R.dist;
    float th[2] = {theta[bid * NP], QUIN ? theta[bid * NP + 1] : 0.0f};
    double e = 0;
    float g[2] = {0.0f, 0.0f};
    for (uint64_t idx = threadIdx.x; idx < T.n; idx += blockDim.x) {
        const uint32_t q2 = (uint32_t)(idx % T.cnt[2]);
        const uint64_t q01 = idx / T.cnt[2];
        const uint32_t q1 = (uint32_t)(q01 % T.cnt[1]), q0 = (uint32_t)(q01 / T.cnt[1]);
        const uint64_t pos = ((uint64_t)(T.lo[0] + q0 * T.step[0]) * R.D[1] + (T.lo[1] + q1 * T.step[1])) * R.D[2] +
                             (T.lo[2] + q2 * T.step[2]);
        float f[3];
        features<QUIN>(work + pos, dist, f);
        const float r = QUIN ? __fmaf_rn(th[1], f[2], __fmaf_rn(th[0], f[1], f[0])) : __fmaf_rn(th[0], f[1], f[0]);
        if (ADJ) {
            g[0] = __fmaf_rn(2.0f * r, f[1], g[0]);
            if (QUIN) g[1] = __fmaf_rn(2.0f * r, f[2], g[1]);
        } else {
            e += (double)r * (double)r;
        }
    }
    const int w = threadIdx.x / kWave;
    if (ADJ) {
#pragma unroll
        for (int j = 0; j < NP; j++) {
            const float v = wave_sum(g[j]);
            if ((threadIdx.x & (kWave - 1)) == 0) s_g[w][j] = v;
        }
    } else {
        e = wave_sum(e);
        if ((threadIdx.x & (kWave - 1)) == 0) s_e[w] = e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (ADJ) {
            const float up = gE[bid];
            for (int j = 0; j < NP; j++)
                g_out[bid * NP + j] = up * ((s_g[0][j] + s_g[1][j]) + (s_g[2][j] + s_g[3][j]));
        } else {
            E_out[bid] = (s_e[0] + s_e[1]) + (s_e[2] + s_e[3]);
        }
    }
}

}  // namespace dev

void launch_energy(const RoundDesc& R, uint64_t nblocks, bool quintic, bool adjoint, const float* work,
                   const float* theta, const float* gE, double* E, float* g, hipStream_t st) {
    const dim3 grid((uint32_t)nblocks), blk(256);
    if (!quintic && !adjoint) hipLaunchKernelGGL((dev::k_energy<false, false>), grid, blk, 0, st, R, work, theta, gE, E, g);
    if (!quintic && adjoint) hipLaunchKernelGGL((dev::k_energy<false, true>), grid, blk, 0, st, R, work, theta, gE, E, g);
    if (quintic && !adjoint) hipLaunchKernelGGL((dev::k_energy<true, false>), grid, blk, 0, st, R, work, theta, gE, E, g);
    if (quintic && adjoint) hipLaunchKernelGGL((dev::k_energy<true, true>), grid, blk, 0, st, R, work, theta, gE, E, g);
}

}  // namespace asz
