// Kernels of automatic plan generation (generate_config.tpp:443-678).
//
// For one candidate round of the reverse greedy search, the reference
// predicts a sub-lattice of the ORIGINAL field (open loop: the neighbours are
// original values, not reconstructions) with the round's best t, quantizes it
// and measures the QB stream (Huffman + Zstd) and the float MSE.  Here one
// launch predicts the samples of every candidate of a round at once
// (k_plan_sample, one grid row per candidate, samples in the reference's loop
// order), one workgroup per candidate replays the reference's sequential
// float MSE sum (k_sq_chain), and k_plan_minmax gives each candidate's QB
// range for the dense histogram.  The Huffman kernels encode the candidates'
// QB arrays afterwards.
#include "device.hpp"
#include "launch_plangen.hpp"

namespace asz {
namespace dev {

// Sample k of candidate S in the reference's nesting (generate_config.tpp:
// 264-437): order 0 ("120") loops i1, i2, i0; order 1 ("021") i0, i2, i1;
// order 2 ("012") i0, i1, i2.
__device__ __forceinline__ void sample_coords(const PlanSampleDesc& S, uint64_t k, uint32_t i[3]) {
    uint32_t q[3];
    if (S.order == 0) {  // innermost axis 0
        q[0] = (uint32_t)(k % S.cnt[0]);
        uint64_t r = k / S.cnt[0];
        q[2] = (uint32_t)(r % S.cnt[2]);
        q[1] = (uint32_t)(r / S.cnt[2]);
    } else if (S.order == 1) {  // innermost axis 1
        q[1] = (uint32_t)(k % S.cnt[1]);
        uint64_t r = k / S.cnt[1];
        q[2] = (uint32_t)(r % S.cnt[2]);
        q[0] = (uint32_t)(r / S.cnt[2]);
    } else {
        q[2] = (uint32_t)(k % S.cnt[2]);
        uint64_t r = k / S.cnt[2];
        q[1] = (uint32_t)(r % S.cnt[1]);
        q[0] = (uint32_t)(r / S.cnt[1]);
    }
#pragma unroll
    for (int a = 0; a < 3; a++) i[a] = S.start[a] + q[a] * S.step[a];
}

// dot_sample_compress (generate_config.tpp:225-254) with GCC's lowering: the
// cubic and edge predictors are those of compress.tpp:14-32 (predict_w), the
// quantizer is q = llroundf((p - x) * (0.5f / eb)) in float, and the error
// term is fmaf(-(float)(2 q), eb, p - x) (the product 2 q in int32).
__global__ __launch_bounds__(256) void k_plan_sample(const PlanSampleDesc* __restrict__ descs,
                                                     const float* __restrict__ x, int32_t* __restrict__ qb,
                                                     float* __restrict__ err) {
    const PlanSampleDesc S = descs[blockIdx.y];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < S.count; k += stride) {
        uint32_t i[3];
        sample_coords(S, k, i);
        const uint64_t pos = ((uint64_t)i[0] * S.D[1] + i[1]) * S.D[2] + i[2];
        const uint32_t id = S.d == 0 ? i[0] : (S.d == 1 ? i[1] : i[2]);
        const uint32_t h = S.h, right = S.d == 0 ? S.D[0] : (S.d == 1 ? S.D[1] : S.D[2]);
        const float* p = x + pos;
        const int64_t dist = (int64_t)S.dist;
        // w[m] = value at id + (2m - 5) h; only the entries the branch reads are loaded
        float w[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        w[2] = p[-dist];
        if (id > h) w[1] = p[-3 * dist];
        if (id + h < right) w[3] = p[dist];
        if (id + 3 * h < right) w[4] = p[3 * dist];
        const float pred = predict_w<false>(id, h, right, w, S.t);
        const float d = pred - p[0];
        const int32_t q = (int32_t)lround_f64(d * S.neg2);
        qb[S.off + k] = q;
        err[S.off + k] = __fmaf_rn(-(float)(int32_t)((uint32_t)q * 2u), S.eb, d);
    }
}

// Per-candidate QB min / max (out2 pre-set to INT32_MAX, INT32_MIN).
__global__ __launch_bounds__(256) void k_plan_minmax(const PlanSampleDesc* __restrict__ descs,
                                                     const int32_t* __restrict__ qb, int32_t* out2) {
    const PlanSampleDesc S = descs[blockIdx.y];
    int32_t mn = INT32_MAX, mx = INT32_MIN;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < S.count;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const int32_t v = qb[S.off + k];
        mn = min(mn, v);
        mx = max(mx, v);
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0 && mn <= mx) {
        atomicMin(&out2[2 * blockIdx.y], mn);
        atomicMax(&out2[2 * blockIdx.y + 1], mx);
    }
}

// The reference accumulates mse += e * e in sample order, one float FMA per
// sample (generate_config.tpp:253).  That chain is sequential, so one lane
// runs it while the workgroup stages the next tile of errors in LDS.
constexpr int kChainTile = 4096;

__device__ __forceinline__ float sq4(const float4 v, float s) {
    s = __fmaf_rn(v.x, v.x, s);
    s = __fmaf_rn(v.y, v.y, s);
    s = __fmaf_rn(v.z, v.z, s);
    return __fmaf_rn(v.w, v.w, s);
}

__global__ __launch_bounds__(256) void k_sq_chain(const PlanSampleDesc* __restrict__ descs,
                                                  const float* __restrict__ err, float* __restrict__ out) {
    __shared__ __align__(16) float buf[2][kChainTile];
    const PlanSampleDesc S = descs[blockIdx.x];
    const float* e = err + S.off;
    const uint64_t n = S.count;
    constexpr int kPer = kChainTile / 256;
    float s = 0.0f;
    float r[kPer];
    // tile 0
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const uint64_t g = (uint64_t)j * 256 + threadIdx.x;
        buf[0][j * 256 + threadIdx.x] = g < n ? e[g] : 0.0f;
    }
    __syncthreads();
    int cur = 0;
    for (uint64_t t0 = 0; t0 < n; t0 += kChainTile) {
        const uint64_t nx = t0 + kChainTile;
#pragma unroll
        for (int j = 0; j < kPer; j++) {
            const uint64_t g = nx + (uint64_t)j * 256 + threadIdx.x;
            r[j] = g < n ? e[g] : 0.0f;
        }
        if (threadIdx.x == 0) {
            const uint32_t m = (uint32_t)min<uint64_t>(kChainTile, n - t0);
            // batches of 128 values: 32 LDS reads in flight, then the FMAs,
            // so the read latency is paid once per batch
            const float4* b4 = reinterpret_cast<const float4*>(buf[cur]);
            const uint32_t m128 = m / 128;
            for (uint32_t k = 0; k < m128; k++) {
                float4 v[32];
#pragma unroll
                for (int u = 0; u < 32; u++) v[u] = b4[32 * k + u];
#pragma unroll
                for (int u = 0; u < 32; u++) s = sq4(v[u], s);
            }
            const float* b = buf[cur];
            for (uint32_t j = 128 * m128; j < m; j++) s = __fmaf_rn(b[j], b[j], s);
        }
#pragma unroll
        for (int j = 0; j < kPer; j++) buf[cur ^ 1][j * 256 + threadIdx.x] = r[j];
        __syncthreads();
        cur ^= 1;
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

}  // namespace dev

void launch_plan_sample(const PlanSampleDesc* descs, uint32_t ncand, uint64_t max_count, const float* x, int32_t* qb,
                        float* err, hipStream_t st) {
    if (!ncand || !max_count) return;
    const uint32_t gx = (uint32_t)std::min<uint64_t>((max_count + 255) / 256, 4096);
    hipLaunchKernelGGL(dev::k_plan_sample, dim3(gx, ncand), dim3(256), 0, st, descs, x, qb, err);
}

void launch_plan_minmax(const PlanSampleDesc* descs, uint32_t ncand, uint64_t max_count, const int32_t* qb,
                        int32_t* out2, hipStream_t st) {
    if (!ncand || !max_count) return;
    const uint32_t gx = (uint32_t)std::min<uint64_t>((max_count + 255) / 256, 1024);
    hipLaunchKernelGGL(dev::k_plan_minmax, dim3(gx, ncand), dim3(256), 0, st, descs, qb, out2);
}

void launch_sq_chain(const PlanSampleDesc* descs, uint32_t ncand, const float* err, float* out, hipStream_t st) {
    if (!ncand) return;
    hipLaunchKernelGGL(dev::k_sq_chain, dim3(ncand), dim3(256), 0, st, descs, err, out);
}

}  // namespace asz
