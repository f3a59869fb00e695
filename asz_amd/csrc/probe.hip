// Kernels of the predictor-family probe (predictor_family.tpp:129-251).
//
// The reference compresses up to eight sampled blocks twice (cubic and
// quintic interior), each in a private copy of its context (the block, the
// blocks just below it on every axis, and a halo of 5 * pivot_stride / 2
// original values), block after block.  Within one round the blocks of a
// context only read values that earlier rounds finalised (coarse lattice
// points, or original halo values), so every block of every probe job runs
// the same round at once here: one workgroup per (job, block), one launch per
// round, with the round's per-block t chosen on the host from keys sampled
// by k_probe_keys, exactly as the reference chooses them.
#include "device.hpp"
#include "launch_family.hpp"

namespace asz {
namespace dev {

// Context copy and pivots (predictor_family.tpp:38-59, 152-159): local
// element e of box b; pivots are the lattice points at or after context_begin.
__global__ __launch_bounds__(256) void k_probe_init(const ProbeBox* __restrict__ boxes, const float* __restrict__ field,
                                                    float* __restrict__ orig, float* __restrict__ recon) {
    const ProbeBox B = boxes[blockIdx.y];
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < B.n; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t l2 = (uint32_t)(e % B.LD[2]);
        const uint64_t r = e / B.LD[2];
        const uint32_t l1 = (uint32_t)(r % B.LD[1]);
        const uint32_t l0 = (uint32_t)(r / B.LD[1]);
        const uint32_t g0 = B.halo[0] + l0, g1 = B.halo[1] + l1, g2 = B.halo[2] + l2;
        const float v = field[((uint64_t)g0 * B.D[1] + g1) * B.D[2] + g2];
        orig[B.base + e] = v;
        const bool pivot = g0 >= B.cb[0] && g1 >= B.cb[1] && g2 >= B.cb[2] && g0 % B.ps[0] == 0 &&
                           g1 % B.ps[1] == 0 && g2 % B.ps[2] == 0;
        float out = v;
        if (pivot) {
            const int32_t q = lround_q((double)daz(v) * B.inv_half_pe);
            out = (float)((double)q * B.two_pe);
        }
        recon[B.base + e] = out;
    }
}

// (A, B) keys of one block's t-selection samples (get_AB_map.tpp:14-22 on
// the local reconstruction, the arithmetic of k_tsel_keys).
__global__ __launch_bounds__(256) void k_probe_keys(const ProbeKeys* __restrict__ units, const float* __restrict__ recon,
                                                    long long* __restrict__ keys) {
    const ProbeKeys U = units[blockIdx.x];
    const uint32_t ns = U.cnt[0] * U.cnt[1] * U.cnt[2];
    long long* out = keys + 2 * U.key_off;
    const int64_t dist = (int64_t)U.dist;
    for (uint32_t idx = threadIdx.x; idx < ns; idx += blockDim.x) {
        const uint32_t q2 = idx % U.cnt[2], q01 = idx / U.cnt[2], q1 = q01 % U.cnt[1], q0 = q01 / U.cnt[1];
        const uint64_t pos = U.base + ((uint64_t)(U.lo[0] + q0 * U.step[0]) * U.LD[1] + (U.lo[1] + q1 * U.step[1])) *
                                          U.LD[2] + (U.lo[2] + q2 * U.step[2]);
        const float* p = recon + pos;
        const float a = p[-dist], c = p[-3 * dist], b = p[dist], e = p[3 * dist];
        const float x0 = daz(p[0]);
        const float A = (a + b) - (c + e);
        const float S = (b + e) + (a + c);
        const float Bv = (float)__fma_rn((double)S, 0.25, -(double)x0);
        out[2 * idx] = lround_f64(U.neg2 * A);
        out[2 * idx + 1] = lround_f64(U.neg2 * Bv);
    }
}

// One round of one block (block_compress_impl, compress.tpp:51-88).  Context
// blocks use the quantizer of block_compress_impl<false> (axis 0 inlined with
// the hoisted reciprocal), target blocks that of the out-of-line
// dot_compress<true> on every axis; target blocks also write their QBs in
// the round's QB order and a (prediction, reconstruction, original) record
// per target in loop order (layer_diagnostics_c::record, asz_param.hpp:93-102).
template <bool QUIN>
__device__ __forceinline__ float probe_predict(const ProbeUnit& U, const float* p, uint32_t id) {
    const int64_t dist = (int64_t)U.dist;
    const uint32_t h = U.h, right = U.right;
    float w[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    w[2] = p[-dist];
    if (QUIN && id >= 5 * h && id + 5 * h < right) {
        w[0] = p[-5 * dist];
        w[1] = p[-3 * dist];
        w[3] = p[dist];
        w[4] = p[3 * dist];
        w[5] = p[5 * dist];
    } else {
        if (id > h) w[1] = p[-3 * dist];
        if (id + h < right) w[3] = p[dist];
        if (id + 3 * h < right) w[4] = p[3 * dist];
    }
    return predict_w<QUIN>(id, h, right, w, U.t);
}

__global__ __launch_bounds__(256) void k_probe_round(const ProbeUnit* __restrict__ units, const float* __restrict__ orig,
                                                     float* __restrict__ recon, int32_t* __restrict__ qb,
                                                     float* __restrict__ rec) {
    const ProbeUnit U = units[blockIdx.x];
    const uint32_t n = U.n[0] * U.n[1] * U.n[2];
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
        const uint32_t j2 = j % U.n[2], j01 = j / U.n[2], j1 = j01 % U.n[1], j0 = j01 / U.n[1];
        const uint32_t i0 = U.start[0] + j0 * U.stride[0], i1 = U.start[1] + j1 * U.stride[1],
                       i2 = U.start[2] + j2 * U.stride[2];
        const uint64_t pos = U.base + ((uint64_t)i0 * U.LD[1] + i1) * U.LD[2] + i2;
        const uint32_t id = U.d == 0 ? i0 : (U.d == 1 ? i1 : i2);
        const float* p = recon + pos;
        const float pred = U.quintic ? probe_predict<true>(U, p, id) : probe_predict<false>(U, p, id);
        const float x = orig[pos];
        const double diff = (double)pred - (double)daz(x);
        const int32_t q = (U.diag || U.d != 0) ? lround_q((diff * 0.5) / U.eb_d) : lround_q(U.inv_half_eb * diff);
        const float r = reconstruct(pred, q, U.neg_two_eb);
        recon[pos] = r;
        if (U.diag) {
            uint32_t qi;
            if (U.order == 0) qi = (j1 * U.n[2] + j2) * U.n[0] + j0;
            else if (U.order == 1) qi = (j0 * U.n[2] + j2) * U.n[1] + j1;
            else qi = j;
            qb[U.qb_off + qi] = q;
            float* o = rec + 3 * (U.rec_off + j);
            o[0] = pred;
            o[1] = r;
            o[2] = x;
        }
    }
}

}  // namespace dev

void launch_probe_init(const ProbeBox* boxes, uint32_t nbox, uint64_t max_n, const float* field, float* orig,
                       float* recon, hipStream_t st) {
    if (!nbox || !max_n) return;
    const uint32_t gx = (uint32_t)std::min<uint64_t>((max_n + 255) / 256, 2048);
    hipLaunchKernelGGL(dev::k_probe_init, dim3(gx, nbox), dim3(256), 0, st, boxes, field, orig, recon);
}

void launch_probe_keys(const ProbeKeys* units, uint32_t nunit, const float* recon, int64_t* keys, hipStream_t st) {
    if (!nunit) return;
    hipLaunchKernelGGL(dev::k_probe_keys, dim3(nunit), dim3(256), 0, st, units, recon, (long long*)keys);
}

void launch_probe_round(const ProbeUnit* units, uint32_t nunit, const float* orig, float* recon, int32_t* qb,
                        float* rec, hipStream_t st) {
    if (!nunit) return;
    hipLaunchKernelGGL(dev::k_probe_round, dim3(nunit), dim3(256), 0, st, units, orig, recon, qb, rec);
}

}  // namespace asz
