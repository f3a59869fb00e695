// Range scan, pivots and the level-synchronous interpolation rounds.
//
// One launch per round covers every block: target (k0, k1, k2) of the round's
// global lattice maps to one thread, k2 (DIM2) fastest so loads and stores
// coalesce along the fastest axis.  The reference runs block -> round
// (compress.tpp:179-208); every round-r read touches points written in rounds
// < r or the pivots, and every target is owned by exactly one block, so the
// round-major order here reads exactly the same values (no block can see
// another block's round-r writes).  The field is updated in place: targets
// hold the original value until their own thread overwrites it, neighbours
// (coarse points) are never written during the round.
#include "device.hpp"
#include "launch.hpp"
#include "layout.hpp"

namespace asz {
namespace dev {

// ------------------------------------------------------------------ range scan

__device__ __forceinline__ uint32_t order_key(float x) {
    x = daz(x);
    if (x == 0.0f) x = 0.0f;  // -0 == +0
    uint32_t b = __float_as_uint(x);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ __launch_bounds__(256) void k_range(const float* __restrict__ x, uint64_t n, uint32_t* out) {
    // out[0] = min key, out[1] = max key, out[2] = non-finite flag
    uint32_t mn = 0xFFFFFFFFu, mx = 0u, bad = 0u;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n4 = n / 4;
    const float4* x4 = reinterpret_cast<const float4*>(x);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        float4 v = x4[i];
        float e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            bad |= !isfinite(e[k]);
            uint32_t key = order_key(e[k]);
            mn = min(mn, key);
            mx = max(mx, key);
        }
    }
    for (uint64_t i = n4 * 4 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        bad |= !isfinite(x[i]);
        uint32_t key = order_key(x[i]);
        mn = min(mn, key);
        mx = max(mx, key);
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    bad = wave_max(bad);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        atomicMin(&out[0], mn);
        atomicMax(&out[1], mx);
        if (bad) atomicOr(&out[2], 1u);
    }
}

// first index whose DAZ value equals zero (min side) / last such index (max side)
__global__ __launch_bounds__(256) void k_zero_index(const float* __restrict__ x, uint64_t n, unsigned long long* out) {
    unsigned long long first = ~0ull, last = 0ull;
    bool any = false;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (daz(x[i]) == 0.0f) {
            first = min(first, (unsigned long long)i);
            last = max(last, (unsigned long long)i);
            any = true;
        }
    }
    first = wave_min(first);
    last = wave_max(last);
    int a = wave_max((int)any);
    if ((threadIdx.x & (kWave - 1)) == 0 && a) {
        atomicMin(&out[0], first);
        atomicMax(&out[1], last);
    }
}

// ------------------------------------------------------------------ accumulators

__device__ __forceinline__ void flush_acc(double err, double sq, int32_t qmin, int32_t qmax, ErrAcc* eacc,
                                          ArrayAcc* aacc) {
    __shared__ double s_err[4], s_sq[4];
    __shared__ int32_t s_min[4], s_max[4];
    err = wave_max(err);
    sq = wave_sum(sq);
    qmin = wave_min(qmin);
    qmax = wave_max(qmax);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) { s_err[w] = err; s_sq[w] = sq; s_min[w] = qmin; s_max[w] = qmax; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x / kWave;
        for (int k = 1; k < nw; k++) {
            err = fmax(err, s_err[k]);
            sq += s_sq[k];
            qmin = min(qmin, s_min[k]);
            qmax = max(qmax, s_max[k]);
        }
        if (qmin <= qmax) {
            atomic_max_pos_double(&eacc->max_err, err);
            atomicAdd(&eacc->sq_sum, sq);
            atomicMin(&aacc->qmin, qmin);
            atomicMax(&aacc->qmax, qmax);
        }
    }
}

// ------------------------------------------------------------------ pivots (compress.tpp:161-175)

__global__ __launch_bounds__(256) void k_pivots_compress(float* __restrict__ work, int32_t* __restrict__ qb,
                                                         PivotDesc P, ErrAcc* eacc, ArrayAcc* aacc) {
    const uint32_t k2 = blockIdx.x * blockDim.x + threadIdx.x, k1 = blockIdx.y, k0 = blockIdx.z;
    double err = 0, sq = 0;
    int32_t qmin = INT32_MAX, qmax = INT32_MIN;
    if (k2 < P.pn[2]) {
        uint64_t pos = ((uint64_t)k0 * P.ps[0] * P.D[1] + (uint64_t)k1 * P.ps[1]) * P.D[2] + (uint64_t)k2 * P.ps[2];
        float x = work[pos];
        int32_t q = lround_q((double)daz(x) * P.inv_half_pe);
        float r = (float)((double)q * P.two_pe);
        work[pos] = r;
        qb[((uint64_t)k0 * P.pn[1] + k1) * P.pn[2] + k2] = q;
        double e = (double)daz(x) - (double)r;
        err = fabs(e);
        sq = e * e;
        qmin = qmax = q;
    }
    flush_acc(err, sq, qmin, qmax, eacc, aacc);
}

__global__ __launch_bounds__(256) void k_pivots_decompress(float* __restrict__ out, const int32_t* __restrict__ qb,
                                                           PivotDesc P) {
    const uint32_t k2 = blockIdx.x * blockDim.x + threadIdx.x, k1 = blockIdx.y, k0 = blockIdx.z;
    if (k2 >= P.pn[2]) return;
    uint64_t pos = ((uint64_t)k0 * P.ps[0] * P.D[1] + (uint64_t)k1 * P.ps[1]) * P.D[2] + (uint64_t)k2 * P.ps[2];
    int32_t q = qb[((uint64_t)k0 * P.pn[1] + k1) * P.pn[2] + k2];
    out[pos] = (float)((double)q * P.two_pe);
}

// ------------------------------------------------------------------ rounds

template <int AX, bool QUIN>
__global__ __launch_bounds__(256) void k_round_compress(RoundDesc R, float* __restrict__ work,
                                                        int32_t* __restrict__ qb, const uint32_t* __restrict__ tabs,
                                                        const float* __restrict__ tblk, ErrAcc* eacc,
                                                        ArrayAcc* aacc, int write_recon) {
    uint32_t k[3] = {blockIdx.z, blockIdx.y, blockIdx.x * blockDim.x + threadIdx.x};
    double err = 0, sq = 0;
    int32_t qmin = INT32_MAX, qmax = INT32_MIN;
    if (k[2] < R.n[2]) {
        uint32_t i[3];
        target_coords(R, AX, k, i);
        TargetLoc L;
        locate(R, AX, k, i, L);
        const uint32_t right = min((L.b[AX] + 1) * R.B[AX] + 1, R.D[AX]);
        const float t = tblk[((uint64_t)L.b[0] * R.nb[1] + L.b[1]) * R.nb[2] + L.b[2]];
        const uint64_t pos = ((uint64_t)i[0] * R.D[1] + i[1]) * R.D[2] + i[2];
        float* rp = work + pos;
        const float p = predict<QUIN>(rp, i[AX], R.h, right, (int64_t)R.dist, t);
        const float x = *rp;
        const int32_t q = quantize<AX>(p, x, R);
        const float xr = reconstruct(p, q, R.neg_two_eb);
        if (write_recon) *rp = xr;
        qb[R.qb_off + qb_slot(R, tabs, L)] = q;
        double e = (double)daz(x) - (double)xr;
        err = fabs(e);
        sq = e * e;
        qmin = qmax = q;
    }
    flush_acc(err, sq, qmin, qmax, eacc, aacc);
}

template <int AX, bool QUIN>
__global__ __launch_bounds__(256) void k_round_decompress(RoundDesc R, float* __restrict__ out,
                                                          const int32_t* __restrict__ qb,
                                                          const uint32_t* __restrict__ tabs,
                                                          const float* __restrict__ tblk) {
    uint32_t k[3] = {blockIdx.z, blockIdx.y, blockIdx.x * blockDim.x + threadIdx.x};
    if (k[2] >= R.n[2]) return;
    uint32_t i[3];
    target_coords(R, AX, k, i);
    TargetLoc L;
    locate(R, AX, k, i, L);
    const uint32_t right = min((L.b[AX] + 1) * R.B[AX] + 1, R.D[AX]);
    const float t = tblk[((uint64_t)L.b[0] * R.nb[1] + L.b[1]) * R.nb[2] + L.b[2]];
    const uint64_t pos = ((uint64_t)i[0] * R.D[1] + i[1]) * R.D[2] + i[2];
    float* rp = out + pos;
    const float p = predict<QUIN>(rp, i[AX], R.h, right, (int64_t)R.dist, t);
    const int32_t q = qb[R.qb_off + qb_slot(R, tabs, L)];
    *rp = reconstruct(p, q, R.neg_two_eb);
}

}  // namespace dev

// ------------------------------------------------------------------ launchers

void launch_range(const float* x, uint64_t n, uint32_t* d_out3, hipStream_t st) {
    uint64_t blocks = std::min<uint64_t>((n / 4 + 255) / 256 + 1, 8192);
    hipLaunchKernelGGL(dev::k_range, dim3((uint32_t)blocks), dim3(256), 0, st, x, n, d_out3);
}
void launch_zero_index(const float* x, uint64_t n, unsigned long long* d_out2, hipStream_t st) {
    uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(dev::k_zero_index, dim3((uint32_t)blocks), dim3(256), 0, st, x, n, d_out2);
}

void launch_pivots(bool compress, float* work, int32_t* qb, const PivotDesc& P, void* eacc, void* aacc,
                   hipStream_t st) {
    dim3 grid((P.pn[2] + 255) / 256, P.pn[1], P.pn[0]);
    if (compress)
        hipLaunchKernelGGL(dev::k_pivots_compress, grid, dim3(256), 0, st, work, qb, P, (dev::ErrAcc*)eacc,
                           (dev::ArrayAcc*)aacc);
    else
        hipLaunchKernelGGL(dev::k_pivots_decompress, grid, dim3(256), 0, st, work, (const int32_t*)qb, P);
}

template <int AX, bool Q>
static void round_c(const RoundDesc& R, float* w, int32_t* qb, const uint32_t* tabs, const float* tblk, void* e,
                    void* a, int wr, hipStream_t st) {
    dim3 grid((R.n[2] + 255) / 256, R.n[1], R.n[0]);
    hipLaunchKernelGGL((dev::k_round_compress<AX, Q>), grid, dim3(256), 0, st, R, w, qb, tabs, tblk,
                       (dev::ErrAcc*)e, (dev::ArrayAcc*)a, wr);
}
template <int AX, bool Q>
static void round_d(const RoundDesc& R, float* w, const int32_t* qb, const uint32_t* tabs, const float* tblk,
                    hipStream_t st) {
    dim3 grid((R.n[2] + 255) / 256, R.n[1], R.n[0]);
    hipLaunchKernelGGL((dev::k_round_decompress<AX, Q>), grid, dim3(256), 0, st, R, w, qb, tabs, tblk);
}

void launch_round_compress(const RoundDesc& R, bool quintic, float* work, int32_t* qb, const uint32_t* tabs,
                           const float* tblk, void* eacc, void* aacc, int write_recon, hipStream_t st) {
    if (R.count == 0) return;
    switch (R.d * 2 + (quintic ? 1 : 0)) {
        case 0: round_c<0, false>(R, work, qb, tabs, tblk, eacc, aacc, write_recon, st); break;
        case 1: round_c<0, true>(R, work, qb, tabs, tblk, eacc, aacc, write_recon, st); break;
        case 2: round_c<1, false>(R, work, qb, tabs, tblk, eacc, aacc, write_recon, st); break;
        case 3: round_c<1, true>(R, work, qb, tabs, tblk, eacc, aacc, write_recon, st); break;
        case 4: round_c<2, false>(R, work, qb, tabs, tblk, eacc, aacc, write_recon, st); break;
        default: round_c<2, true>(R, work, qb, tabs, tblk, eacc, aacc, write_recon, st); break;
    }
}

void launch_round_decompress(const RoundDesc& R, bool quintic, float* out, const int32_t* qb, const uint32_t* tabs,
                             const float* tblk, hipStream_t st) {
    if (R.count == 0) return;
    switch (R.d * 2 + (quintic ? 1 : 0)) {
        case 0: round_d<0, false>(R, out, qb, tabs, tblk, st); break;
        case 1: round_d<0, true>(R, out, qb, tabs, tblk, st); break;
        case 2: round_d<1, false>(R, out, qb, tabs, tblk, st); break;
        case 3: round_d<1, true>(R, out, qb, tabs, tblk, st); break;
        case 4: round_d<2, false>(R, out, qb, tabs, tblk, st); break;
        default: round_d<2, true>(R, out, qb, tabs, tblk, st); break;
    }
}

}  // namespace asz
