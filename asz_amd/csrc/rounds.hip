// Range scan, pivots and the level-synchronous interpolation rounds.
//
// One launch per round covers every block.  The reference runs block -> round
// (compress.tpp:179-208); every round-r read touches points written in rounds
// < r or the pivots, and every target is owned by exactly one block, so the
// round-major order here reads exactly the same values.  The field is updated
// in place: targets hold the original value until their own thread rewrites
// them, neighbours (coarse points, even multiples of h along the round axis)
// are never written during the round.
//
// Two kernel shapes:
//  * k_round_sweep (round axis 0 or 1): a wave owns 64 consecutive lattice
//    columns along DIM2 and walks a segment (<= kSeg targets) of one block
//    along the round axis, keeping the coarse neighbours in a register window
//    (each coarse value is loaded once per segment).  QB values go straight
//    to their slots when the serialization order runs along DIM2 ("012"), are
//    staged in LDS and written as contiguous runs when it runs along the round
//    axis ("120" on axis 0, "021" on axis 1), and are scattered otherwise.
//  * k_round_direct (round axis 2, and every order): one thread per target,
//    lanes along DIM2.
// Error statistics and QB ranges are reduced per workgroup into partials and
// folded by one small kernel after the last round.
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
    __shared__ uint32_t s_mn[4], s_mx[4], s_bad[4];
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
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) { s_mn[w] = mn; s_mx[w] = mx; s_bad[w] = bad; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) { mn = min(mn, s_mn[k]); mx = max(mx, s_mx[k]); bad |= s_bad[k]; }
        atomicMin(&out[0], mn);
        atomicMax(&out[1], mx);
        if (bad) atomicOr(&out[2], 1u);
    }
}

// first index whose DAZ value equals zero (min side) / last such index (max side)
__global__ __launch_bounds__(256) void k_zero_index(const float* __restrict__ x, uint64_t n, unsigned long long* out) {
    unsigned long long first = ~0ull, last = 0ull;
    int any = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (daz(x[i]) == 0.0f) {
            first = min(first, (unsigned long long)i);
            last = max(last, (unsigned long long)i);
            any = 1;
        }
    }
    first = wave_min(first);
    last = wave_max(last);
    any = wave_max(any);
    if ((threadIdx.x & (kWave - 1)) == 0 && any) {
        atomicMin(&out[0], first);
        atomicMax(&out[1], last);
    }
}

// ------------------------------------------------------------------ partial reductions

// One Partial per workgroup: max |x - x'|, sum (x - x')^2, QB range.
__device__ __forceinline__ void write_partial(double err, double sq, int32_t qmin, int32_t qmax, Partial* out) {
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
        for (int k = 1; k < 4; k++) {
            err = fmax(err, s_err[k]);
            sq += s_sq[k];
            qmin = min(qmin, s_min[k]);
            qmax = max(qmax, s_max[k]);
        }
        const uint64_t wg = blockIdx.x + (uint64_t)gridDim.x * (blockIdx.y + (uint64_t)gridDim.y * blockIdx.z);
        out[wg] = Partial{err, sq, qmin, qmax};
    }
}

// One workgroup per round: fold that round's partials into arr[2 + r] / rerr[r].
__global__ __launch_bounds__(256) void k_reduce_partials(const Partial* __restrict__ part,
                                                         const uint64_t* __restrict__ ranges, DevAcc* acc) {
    const uint32_t r = blockIdx.x;
    const uint64_t off = ranges[2 * r], n = ranges[2 * r + 1];
    double err = 0, sq = 0;
    int32_t qmin = INT32_MAX, qmax = INT32_MIN;
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
        Partial p = part[off + i];
        err = fmax(err, p.max_err);
        sq += p.sq;
        qmin = min(qmin, p.qmin);
        qmax = max(qmax, p.qmax);
    }
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
        for (int k = 1; k < 4; k++) {
            err = fmax(err, s_err[k]);
            sq += s_sq[k];
            qmin = min(qmin, s_min[k]);
            qmax = max(qmax, s_max[k]);
        }
        acc->rerr[r] = ErrAcc{err, sq};
        acc->arr[2 + r].qmin = qmin;
        acc->arr[2 + r].qmax = qmax;
    }
}

// ------------------------------------------------------------------ pivots (compress.tpp:161-175)

__device__ __forceinline__ void flush_atomic(double err, double sq, int32_t qmin, int32_t qmax, ErrAcc* eacc,
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
        for (int k = 1; k < 4; k++) {
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
    flush_atomic(err, sq, qmin, qmax, eacc, aacc);
}

__global__ __launch_bounds__(256) void k_pivots_decompress(float* __restrict__ out, const int32_t* __restrict__ qb,
                                                           PivotDesc P) {
    const uint32_t k2 = blockIdx.x * blockDim.x + threadIdx.x, k1 = blockIdx.y, k0 = blockIdx.z;
    if (k2 >= P.pn[2]) return;
    uint64_t pos = ((uint64_t)k0 * P.ps[0] * P.D[1] + (uint64_t)k1 * P.ps[1]) * P.D[2] + (uint64_t)k2 * P.ps[2];
    int32_t q = qb[((uint64_t)k0 * P.pn[1] + k1) * P.pn[2] + k2];
    out[pos] = (float)((double)q * P.two_pe);
}

// ------------------------------------------------------------------ direct rounds

template <int AX, bool QUIN, bool DEC>
__global__ __launch_bounds__(256) void k_round_direct(RoundDesc R, float* __restrict__ work, int32_t* __restrict__ qb,
                                                      const uint32_t* __restrict__ tabs,
                                                      const float* __restrict__ tblk, Partial* part, int write_recon) {
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
        const uint64_t slot = R.qb_off + qb_slot(R, tabs, L);
        if (DEC) {
            *rp = reconstruct(p, qb[slot], R.neg_two_eb);
        } else {
            const float x = *rp;
            const int32_t q = quantize<AX>(p, x, R);
            const float xr = reconstruct(p, q, R.neg_two_eb);
            if (write_recon) *rp = xr;
            qb[slot] = q;
            double e = (double)daz(x) - (double)xr;
            err = fabs(e);
            sq = e * e;
            qmin = qmax = q;
        }
    }
    if (!DEC) write_partial(err, sq, qmin, qmax, part + R.part_off);
}

// ------------------------------------------------------------------ sweep rounds (axis 0 / 1)

// The reference predictor over a register window w[m] = coarse value at
// i + (2m - 5) h, m = 0..5 (w[0] = i-5h ... w[5] = i+5h).  Same branches and
// arithmetic as predict<QUIN>.
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

// WMODE 0: order "012" (QB runs along DIM2)      -> direct per-lane writes
// WMODE 1: order runs along the round axis       -> LDS tile, contiguous runs
// WMODE 2: order runs along the other axis       -> scattered per-lane writes
template <int AX, bool QUIN, bool DEC, int WMODE>
__global__ __launch_bounds__(256) void k_round_sweep(RoundDesc R, float* __restrict__ work, int32_t* __restrict__ qb,
                                                     const uint32_t* __restrict__ tabs,
                                                     const float* __restrict__ tblk, const SweepSeg* __restrict__ segs,
                                                     Partial* part, int write_recon) {
    constexpr int OA = 1 - AX;  // the non-round axis other than DIM2
    __shared__ int32_t tile[WMODE == 1 ? 4 : 1][WMODE == 1 ? kSweepSeg : 1][65];
    __shared__ uint64_t lbase[WMODE == 1 ? 4 : 1][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const SweepSeg S = segs[R.seg_off + blockIdx.z];
    const uint32_t k2 = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t ko = blockIdx.y;
    const bool valid = k2 < R.n[2];
    const uint32_t kk2 = valid ? k2 : 0;
    const uint32_t h = R.h;

    // block coordinates and local indices of this lane's column
    uint32_t b[3], j[3];
    b[AX] = S.b;
    j[AX] = S.k0 - S.b * (R.B[AX] / (2 * h));
    {
        const uint32_t i2 = R.s[2] * kk2;
        b[2] = i2 == 0 ? 0 : (i2 - 1) / R.B[2];
        j[2] = i2 == 0 ? 0 : (b[2] == 0 ? kk2 : kk2 - b[2] * (R.B[2] / R.s[2]) - 1);
        const uint32_t io = R.s[OA] * ko;
        b[OA] = io == 0 ? 0 : (io - 1) / R.B[OA];
        j[OA] = io == 0 ? 0 : (b[OA] == 0 ? ko : ko - b[OA] * (R.B[OA] / R.s[OA]) - 1);
    }
    uint32_t cnt[3], pre[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        cnt[a] = tabs[R.tab_off[a] + b[a]];
        pre[a] = tabs[R.tab_off[a] + R.nb[a] + b[a]];
    }
    const uint64_t base = R.qb_off + (uint64_t)pre[0] * R.C[1] * R.C[2] +
                          (uint64_t)cnt[0] * ((uint64_t)pre[1] * R.C[2] + (uint64_t)cnt[1] * pre[2]);
    const uint32_t right = min((S.b + 1) * R.B[AX] + 1, R.D[AX]);
    const float t = tblk[((uint64_t)b[0] * R.nb[1] + b[1]) * R.nb[2] + b[2]];
    const uint64_t sd = AX == 0 ? (uint64_t)R.D[1] * R.D[2] : (uint64_t)R.D[2];
    const uint64_t so = OA == 0 ? (uint64_t)R.D[1] * R.D[2] : (uint64_t)R.D[2];
    const uint64_t col = (uint64_t)(R.s[OA] * ko) * so + (uint64_t)R.s[2] * kk2;
    const int64_t Dd = R.D[AX];

    // QB slot of target c of the segment (orders of asz_param.hpp:122-129)
    auto slot_of = [&](uint32_t c) -> uint64_t {
        uint32_t jj[3] = {j[0], j[1], j[2]};
        jj[AX] += c;
        if (R.order == 0) return base + ((uint64_t)jj[1] * cnt[2] + jj[2]) * cnt[0] + jj[0];
        if (R.order == 1) return base + ((uint64_t)jj[0] * cnt[2] + jj[2]) * cnt[1] + jj[1];
        return base + ((uint64_t)jj[0] * cnt[1] + jj[1]) * cnt[2] + jj[2];
    };
    if (WMODE == 1) {
        lbase[wave][lane] = slot_of(0);
        if (DEC) {
            __syncthreads();
            for (uint32_t e = lane; e < 64 * S.cnt; e += 64) {
                const uint32_t jl = e / S.cnt, c = e - jl * S.cnt;
                const uint32_t kl = blockIdx.x * blockDim.x + wave * 64 + jl;
                if (kl < R.n[2]) tile[wave][c][jl] = qb[lbase[wave][jl] + c];
            }
        }
        __syncthreads();
    }

    double err = 0, sq = 0;
    int32_t qmin = INT32_MAX, qmax = INT32_MIN;
    if (valid) {
        // Every load of the segment is issued before any use: cw[m] holds the
        // coarse value at i0 + (2m - 5) h, xs[c] the original value of target c.
        constexpr int SEG = (int)kSweepSeg;
        const uint32_t cnt_s = S.cnt;
        const uint32_t i0 = h + 2 * h * S.k0;
        const int64_t hh = (int64_t)h;
        float cw[SEG + 5];
#pragma unroll
        for (int m = 0; m < SEG + 5; m++) {
            const int64_t c = (int64_t)i0 + (2 * m - 5) * hh;
            const bool need = QUIN ? m <= (int)cnt_s + 4 : (m >= 1 && m <= (int)cnt_s + 3);
            cw[m] = (need && c >= 0 && c < Dd) ? work[col + (uint64_t)c * sd] : 0.0f;
        }
        float xs[SEG];
        if (!DEC) {
#pragma unroll
            for (int c = 0; c < SEG; c++)
                xs[c] = c < (int)cnt_s ? work[col + (uint64_t)(i0 + 2 * h * c) * sd] : 0.0f;
        }
#pragma unroll
        for (int c = 0; c < SEG; c++) {
            if (c >= (int)cnt_s) break;
            const uint32_t i = i0 + 2 * h * c;
            float* rp = work + col + (uint64_t)i * sd;
            const float w[6] = {cw[c], cw[c + 1], cw[c + 2], cw[c + 3], cw[c + 4], cw[c + 5]};
            const float p = predict_w<QUIN>(i, h, right, w, t);
            if (DEC) {
                const int32_t q = WMODE == 1 ? tile[wave][c][lane] : qb[slot_of(c)];
                *rp = reconstruct(p, q, R.neg_two_eb);
            } else {
                const float x = xs[c];
                const int32_t q = quantize<AX>(p, x, R);
                const float xr = reconstruct(p, q, R.neg_two_eb);
                if (write_recon) *rp = xr;
                if (WMODE == 1) tile[wave][c][lane] = q;
                else qb[slot_of(c)] = q;
                const double e = (double)daz(x) - (double)xr;
                err = fmax(err, fabs(e));
                sq += e * e;
                qmin = min(qmin, q);
                qmax = max(qmax, q);
            }
        }
    }
    if (!DEC) {
        if (WMODE == 1) {
            __syncthreads();
            for (uint32_t e = lane; e < 64 * S.cnt; e += 64) {
                const uint32_t jl = e / S.cnt, c = e - jl * S.cnt;
                const uint32_t kl = blockIdx.x * blockDim.x + wave * 64 + jl;
                if (kl < R.n[2]) qb[lbase[wave][jl] + c] = tile[wave][c][jl];
            }
        }
        write_partial(err, sq, qmin, qmax, part + R.part_off);
    }
}

}  // namespace dev

// ------------------------------------------------------------------ launchers

void launch_range(const float* x, uint64_t n, uint32_t* d_out3, hipStream_t st) {
    uint64_t blocks = std::min<uint64_t>((n / 4 + 255) / 256 + 1, 2048);
    hipLaunchKernelGGL(dev::k_range, dim3((uint32_t)blocks), dim3(256), 0, st, x, n, d_out3);
}
void launch_zero_index(const float* x, uint64_t n, unsigned long long* d_out2, hipStream_t st) {
    uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(dev::k_zero_index, dim3((uint32_t)blocks), dim3(256), 0, st, x, n, d_out2);
}

void launch_pivots(bool compress, float* work, int32_t* qb, const PivotDesc& P, void* eacc, void* aacc,
                   hipStream_t st) {
    dim3 grid((P.pn[2] + 255) / 256, P.pn[1], P.pn[0]);
    if (compress)
        hipLaunchKernelGGL(dev::k_pivots_compress, grid, dim3(256), 0, st, work, qb, P, (ErrAcc*)eacc,
                           (ArrayAcc*)aacc);
    else
        hipLaunchKernelGGL(dev::k_pivots_decompress, grid, dim3(256), 0, st, work, (const int32_t*)qb, P);
}

void launch_reduce_partials(const Partial* part, const uint64_t* ranges, uint32_t rounds, DevAcc* acc,
                            hipStream_t st) {
    if (!rounds) return;
    hipLaunchKernelGGL(dev::k_reduce_partials, dim3(rounds), dim3(256), 0, st, part, ranges, acc);
}

uint64_t round_workgroups(const RoundDesc& R) {
    const uint64_t gx = (R.n[2] + 255) / 256;
    if (R.kind == kRoundSweep) return gx * R.n[1 - R.d] * R.nseg;
    return gx * R.n[1] * R.n[0];
}

namespace {

template <int AX, bool Q, bool DEC>
void direct(const RoundDesc& R, float* w, int32_t* qb, const uint32_t* tabs, const float* tblk, Partial* part, int wr,
            hipStream_t st) {
    dim3 grid((R.n[2] + 255) / 256, R.n[1], R.n[0]);
    hipLaunchKernelGGL((dev::k_round_direct<AX, Q, DEC>), grid, dim3(256), 0, st, R, w, qb, tabs, tblk, part, wr);
}

template <int AX, bool Q, bool DEC, int WM>
void sweep(const RoundDesc& R, float* w, int32_t* qb, const uint32_t* tabs, const float* tblk, const SweepSeg* segs,
           Partial* part, int wr, hipStream_t st) {
    dim3 grid((R.n[2] + 255) / 256, R.n[1 - AX], R.nseg);
    hipLaunchKernelGGL((dev::k_round_sweep<AX, Q, DEC, WM>), grid, dim3(256), 0, st, R, w, qb, tabs, tblk, segs,
                       part, wr);
}

template <bool DEC>
void dispatch(const RoundDesc& R, bool q, float* w, int32_t* qb, const uint32_t* tabs, const float* tblk,
              const SweepSeg* segs, Partial* part, int wr, hipStream_t st) {
    if (R.kind == kRoundDirect) {
        switch (R.d * 2 + (q ? 1 : 0)) {
            case 0: direct<0, false, DEC>(R, w, qb, tabs, tblk, part, wr, st); break;
            case 1: direct<0, true, DEC>(R, w, qb, tabs, tblk, part, wr, st); break;
            case 2: direct<1, false, DEC>(R, w, qb, tabs, tblk, part, wr, st); break;
            case 3: direct<1, true, DEC>(R, w, qb, tabs, tblk, part, wr, st); break;
            case 4: direct<2, false, DEC>(R, w, qb, tabs, tblk, part, wr, st); break;
            default: direct<2, true, DEC>(R, w, qb, tabs, tblk, part, wr, st); break;
        }
        return;
    }
    // sweep: write mode from the serialization order
    const int fastest = R.order == 2 ? 2 : (R.order == 0 ? 0 : 1);
    const int wm = fastest == 2 ? 0 : (fastest == (int)R.d ? 1 : 2);
    const int key = ((int)R.d * 2 + (q ? 1 : 0)) * 3 + wm;
    switch (key) {
        case 0: sweep<0, false, DEC, 0>(R, w, qb, tabs, tblk, segs, part, wr, st); break;
        case 1: sweep<0, false, DEC, 1>(R, w, qb, tabs, tblk, segs, part, wr, st); break;
        case 2: sweep<0, false, DEC, 2>(R, w, qb, tabs, tblk, segs, part, wr, st); break;
        case 3: sweep<0, true, DEC, 0>(R, w, qb, tabs, tblk, segs, part, wr, st); break;
        case 4: sweep<0, true, DEC, 1>(R, w, qb, tabs, tblk, segs, part, wr, st); break;
        case 5: sweep<0, true, DEC, 2>(R, w, qb, tabs, tblk, segs, part, wr, st); break;
        case 6: sweep<1, false, DEC, 0>(R, w, qb, tabs, tblk, segs, part, wr, st); break;
        case 7: sweep<1, false, DEC, 1>(R, w, qb, tabs, tblk, segs, part, wr, st); break;
        case 8: sweep<1, false, DEC, 2>(R, w, qb, tabs, tblk, segs, part, wr, st); break;
        case 9: sweep<1, true, DEC, 0>(R, w, qb, tabs, tblk, segs, part, wr, st); break;
        case 10: sweep<1, true, DEC, 1>(R, w, qb, tabs, tblk, segs, part, wr, st); break;
        default: sweep<1, true, DEC, 2>(R, w, qb, tabs, tblk, segs, part, wr, st); break;
    }
}

}  // namespace

void launch_round_compress(const RoundDesc& R, bool quintic, float* work, int32_t* qb, const uint32_t* tabs,
                           const float* tblk, const SweepSeg* segs, Partial* part, int write_recon, hipStream_t st) {
    if (R.count == 0) return;
    dispatch<false>(R, quintic, work, qb, tabs, tblk, segs, part, write_recon, st);
}

void launch_round_decompress(const RoundDesc& R, bool quintic, float* out, const int32_t* qb, const uint32_t* tabs,
                             const float* tblk, const SweepSeg* segs, hipStream_t st) {
    if (R.count == 0) return;
    dispatch<true>(R, quintic, out, const_cast<int32_t*>(qb), tabs, tblk, segs, nullptr, 1, st);
}

}  // namespace asz
