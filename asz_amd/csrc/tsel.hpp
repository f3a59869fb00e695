// Device code of the per-block t selection (compress.tpp:90-144), shared by
// the per-round selection kernels (tsel.hip) and the block-round kernel
// (coarse.hip).
//
// One workgroup (256 threads, or 512 in a block round) scores one block.
// The per-sample arithmetic is the reference's, operation for operation
// (get_AB_map.tpp:18-21 and get_h_psnr.tpp:12-20 as GCC lowers them):
// A = (a+b)-(c+e),
// B = (float)fma(S, 0.25, -x0) with S = (b+e)+(a+c), keys = lroundf(A or B / 2 sigma),
// err = fmaf(Bk, 2, Ak*2t)*sigma, q = lroundf(err/2eb), de = err - q*2eb.
// The reference then sums over distinct keys in unordered_map order in float;
// here the sums run over samples in double (the same terms, count-weighted),
// so H1 and PSNR agree with the reference to float rounding and the chosen t
// agrees except where two candidates score within that rounding of each other.
#pragma once

#include <algorithm>
#include <climits>

#include "device.hpp"
#include "launch.hpp"

// Phase marks for timing builds (the includer defines TSEL_MARK; none here).
#ifndef TSEL_MARK
#define TSEL_MARK(k)
#endif

namespace asz {
namespace dev {

constexpr int kTselThreads = 256;  // the selection kernels' workgroup
constexpr int kTselMaxWaves = 8;   // the widest workgroup that runs a selection: 512 threads
constexpr int kNCand = 5;          // block candidates {1/16, 1/4, 9/32, 5/16, 11/32} (asz_param.hpp:203)
constexpr int kDirect = 1024;      // direct LDS bins per candidate: QB in [-512, 512)
constexpr int kHashSlots = 512;    // per candidate, for QBs outside the direct bins
constexpr int kMaxKeep = 2048;     // samples whose keys stay in LDS; more are recomputed

// NC: candidates scored per block (5 for the discrete scan, 3 for the
// gradient-tuned form, whose smaller footprint lets more workgroups share a CU).
template <int NC>
struct alignas(16) TselShared {
    float af[kMaxKeep], bf[kMaxKeep];
    uint32_t bins[NC][kDirect];
    long long hkey[NC][kHashSlots];
    uint32_t hcnt[NC][kHashSlots];
    uint32_t ovf[NC];              // QBs that found no hash slot (counted as distinct symbols)
    double red[kTselMaxWaves][2 * NC];
};

// The five values sample idx reads: a = x[-h], c = x[-3h], b = x[+h], e = x[+3h], x0.
__device__ __forceinline__ void sample_load(const TselDesc& T, const float* __restrict__ work, const uint32_t lo[3],
                                            const uint32_t step[3], const uint32_t cnt[3], uint32_t idx,
                                            float v[5]) {
    const RoundDesc& R = T.R;
    uint32_t q2 = idx % cnt[2];
    uint32_t q01 = idx / cnt[2];
    uint32_t q1 = q01 % cnt[1];
    uint32_t q0 = q01 / cnt[1];
    uint64_t pos = ((uint64_t)(lo[0] + q0 * step[0]) * R.D[1] + (lo[1] + q1 * step[1])) * R.D[2] + (lo[2] + q2 * step[2]);
    const int64_t dist = (int64_t)R.dist;
    const gptr<const float> p = to_global(work + pos);
    v[0] = p[-dist];
    v[1] = p[-3 * dist];
    v[2] = p[dist];
    v[3] = p[3 * dist];
    v[4] = p[0];
}

// Keys and (A, B) of a sample from its five values.
__device__ __forceinline__ void sample_from(const float v[5], float neg2, float& af, float& bf, float& Araw,
                                            float& Braw) {
    const float a = v[0], c = v[1], b = v[2], e = v[3];
    const float x0 = daz(v[4]);
    float A = (a + b) - (c + e);
    float S = (b + e) + (a + c);
    float B = (float)__fma_rn((double)S, 0.25, -(double)x0);
    af = (float)lround_f64(neg2 * A);
    bf = (float)lround_f64(neg2 * B);
    Araw = A;
    Braw = B;
}

// Keys (lroundf(A / 2 sigma), lroundf(B / 2 sigma)) of sample idx; A and B
// themselves in Araw / Braw (the residual of the interior 1D4R predictor at
// t is A t + B).
__device__ __forceinline__ void sample_keys(const TselDesc& T, const float* __restrict__ work, const uint32_t lo[3],
                                            const uint32_t step[3], const uint32_t cnt[3], uint32_t idx,
                                            float neg2, float& af, float& bf, float& Araw, float& Braw) {
    float v[5];
    sample_load(T, work, lo, step, cnt, idx, v);
    sample_from(v, neg2, af, bf, Araw, Braw);
}

template <int NC>
__device__ __forceinline__ void count_qb(TselShared<NC>& sh, int k, long long q) {
    if (q >= -kDirect / 2 && q < kDirect / 2) {
        atomicAdd(&sh.bins[k][(int)q + kDirect / 2], 1u);
        return;
    }
    uint32_t hsh = (uint32_t)((unsigned long long)q * 0x9E3779B97F4A7C15ull >> 55) & (kHashSlots - 1);
    for (int probe = 0; probe < kHashSlots; probe++) {
        const uint32_t slot = (hsh + probe) & (kHashSlots - 1);
        const long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(&sh.hkey[k][slot]),
                                         (unsigned long long)LLONG_MIN, (unsigned long long)q);
        if (prev == LLONG_MIN || prev == q) {
            atomicAdd(&sh.hcnt[k][slot], 1u);
            return;
        }
    }
    atomicAdd(&sh.ovf[k], 1u);
}

__device__ __forceinline__ double entropy_term(uint32_t cnt, float rec) {
    const float p = (float)cnt * rec;
    return (double)p * (double)log2f(p);
}

// One workgroup per block: samples the block once (compress.tpp:102-120,
// get_AB_map.tpp:9-22) and scores all candidates in the same pass over the
// samples (get_h_psnr.tpp:12-30, one QB histogram per candidate), then
// replays the reference's sequential acceptance rule (compress.tpp:127-142)
// on the (H1, PSNR) pairs, starting from t = 1/4.
//
// GRAD = false: the five block candidates {1/16, 1/4, 9/32, 5/16, 11/32}.
// GRAD = true (gradient-tuned form): the sampled residual energy
// E(t) = sum (A t + B)^2 is minimised by projected Newton steps on
// [t_min, t_max] (E is quadratic: dE/dt = 2 sum A (A t + B), d2E/dt2 =
// 2 sum A^2, so the step is exact), the minimiser is snapped to the two
// neighbouring points of the 1/t_precision grid, and {floor, 1/4, ceil}
// are judged by the exact proxy.  Any t on the grid decodes with the
// reference decoder (decompress.tpp:185).
// Block bid of field f (f = 0 unless fs.fd is set).
// The workgroup's LDS for one block's selection (the calling kernel declares
// it); best holds the chosen t for every thread after the selection.
template <bool GRAD>
struct TselLds {
    static constexpr int NC = GRAD ? 3 : kNCand;  // GRAD: 1/4 and the two grid points around the minimiser
    TselShared<NC> sh;
    float s_cand[NC];
    float h1[NC], psnr[NC];
    int s_nc;
    float best;
};

template <bool GRAD, int NT = kTselThreads>
__device__ __forceinline__ void tsel_block(const TselDesc& T, const float* __restrict__ work,
                                           const uint32_t* __restrict__ ss_tab, float* tblk, int32_t* t_int,
                                           DevAcc* acc, const FieldSet& fs, uint32_t bid, uint32_t f,
                                           TselLds<GRAD>& L) {
    constexpr int NC = TselLds<GRAD>::NC;
    static_assert(NT % kWave == 0 && NT / kWave <= kTselMaxWaves, "selection workgroup shape");
    TselShared<NC>& sh = L.sh;
    float* s_cand = L.s_cand;
    int& s_nc = L.s_nc;
    const RoundDesc& R = T.R;
    // the field's layer error bound, layer slope and value range (kept out of
    // T, so that T stays in kernel arguments)
    float eb = R.eb, layer_slope = R.layer_slope, range = T.range;
    if (fs.fd) {  // field f of a multi-field call
        const gptr<const FieldDesc> F = to_global(fs.fd + f);
        work = F->work;
        tblk += f * fs.tblk;
        t_int += f * fs.sym;
        acc += f;
        eb = F->r[R.layer].eb;
        layer_slope = F->r[R.layer].layer_slope;
        range = F->range;
    }
    const uint32_t b2 = bid % R.nb[2], b1 = (bid / R.nb[2]) % R.nb[1], b0 = bid / (R.nb[2] * R.nb[1]);
    const uint32_t bb[3] = {b0, b1, b2};
    const uint32_t ss = ss_tab[bid];
    TSEL_MARK(0);
    float best_t = 0.25f;  // insufficient_sample_t (asz_param.hpp:167)
    const bool fallback = ss == 0;
    if (!fallback) {
        const int d = (int)R.d;
        uint32_t lo[3], hi[3], step[3], cnt[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = bb[a] * R.B[a];
            hi[a] = min(lo[a] + R.B[a] + 1, R.D[a]);
            if (a != d && lo[a] != 0) lo[a] += R.s[a];
            step[a] = R.s[a] * ss;
        }
        lo[d] += 3 * R.s[d] / 2;
        hi[d] -= 3 * R.s[d] / 2;
        uint32_t ns = 1;
        for (int a = 0; a < 3; a++) {
            cnt[a] = (hi[a] - lo[a] + step[a] - 1) / step[a];
            ns *= cnt[a];
        }
        const float stat_eb = T.stat_eb_ratio * eb;
        const float neg2 = 0.5f / stat_eb;
        const float neg_2_eb = 0.5f / eb;
        const float two_eb = eb * 2.0f;
        // 16-byte LDS stores: the direct bins and hash counts to 0, the hash keys to LLONG_MIN
        {
            uint4* b4 = reinterpret_cast<uint4*>(&sh.bins[0][0]);
            for (int i = threadIdx.x; i < NC * kDirect / 4; i += NT) b4[i] = make_uint4(0, 0, 0, 0);
            uint4* k4 = reinterpret_cast<uint4*>(&sh.hkey[0][0]);
            for (int i = threadIdx.x; i < NC * kHashSlots / 2; i += NT) k4[i] = make_uint4(0, 0x80000000u, 0, 0x80000000u);
            uint4* c4 = reinterpret_cast<uint4*>(&sh.hcnt[0][0]);
            for (int i = threadIdx.x; i < NC * kHashSlots / 4; i += NT) c4[i] = make_uint4(0, 0, 0, 0);
        }
        if (threadIdx.x < NC) sh.ovf[threadIdx.x] = 0;
        TSEL_MARK(1);
        double saa = 0, sab = 0;  // GRAD: sum A^2, sum A B over the samples
        // four samples per thread at a time: all twenty loads in flight before any use
        const uint32_t ns_pass = GRAD ? ns : min(ns, (uint32_t)kMaxKeep);
        for (uint32_t idx0 = threadIdx.x; idx0 < ns_pass; idx0 += 4 * NT) {
            float v[4][5];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t idx = idx0 + u * NT;
                if (idx < ns_pass) sample_load(T, work, lo, step, cnt, idx, v[u]);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t idx = idx0 + u * NT;
                if (idx >= ns_pass) break;
                float af, bf, A, B;
                sample_from(v[u], neg2, af, bf, A, B);
                if (idx < kMaxKeep) { sh.af[idx] = af; sh.bf[idx] = bf; }
                if (GRAD) { saa += (double)A * (double)A; sab += (double)A * (double)B; }
            }
        }
        TSEL_MARK(2);
        if (GRAD) {
            const int w0 = threadIdx.x / kWave;
            saa = wave_sum(saa);
            sab = wave_sum(sab);
            if ((threadIdx.x & (kWave - 1)) == 0) { sh.red[w0][0] = saa; sh.red[w0][1] = sab; }
            __syncthreads();
            if (threadIdx.x == 0) {
                double a2 = 0, ab = 0;
                for (int v = 0; v < NT / kWave; v++) { a2 += sh.red[v][0]; ab += sh.red[v][1]; }
                // projected Newton on [t_min, t_max] from t = 1/4
                double t = 0.25;
                for (int it = 0; it < 3 && a2 > 0; it++) {
                    const double g = 2.0 * (a2 * t + ab), hs = 2.0 * a2;
                    t = fmin(fmax(t - g / hs, (double)T.t_min), (double)T.t_max);
                }
                const float P = T.t_precision;
                float lo_t = floorf((float)t * P) / P, hi_t = ceilf((float)t * P) / P;
                lo_t = fmaxf(lo_t, T.t_min);
                hi_t = fminf(hi_t, T.t_max);
                int nc = 0;
                s_cand[nc++] = 0.25f;
                if (lo_t != 0.25f) s_cand[nc++] = lo_t;
                if (hi_t != 0.25f && hi_t != lo_t) s_cand[nc++] = hi_t;
                s_nc = nc;
            }
        } else if (threadIdx.x == 0) {
            const float c5[kNCand] = {1.0f / 16, 1.0f / 4, 9.0f / 32, 5.0f / 16, 11.0f / 32};
            for (int k = 0; k < NC; k++) s_cand[k] = c5[k];
            s_nc = NC;
        }
        __syncthreads();
        const int nc = s_nc;
        float cand[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) cand[k] = k < nc ? s_cand[k] : 0.0f;
        double mse[NC] = {};
        for (uint32_t idx = threadIdx.x; idx < ns; idx += NT) {
            float af, bf, A, B;
            if (idx < kMaxKeep) { af = sh.af[idx]; bf = sh.bf[idx]; }
            else sample_keys(T, work, lo, step, cnt, idx, neg2, af, bf, A, B);
#pragma unroll
            for (int k = 0; k < NC; k++) {
                if (k >= nc) break;
                const float err = __fmaf_rn(bf, 2.0f, af * (cand[k] * 2.0f)) * stat_eb;
                const long long q = lround_f64(err * neg_2_eb);
                const float de = __fmaf_rn(-(float)q, two_eb, err);
                mse[k] += (double)de * (double)de;
                count_qb(sh, k, q);
            }
        }
        __syncthreads();
        TSEL_MARK(3);
        const float dn = (float)ns;
        const float rec = 1.0f / dn;
        // the first 256 threads: four direct bins and two hash slots per
        // candidate each (16- and 8-byte LDS reads)
        static_assert(kDirect == 4 * 256 && kHashSlots == 2 * 256 && NT >= 256, "entropy read shape");
        double h[NC] = {};
        if (threadIdx.x < 256) {
#pragma unroll
            for (int k = 0; k < NC; k++) {
                const uint4 d4 = reinterpret_cast<const uint4*>(&sh.bins[k][0])[threadIdx.x];
                const uint2 c2 = reinterpret_cast<const uint2*>(&sh.hcnt[k][0])[threadIdx.x];
                const uint32_t cs[6] = {d4.x, d4.y, d4.z, d4.w, c2.x, c2.y};
#pragma unroll
                for (int e = 0; e < 6; e++)
                    if (cs[e]) h[k] -= entropy_term(cs[e], rec);
                if (threadIdx.x == 0 && sh.ovf[k]) h[k] -= (double)sh.ovf[k] * entropy_term(1, rec);
            }
        }
        const int w = threadIdx.x / kWave;
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const double hs = wave_sum(h[k]), ms = wave_sum(mse[k]);
            if ((threadIdx.x & (kWave - 1)) == 0) { sh.red[w][k] = hs; sh.red[w][NC + k] = ms; }
        }
        __syncthreads();
        TSEL_MARK(4);
        // lane k of the first wave: candidate k's H1 and PSNR
        if (threadIdx.x < (uint32_t)nc) {
            const int k = threadIdx.x;
            double hs = 0, ms = 0;
            for (int v = 0; v < NT / kWave; v++) { hs += sh.red[v][k]; ms += sh.red[v][NC + k]; }
            L.h1[k] = (float)hs;
            L.psnr[k] = log10f((range * range / (float)ms) * dn) * 10.0f;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            float h1[NC], psnr[NC];
            for (int k = 0; k < nc; k++) {
                h1[k] = L.h1[k];
                psnr[k] = L.psnr[k];
            }
            // start from 1/4, then the candidates in order (compress.tpp:127-142)
            int best = GRAD ? 0 : 1;
            for (int k = 0; k < nc; k++) {
                if (cand[k] == cand[best] || psnr[best] > FLT_MAX) continue;
                const float sc = __fmaf_rn(h1[best] - h1[k], layer_slope, psnr[k] - psnr[best]);
                if (sc > 0.0f || psnr[k] > FLT_MAX) best = k;
            }
            best_t = cand[best];
        }
    }
    if (threadIdx.x == 0) {
        tblk[bid] = best_t;
        t_int[(uint64_t)bid * T.L + T.r] = (int32_t)lround_f64(best_t * T.t_precision);
        if (fallback) atomicAdd(&acc->fallbacks, 1u);
        L.best = best_t;
    }
    __syncthreads();
    TSEL_MARK(5);
}

}  // namespace dev
}  // namespace asz
