// Per-block t selection (compress.tpp:90-144) and AB-key sampling for the
// default slope (get_default_slope.tpp:15-25).
//
// One 256-thread workgroup scores one block.  The per-sample arithmetic is the
// reference's, operation for operation (get_AB_map.tpp:18-21 and
// get_h_psnr.tpp:12-20 as GCC lowers them): A = (a+b)-(c+e),
// B = (float)fma(S, 0.25, -x0) with S = (b+e)+(a+c), keys = lroundf(A or B / 2 sigma),
// err = fmaf(Bk, 2, Ak*2t)*sigma, q = lroundf(err/2eb), de = err - q*2eb.
// The reference then sums over distinct keys in unordered_map order in float;
// here the sums run over samples in double (the same terms, count-weighted),
// so H1 and PSNR agree with the reference to float rounding and the chosen t
// agrees except where two candidates score within that rounding of each other.
#include <algorithm>

#include "device.hpp"
#include "fold.hpp"
#include "launch.hpp"

namespace asz {
namespace dev {

constexpr int kTselThreads = 256;
constexpr int kNCand = 5;          // block candidates {1/16, 1/4, 9/32, 5/16, 11/32} (asz_param.hpp:203)
constexpr int kDirect = 1024;      // direct LDS bins per candidate: QB in [-512, 512)
constexpr int kHashSlots = 512;    // per candidate, for QBs outside the direct bins
constexpr int kMaxKeep = 2048;     // samples whose keys stay in LDS; more are recomputed

// NC: candidates scored per block (5 for the discrete scan, 3 for the
// gradient-tuned form, whose smaller footprint lets more workgroups share a CU).
template <int NC>
struct TselShared {
    float af[kMaxKeep], bf[kMaxKeep];
    uint32_t bins[NC][kDirect];
    long long hkey[NC][kHashSlots];
    uint32_t hcnt[NC][kHashSlots];
    uint32_t ovf[NC];              // QBs that found no hash slot (counted as distinct symbols)
    double red[kTselThreads / kWave][2 * NC];
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
template <bool GRAD>
__device__ __forceinline__ void tsel_block(const TselDesc& T, const float* __restrict__ work,
                                           const uint32_t* __restrict__ ss_tab, float* tblk, int32_t* t_int,
                                           DevAcc* acc, const FieldSet& fs, uint32_t bid, uint32_t f) {
    constexpr int NC = GRAD ? 3 : kNCand;  // GRAD: 1/4 and the two grid points around the minimiser
    __shared__ TselShared<NC> sh;
    __shared__ float s_cand[NC];
    __shared__ int s_nc;
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
        for (int i = threadIdx.x; i < NC * kDirect; i += kTselThreads) (&sh.bins[0][0])[i] = 0;
        for (int i = threadIdx.x; i < NC * kHashSlots; i += kTselThreads) {
            (&sh.hkey[0][0])[i] = LLONG_MIN;
            (&sh.hcnt[0][0])[i] = 0;
        }
        if (threadIdx.x < NC) sh.ovf[threadIdx.x] = 0;
        double saa = 0, sab = 0;  // GRAD: sum A^2, sum A B over the samples
        // four samples per thread at a time: all twenty loads in flight before any use
        const uint32_t ns_pass = GRAD ? ns : min(ns, (uint32_t)kMaxKeep);
        for (uint32_t idx0 = threadIdx.x; idx0 < ns_pass; idx0 += 4 * kTselThreads) {
            float v[4][5];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t idx = idx0 + u * kTselThreads;
                if (idx < ns_pass) sample_load(T, work, lo, step, cnt, idx, v[u]);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t idx = idx0 + u * kTselThreads;
                if (idx >= ns_pass) break;
                float af, bf, A, B;
                sample_from(v[u], neg2, af, bf, A, B);
                if (idx < kMaxKeep) { sh.af[idx] = af; sh.bf[idx] = bf; }
                if (GRAD) { saa += (double)A * (double)A; sab += (double)A * (double)B; }
            }
        }
        if (GRAD) {
            const int w0 = threadIdx.x / kWave;
            saa = wave_sum(saa);
            sab = wave_sum(sab);
            if ((threadIdx.x & (kWave - 1)) == 0) { sh.red[w0][0] = saa; sh.red[w0][1] = sab; }
            __syncthreads();
            if (threadIdx.x == 0) {
                double a2 = 0, ab = 0;
                for (int v = 0; v < kTselThreads / kWave; v++) { a2 += sh.red[v][0]; ab += sh.red[v][1]; }
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
        for (uint32_t idx = threadIdx.x; idx < ns; idx += kTselThreads) {
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
        const float dn = (float)ns;
        const float rec = 1.0f / dn;
        double h[NC] = {};
#pragma unroll
        for (int k = 0; k < NC; k++) {
            for (int i = threadIdx.x; i < kDirect; i += kTselThreads)
                if (sh.bins[k][i]) h[k] -= entropy_term(sh.bins[k][i], rec);
            for (int i = threadIdx.x; i < kHashSlots; i += kTselThreads)
                if (sh.hcnt[k][i]) h[k] -= entropy_term(sh.hcnt[k][i], rec);
            if (threadIdx.x == 0 && sh.ovf[k]) h[k] -= (double)sh.ovf[k] * entropy_term(1, rec);
        }
        const int w = threadIdx.x / kWave;
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const double hs = wave_sum(h[k]), ms = wave_sum(mse[k]);
            if ((threadIdx.x & (kWave - 1)) == 0) { sh.red[w][k] = hs; sh.red[w][NC + k] = ms; }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            float h1[NC], psnr[NC];
            const float r2 = range * range;
            for (int k = 0; k < nc; k++) {
                double hs = 0, ms = 0;
                for (int v = 0; v < kTselThreads / kWave; v++) { hs += sh.red[v][k]; ms += sh.red[v][NC + k]; }
                h1[k] = (float)hs;
                psnr[k] = log10f((r2 / (float)ms) * dn) * 10.0f;
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
    }
}

template <bool GRAD>
__global__ __launch_bounds__(kTselThreads) void k_tsel(TselDesc T, const float* __restrict__ work,
                                                       const uint32_t* __restrict__ ss_tab, float* tblk,
                                                       int32_t* t_int, DevAcc* acc, FieldSet fs) {
    tsel_block<GRAD>(T, work, ss_tab, tblk, t_int, acc, fs, blockIdx.x, blockIdx.y);
}

// The fold of rounds r0 .. r0 + nr - 1 beside round T.r's t selection, in one
// launch: the two are independent (the fold reads the finished rounds'
// partials, the selection reads the field), so the fold no longer adds a
// launch of its own to the chain of rounds.  Grid x: the nbl blocks' t
// selection, then mg x nr fold workgroups; grid y: fields.
template <bool GRAD>
__global__ __launch_bounds__(kTselThreads) void k_fold_tsel(TselDesc T, const float* __restrict__ work,
                                                            const uint32_t* __restrict__ ss_tab, float* tblk,
                                                            int32_t* t_int, DevAcc* acc, FieldSet fs, uint32_t nbl,
                                                            const FoldDesc* __restrict__ descs, uint32_t r0,
                                                            uint32_t mg, const Partial* __restrict__ part,
                                                            const uint32_t* __restrict__ phist, Partial* fpart,
                                                            uint32_t* ghist, uint32_t* hres, uint32_t epoch) {
    const uint32_t f = blockIdx.y;
    if (blockIdx.x < nbl) {
        tsel_block<GRAD>(T, work, ss_tab, tblk, t_int, acc, fs, blockIdx.x, f);
        return;
    }
    const uint32_t g = blockIdx.x - nbl;
    if (fs.fd) {
        part += f * fs.part;
        phist += f * fs.phist;
        fpart += f * fs.fpart;
        acc += f;
        ghist += f * fs.ghist;
        hres += f * fs.hres;
    }
    fold_group(descs, r0 + g / mg, g % mg, part, phist, fpart, acc, ghist, hres, epoch);
}

// Keys of every sample of every block, in the reference's loop order, for
// the host scoring of ASZ_TSEL_EXACT (compress.tpp:102-120).
__global__ __launch_bounds__(256) void k_tsel_keys(TselDesc T, const float* __restrict__ work,
                                                   const uint32_t* __restrict__ ss_tab,
                                                   const uint64_t* __restrict__ key_off, long long* keys) {
    const RoundDesc& R = T.R;
    const uint32_t bid = blockIdx.x;
    const uint32_t ss = ss_tab[bid];
    if (ss == 0) return;
    const uint32_t bb[3] = {bid / (R.nb[2] * R.nb[1]), (bid / R.nb[2]) % R.nb[1], bid % R.nb[2]};
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
    const float neg2 = 0.5f / (T.stat_eb_ratio * R.eb);
    long long* out = keys + 2 * key_off[bid];
    for (uint32_t idx = threadIdx.x; idx < ns; idx += blockDim.x) {
        uint32_t q2 = idx % cnt[2], q01 = idx / cnt[2], q1 = q01 % cnt[1], q0 = q01 / cnt[1];
        uint64_t pos = ((uint64_t)(lo[0] + q0 * step[0]) * R.D[1] + (lo[1] + q1 * step[1])) * R.D[2] +
                       (lo[2] + q2 * step[2]);
        const int64_t dist = (int64_t)R.dist;
        const float* p = work + pos;
        float a = p[-dist], c = p[-3 * dist], b = p[dist], e = p[3 * dist];
        float x0 = daz(p[0]);
        float A = (a + b) - (c + e);
        float S = (b + e) + (a + c);
        float B = (float)__fma_rn((double)S, 0.25, -(double)x0);
        out[2 * idx] = lround_f64(neg2 * A);
        out[2 * idx + 1] = lround_f64(neg2 * B);
    }
}

// Keys (A_key, B_key) of sample idx, in the reference's loop order (i0, i1, i2).
__device__ __forceinline__ void slope_key(const SlopeSampleDesc& S, float scale, const float* __restrict__ f,
                                          uint64_t idx, long long* keys) {
    uint32_t q2 = idx % S.cnt[2];
    uint64_t q01 = idx / S.cnt[2];
    uint32_t q1 = q01 % S.cnt[1];
    uint32_t q0 = q01 / S.cnt[1];
    uint64_t pos = ((uint64_t)(S.lo[0] + q0 * S.step[0]) * S.D[1] + (S.lo[1] + q1 * S.step[1])) * S.D[2] +
                   (S.lo[2] + q2 * S.step[2]);
    const int64_t dist = (int64_t)S.dist;
    const float* p = f + pos;
    float a = daz(p[-dist]), c = daz(p[-3 * dist]), b = daz(p[dist]), e = daz(p[3 * dist]);
    float x0 = daz(p[0]);
    float A = (a + b) - (c + e);
    float Ssum = (b + e) + (a + c);
    float B = (float)__fma_rn((double)Ssum, 0.25, -(double)x0);
    keys[2 * idx] = lround_f64(scale * A);
    keys[2 * idx + 1] = lround_f64(scale * B);
}

__global__ __launch_bounds__(256) void k_slope_sample(SlopeSampleDesc S, const float* __restrict__ f,
                                                      long long* keys) {
    const uint64_t ns = (uint64_t)S.cnt[0] * S.cnt[1] * S.cnt[2];
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ns) return;
    slope_key(S, S.neg_2_stat_eb, f, idx, keys);
}

// The three axes' samples of up to kSlopeMaxFields fields of one shape: grid
// y = axis, z = field; field f's keys of axis d at keys + f per + off[d].
__global__ __launch_bounds__(256) void k_slope_sample_fields(SlopeFieldsDesc B, long long* keys) {
    const uint32_t d = blockIdx.y, f = blockIdx.z;
    const SlopeSampleDesc& S = B.S[d];
    const uint64_t ns = (uint64_t)S.cnt[0] * S.cnt[1] * S.cnt[2];
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ns) return;
    slope_key(S, B.scale[f], B.f[f], idx, keys + f * B.per + B.off[d]);
}

}  // namespace dev

void launch_tsel(const TselDesc& T, const float* work, const uint32_t* ss, float* tblk, int32_t* t_int, DevAcc* acc,
                 uint64_t nblocks, bool grad, hipStream_t st, const FieldSet& fs, uint32_t nfields) {
    const dim3 grid((uint32_t)nblocks, nfields);
    if (grad)
        hipLaunchKernelGGL(dev::k_tsel<true>, grid, dim3(dev::kTselThreads), 0, st, T, work, ss, tblk, t_int, acc, fs);
    else
        hipLaunchKernelGGL(dev::k_tsel<false>, grid, dim3(dev::kTselThreads), 0, st, T, work, ss, tblk, t_int, acc,
                           fs);
}

void launch_fold_tsel(const TselDesc& T, const float* work, const uint32_t* ss, float* tblk, int32_t* t_int,
                      DevAcc* acc, uint64_t nblocks, bool grad, const FoldDesc* descs, uint32_t r0, uint32_t nr,
                      uint32_t max_groups, const Partial* part, const uint32_t* phist, Partial* fpart,
                      uint32_t* ghist, uint32_t* hres, uint32_t epoch, hipStream_t st, const FieldSet& fs,
                      uint32_t nfields) {
    const uint64_t nx = nblocks + (uint64_t)max_groups * nr;
    if (nx > 0x7FFFFFFFull) throw Error(ASZ_ERR_ARG, "field dimensions exceed the launch grid");
    const dim3 grid((uint32_t)nx, nfields);
    if (grad)
        hipLaunchKernelGGL(dev::k_fold_tsel<true>, grid, dim3(dev::kTselThreads), 0, st, T, work, ss, tblk, t_int,
                           acc, fs, (uint32_t)nblocks, descs, r0, max_groups, part, phist, fpart, ghist, hres,
                           epoch);
    else
        hipLaunchKernelGGL(dev::k_fold_tsel<false>, grid, dim3(dev::kTselThreads), 0, st, T, work, ss, tblk, t_int,
                           acc, fs, (uint32_t)nblocks, descs, r0, max_groups, part, phist, fpart, ghist, hres,
                           epoch);
}

void launch_tsel_keys(const TselDesc& T, const float* work, const uint32_t* ss, const uint64_t* key_off,
                      int64_t* keys, uint64_t nblocks, hipStream_t st) {
    hipLaunchKernelGGL(dev::k_tsel_keys, dim3((uint32_t)nblocks), dim3(256), 0, st, T, work, ss, key_off,
                       (long long*)keys);
}

void launch_slope_sample_fields(const SlopeFieldsDesc& B, int64_t* keys, hipStream_t st) {
    uint64_t nx = 0;
    for (int d = 0; d < 3; d++) nx = std::max<uint64_t>(nx, (uint64_t)B.S[d].cnt[0] * B.S[d].cnt[1] * B.S[d].cnt[2]);
    if (!nx || !B.nf) return;
    if (B.nf > kSlopeMaxFields) throw Error(ASZ_ERR_ARG, "too many fields for one slope-sample launch");
    hipLaunchKernelGGL(dev::k_slope_sample_fields, dim3((uint32_t)((nx + 255) / 256), 3, B.nf), dim3(256), 0, st, B,
                       (long long*)keys);
}

void launch_slope_sample(const SlopeSampleDesc& S, const float* field, int64_t* keys, hipStream_t st) {
    uint64_t ns = (uint64_t)S.cnt[0] * S.cnt[1] * S.cnt[2];
    if (!ns) return;
    hipLaunchKernelGGL(dev::k_slope_sample, dim3((uint32_t)((ns + 255) / 256)), dim3(256), 0, st, S, field,
                       (long long*)keys);
}

}  // namespace asz
