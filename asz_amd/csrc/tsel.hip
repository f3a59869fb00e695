// Per-block t selection kernels (compress.tpp:90-144; the per-block scoring
// itself is in tsel.hpp) and AB-key sampling for the exact host scoring and
// the default slope (get_default_slope.tpp:15-25).
#include <algorithm>

#include "fold.hpp"
#include "tsel.hpp"

namespace asz {
namespace dev {

template <bool GRAD>
__global__ __launch_bounds__(kTselThreads) void k_tsel(TselDesc T, const float* __restrict__ work,
                                                       const uint32_t* __restrict__ ss_tab, float* tblk,
                                                       int32_t* t_int, DevAcc* acc, FieldSet fs) {
    __shared__ TselLds<GRAD> L;
    tsel_block<GRAD>(T, work, ss_tab, tblk, t_int, acc, fs, blockIdx.x, blockIdx.y, L);
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
        __shared__ TselLds<GRAD> L;
        tsel_block<GRAD>(T, work, ss_tab, tblk, t_int, acc, fs, blockIdx.x, f, L);
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
