// Block rounds: a coarse round (at most kCoarseTargets targets per block)
// with its per-block t selection in one launch, one workgroup per block.
//
// The reference runs block after block, each through every round
// (compress.tpp:179-208); a round's targets read only coarse points that
// earlier rounds (or the pivots) finalised, so all blocks of a round can run
// at once after the previous round (the level-synchronous schedule of the
// other round kernels).  For the coarse rounds the per-round work is small
// and two launches per round (selection, then prediction) cost more than the
// work: here one workgroup selects its block's t from the block's samples
// (tsel.hpp, exactly as k_tsel) and then predicts, quantizes and
// reconstructs the block's targets of the round, leaving the block's partial
// and histogram window for the fold like any round kernel's workgroup.
#ifdef ASZ_COARSE_TRACE
#include <hip/hip_runtime.h>
// timing builds: per (round, block) phase marks of the block-round kernel (s_memrealtime)
__device__ unsigned long long g_tsel_trace[16 * 4096 * 8];
#define TSEL_MARK(k)                                          \
    if (threadIdx.x == 0 && blockIdx.x < 4096 && T.r < 16)    \
    g_tsel_trace[((unsigned long long)T.r * 4096 + blockIdx.x) * 8 + (k)] = __builtin_amdgcn_s_memrealtime()
#endif
#include "tsel.hpp"
#include "layout.hpp"

namespace asz {
namespace dev {

static_assert(kHistWinMax <= 2 * kMaxKeep, "the round's LDS histogram reuses the selection's key buffers");

constexpr int kBlockBatch = 6;      // targets per thread whose loads are issued together
constexpr int kBlockThreads = 512;  // 8 waves: two such workgroups (selection LDS ~66 KB each) share a CU

// n / d for n, d < 2^24 with rd = 1.0f / d: the float quotient is within one
// of the exact one, and one step corrects it.
__device__ __forceinline__ uint32_t udiv24(uint32_t n, uint32_t d, float rd) {
    uint32_t q = (uint32_t)((float)n * rd);
    const int32_t r = (int32_t)(n - q * d);
    if (r < 0) q--;
    else if ((uint32_t)r >= d) q++;
    return q;
}

// Grid: x = block, y = field; NT threads.  tsel: 0 takes each block's t from
// tblk, 1 selects it with the five-candidate scan, 2 with the gradient-tuned
// form.
template <bool QUIN, bool GRAD, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT / 128))) void k_round_block(TselDesc T, uint32_t tsel, float* work, int32_t* sym,
                                                    int32_t* qb, const uint32_t* __restrict__ tabs,
                                                    const uint32_t* __restrict__ ss, float* tblk, Partial* part,
                                                    uint32_t* phist, uint32_t* gwin, DevAcc* acc, FieldSet fs) {
    __shared__ TselLds<GRAD> L;
    const uint32_t b = blockIdx.x, f = blockIdx.y;
    const RoundDesc& R = T.R;
    // the field's layer bound (its range gives it; the plan's is field 0's)
    double inv_half_eb = R.inv_half_eb, eb_d = R.eb_d, neg_two_eb = R.neg_two_eb;
    float* wk = work;
    if (fs.fd) {  // field f of a multi-field (or device-range) call
        const FieldDesc& F = fs.fd[f];
        const FieldRound& fr = F.r[R.layer];
        inv_half_eb = fr.inv_half_eb;
        eb_d = fr.eb_d;
        neg_two_eb = fr.neg_two_eb;
        wk = F.work;
        qb += (uint64_t)f * fs.sym;
        part += (uint64_t)f * fs.part;
        phist += (uint64_t)f * fs.phist;
        gwin += (uint64_t)f * fs.ghist;
    }
    // the block's t: selected here (every sample it reads is final) or given
    float t;
    if (tsel) {
        tsel_block<GRAD, NT>(T, work, ss, tblk, sym, acc, fs, b, f, L);
        t = L.best;
        TSEL_MARK(6);
    } else {
        t = tblk[(uint64_t)f * fs.tblk + b];
    }
    // the round's LDS histogram window in the selection's key buffers
    uint32_t* lh = reinterpret_cast<uint32_t*>(L.sh.af);
    const uint32_t win = R.lds_win;
    for (uint32_t i = threadIdx.x; i < win; i += NT) lh[i] = 0;
    __syncthreads();

    // the block's targets: along d from bB + h in steps of 2h, along another
    // axis from 0 (block 0) or bB + s in steps of s (the plan's per-axis
    // counts and prefixes give the block's shape and QB base)
    const uint32_t d = R.d;
    const uint32_t bb[3] = {b / (R.nb[1] * R.nb[2]), (b / R.nb[2]) % R.nb[1], b % R.nb[2]};
    uint32_t cnt[3], pre[3], first[3], step[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        cnt[a] = tabs[R.tab_off[a] + bb[a]];
        pre[a] = tabs[R.tab_off[a] + R.nb[a] + bb[a]];
        if ((uint32_t)a == d) {
            first[a] = bb[a] * R.B[a] + R.h;
            step[a] = 2 * R.h;
        } else {
            first[a] = bb[a] == 0 ? 0 : bb[a] * R.B[a] + R.s[a];
            step[a] = R.s[a];
        }
    }
    const uint32_t right = (uint32_t)min<uint64_t>((uint64_t)(bb[d] + 1) * R.B[d] + 1, R.D[d]);
    const uint64_t base = R.qb_off + block_base(R, cnt, pre);
    const uint32_t c12 = cnt[1] * cnt[2];
    const uint32_t ntar = cnt[0] * c12;  // below 2^24: at most kCoarseTargets per block on average
    const float rc12 = 1.0f / (float)max(c12, 1u), rc2 = 1.0f / (float)max(cnt[2], 1u);
    const int64_t dist = (int64_t)R.dist, hh = R.h, Dd = R.D[d];
    // positions relative to the block's first target (below 2^31: the plan
    // makes block rounds only of blocks whose slab is that small)
    float* const wb = wk + ((uint64_t)first[0] * R.D[1] + first[1]) * R.D[2] + first[2];
    const uint32_t s0 = step[0] * R.D[1] * R.D[2], s1 = step[1] * R.D[2];
    const uint32_t first_d = d == 0 ? first[0] : (d == 1 ? first[1] : first[2]);
    const uint32_t step_d = d == 0 ? step[0] : (d == 1 ? step[1] : step[2]);
    double err = 0, sq = 0;
    int32_t qmin = INT32_MAX, qmax = INT32_MIN;
    for (uint32_t k0 = threadIdx.x; k0 < ntar; k0 += kBlockBatch * NT) {
        // every load of the batch first (the stores below touch only targets,
        // which no load of the round reads but their own); a target's field
        // position, QB slot and coordinate along d are kept for its update
        float w[kBlockBatch][6], x[kBlockBatch];
        uint32_t pos[kBlockBatch], slot[kBlockBatch], id[kBlockBatch];
#pragma unroll
        for (int u = 0; u < kBlockBatch; u++) {
            const uint32_t k = min(k0 + u * NT, ntar - 1);
            uint32_t j[3];
            j[0] = udiv24(k, c12, rc12);
            const uint32_t rem = k - j[0] * c12;
            j[1] = udiv24(rem, cnt[2], rc2);
            j[2] = rem - j[1] * cnt[2];
            id[u] = first_d + (d == 0 ? j[0] : (d == 1 ? j[1] : j[2])) * step_d;
            pos[u] = j[0] * s0 + j[1] * s1 + j[2] * step[2];
            slot[u] = (uint32_t)local_slot(R.order, j, cnt);
            // coarse window w[m] at id + (2m - 5) h; an entry outside the field
            // (which the stencil never uses) reads the target itself and is 0
#pragma unroll
            for (int m = 0; m < 6; m++) {
                if (!QUIN && (m == 0 || m == 5)) {
                    w[u][m] = 0.0f;
                    continue;
                }
                const int64_t c = (int64_t)id[u] + (2 * m - 5) * hh;
                const bool in = c >= 0 && c < Dd;
                const float v = wb[(int64_t)pos[u] + (in ? (2 * m - 5) * dist : 0)];
                w[u][m] = in ? v : 0.0f;
            }
            x[u] = wb[pos[u]];
        }
#pragma unroll
        for (int u = 0; u < kBlockBatch; u++) {
            if (k0 + u * NT >= ntar) break;
            const float p = predict_w<QUIN>(id[u], R.h, right, w[u], t);
            const int32_t q = d == 0 ? quantize_eb<0>(p, x[u], inv_half_eb, eb_d)
                                     : quantize_eb<1>(p, x[u], inv_half_eb, eb_d);
            const float xr = reconstruct(p, q, neg_two_eb);
            qb[base + slot[u]] = q;
            wb[pos[u]] = xr;
            const double e = (double)daz(x[u]) - (double)xr;
            err = fmax(err, fabs(e));
            sq += e * e;
            qmin = min(qmin, q);
            qmax = max(qmax, q);
            hist_add(lh, win, gwin, q);
        }
    }
    if (R.nwg) finish_wg<NT / kWave>(err, sq, qmin, qmax, R, b, part, lh, phist);
    TSEL_MARK(7);
}

}  // namespace dev

#ifdef ASZ_COARSE_TRACE
extern "C" int asz_coarse_trace(unsigned long long* out, uint64_t n) {
    return hipMemcpyFromSymbol(out, HIP_SYMBOL(g_tsel_trace), std::min<uint64_t>(n, 16 * 4096 * 8) * 8, 0,
                               hipMemcpyDeviceToHost) == hipSuccess ? 0 : 1;
}
#endif

void launch_round_block(const TselDesc& T, uint32_t tsel, bool quintic, float* work, int32_t* sym, int32_t* qb,
                        const uint32_t* tabs, const uint32_t* ss, float* tblk, Partial* part, uint32_t* phist,
                        uint32_t* gwin, DevAcc* acc, hipStream_t st, const FieldSet& fs, uint32_t nfields) {
    const uint64_t nbl = (uint64_t)T.R.nb[0] * T.R.nb[1] * T.R.nb[2];
    if (!nbl || !nfields) return;
    if (nbl > 0x7FFFFFFFull || nfields > 65535) throw Error(ASZ_ERR_ARG, "too many blocks for one launch");
    const bool grad = tsel == 2;
    using K = void (*)(TselDesc, uint32_t, float*, int32_t*, int32_t*, const uint32_t*, const uint32_t*, float*,
                       Partial*, uint32_t*, uint32_t*, DevAcc*, FieldSet);
    constexpr int NT = dev::kBlockThreads;
    const K k = quintic ? (grad ? dev::k_round_block<true, true, NT> : dev::k_round_block<true, false, NT>)
                        : (grad ? dev::k_round_block<false, true, NT> : dev::k_round_block<false, false, NT>);
    hipLaunchKernelGGL(k, dim3((uint32_t)nbl, nfields), dim3(NT), 0, st, T, tsel, work, sym, qb, tabs, ss, tblk, part,
                       phist, gwin, acc, fs);
}

}  // namespace asz
