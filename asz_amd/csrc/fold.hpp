// Device side of the histogram folds: shared by k_fold_rounds / k_hist_tail
// (huffman.hip) and the launch that runs a round's fold beside the next
// round's t selection (tsel.hip).
#pragma once

#include "device.hpp"
#include "launch.hpp"

namespace asz {
namespace dev {

// Finishes array k for the host (whole workgroup): the tail, then the bins
// of [qmin, qmax] from the global window gw when they lie inside it, into the
// host slot; after a system-scope fence, the epoch into the slot's flag.
__device__ inline void publish_host(const uint32_t* gw, HistTail t, uint32_t* hslot, uint32_t epoch) {
    if (t.qmin <= t.qmax && t.qmin >= -kHistCentre && t.qmax < kHistCentre) {
        const uint32_t n = (uint32_t)(t.qmax - t.qmin + 1);
        const uint32_t* src = gw + (t.qmin + kHistCentre);
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) hslot[kHostSlotBins + i] = src[i];
    }
    if (threadIdx.x == 0) {
        uint32_t* o = hslot + kHostSlotHead;
        o[0] = (uint32_t)t.qmin;
        o[1] = (uint32_t)t.qmax;
        o[2] = t.fallbacks;
        o[3] = 0;
        const uint64_t e = (uint64_t)__double_as_longlong(t.max_err), q = (uint64_t)__double_as_longlong(t.sq);
        o[4] = (uint32_t)e;
        o[5] = (uint32_t)(e >> 32);
        o[6] = (uint32_t)q;
        o[7] = (uint32_t)(q >> 32);
    }
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence_system();
        __hip_atomic_store(hslot, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Workgroup-wide fold of Partials [p0, p1) (every thread gets the result).
__device__ inline Partial fold_partials(const Partial* __restrict__ part, uint64_t p0, uint64_t p1) {
    __shared__ double s_err[4], s_sq[4];
    __shared__ int32_t s_min[4], s_max[4];
    double err = 0, sq = 0;
    int32_t qmin = INT32_MAX, qmax = INT32_MIN;
    for (uint64_t i = p0 + threadIdx.x; i < p1; i += blockDim.x) {
        const Partial p = part[i];
        err = fmax(err, p.max_err);
        sq += p.sq;
        qmin = min(qmin, p.qmin);
        qmax = max(qmax, p.qmax);
    }
    err = wave_max(err);
    sq = wave_sum(sq);
    qmin = wave_min(qmin);
    qmax = wave_max(qmax);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) { s_err[w] = err; s_sq[w] = sq; s_min[w] = qmin; s_max[w] = qmax; }
    __syncthreads();
    Partial r{s_err[0], s_sq[0], s_min[0], s_max[0]};
    for (int k = 1; k < (int)(blockDim.x / kWave); k++) {
        r.max_err = fmax(r.max_err, s_err[k]);
        r.sq += s_sq[k];
        r.qmin = min(r.qmin, s_min[k]);
        r.qmax = max(r.qmax, s_max[k]);
    }
    __syncthreads();  // the shared slots may be reused by a later call
    return r;
}

// One fold workgroup grp of round r (one launch after a round kernel covers
// one round, or a batch of finished rounds; the pointers are the field's).
// grp < F.slices: bins [256 grp, 256 grp + 256) of the LDS window, summed
// over the round's F.copies histogram copies into the array's global window
// (coalesced along the bins).  Otherwise partial group pg = grp - F.slices:
// the Partials of round workgroups [pg kFoldPartials, (pg + 1) kFoldPartials)
// folded into fpart[F.fpart_off + pg].  The last workgroup of the round to
// finish (counted in acc->fold_count[r], which it resets) folds those into
// acc->arr[2 + r] and the window's tail.
__device__ inline void fold_group(const FoldDesc* __restrict__ descs, uint32_t r, uint32_t grp,
                                  const Partial* __restrict__ part, const uint32_t* __restrict__ phist,
                                  Partial* fpart, DevAcc* acc, uint32_t* ghist, uint32_t* hres, uint32_t epoch) {
    __shared__ int s_last;
    const FoldDesc F = descs[r];
    if (grp >= F.groups) return;
    uint32_t* gw = ghist + (uint64_t)(2 + r) * kHistStride;
    Partial* fp = fpart + F.fpart_off;
    if (grp < F.slices) {
        // bin slice grp: one bin per thread, summed over the copies; the
        // global window's bins inside the LDS window receive nothing else
        const uint32_t bin = grp * 256 + threadIdx.x;
        const uint32_t* src = phist + F.hist_off + bin;
        uint32_t s = 0;
#pragma unroll 8
        for (uint32_t k = 0; k < F.copies; k++) s += src[(uint64_t)k * F.win];
        gw[kHistCentre - F.win / 2 + bin] = s;
    } else {  // uniform per workgroup
        const uint32_t pg = grp - F.slices;
        const uint64_t p0 = (uint64_t)pg * kFoldPartials, p1 = min<uint64_t>(F.nwg, p0 + kFoldPartials);
        const Partial mine = fold_partials(part + F.part_off, p0, p1);
        if (threadIdx.x == 0) fp[pg] = mine;
    }
    __syncthreads();  // every wave's stores done before the release below
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(&acc->fold_count[r], 1u) == F.groups - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    const Partial all = fold_partials(fp, 0, F.pgroups);
    if (threadIdx.x == 0) {
        acc->rerr[r] = ErrAcc{all.max_err, all.sq};
        acc->arr[2 + r].qmin = all.qmin;
        acc->arr[2 + r].qmax = all.qmax;
        acc->fold_count[r] = 0;
    }
    publish_host(gw, HistTail{all.qmin, all.qmax, 0, 0, all.max_err, all.sq}, hres + (2 + r) * kHostSlotWords,
                 epoch);
}

}  // namespace dev
}  // namespace asz
