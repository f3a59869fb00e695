// Target -> (block, local index) -> QB slot mapping of one round, shared by
// the round kernels and the host-side layout query (asz_round_layout).
//
// Along the round axis d, targets are i = h + 2h k; block b = i / B holds the
// targets in [bB, (b+1)B).  Along another axis, targets are i = s k; block
// b > 0 owns (bB, (b+1)B] (its lower face belongs to block b-1), block 0 owns
// [0, B].  This is the ownership of compress.tpp:57-59 with the blocks of
// compress.tpp:183-188.  The slot is the block-major concatenation
// (compress.tpp:61-62) of each block's order-permuted array
// (asz_param.hpp:122-129).
#pragma once

#include <cstdint>

#include "internal.hpp"

#if defined(__HIP__)
#define ASZ_HD __host__ __device__ __forceinline__
#else
#define ASZ_HD inline
#endif

namespace asz {

struct TargetLoc {
    uint32_t b[3], j[3];
};

// Block b and local index j of lattice coordinate i = s k along an axis that
// is not the round axis: block b > 0 owns (bB, (b+1)B], block 0 owns [0, B].
ASZ_HD void side_loc(uint32_t i, uint32_t k, uint32_t B, uint32_t s, uint32_t& b, uint32_t& j) {
    if (i == 0) {
        b = 0;
        j = 0;
        return;
    }
    b = (i - 1) / B;
    j = b == 0 ? k : k - b * (B / s) - 1;
}

ASZ_HD void locate(const RoundDesc& R, int ax, const uint32_t k[3], const uint32_t i[3], TargetLoc& L) {
    for (int a = 0; a < 3; a++) {
        if (a == ax) {
            uint32_t b = i[a] / R.B[a];
            L.b[a] = b;
            L.j[a] = k[a] - b * (R.B[a] / (2 * R.h));
        } else {
            side_loc(i[a], k[a], R.B[a], R.s[a], L.b[a], L.j[a]);
        }
    }
}

// First slot of block (b0,b1,b2) in the round's array, from the per-axis
// target counts cnt[a] and their exclusive prefixes pre[a] at that block.
ASZ_HD uint64_t block_base(const RoundDesc& R, const uint32_t cnt[3], const uint32_t pre[3]) {
    return (uint64_t)pre[0] * R.C[1] * R.C[2] +
           (uint64_t)cnt[0] * ((uint64_t)pre[1] * R.C[2] + (uint64_t)cnt[1] * pre[2]);
}

// Slot of local index j inside a block of shape cnt (get_qb_index,
// asz_param.hpp:122-129): order 0 = "120", 1 = "021", 2 = "012".
ASZ_HD uint64_t local_slot(uint32_t order, const uint32_t j[3], const uint32_t cnt[3]) {
    if (order == 0) return ((uint64_t)j[1] * cnt[2] + j[2]) * cnt[0] + j[0];
    if (order == 1) return ((uint64_t)j[0] * cnt[2] + j[2]) * cnt[1] + j[1];
    return ((uint64_t)j[0] * cnt[1] + j[1]) * cnt[2] + j[2];
}

ASZ_HD uint64_t qb_slot(const RoundDesc& R, const uint32_t* tabs, const TargetLoc& L) {
    uint32_t cnt[3], pre[3];
    for (int a = 0; a < 3; a++) {
        cnt[a] = tabs[R.tab_off[a] + L.b[a]];
        pre[a] = tabs[R.tab_off[a] + R.nb[a] + L.b[a]];
    }
    return block_base(R, cnt, pre) + local_slot(R.order, L.j, cnt);
}

// Coordinates of target k of round R.
ASZ_HD void target_coords(const RoundDesc& R, int ax, const uint32_t k[3], uint32_t i[3]) {
    for (int a = 0; a < 3; a++) i[a] = a == ax ? R.h + 2 * R.h * k[a] : R.s[a] * k[a];
}

}  // namespace asz
