// Launchers of the plan-generation kernels (plansample.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace asz {

// One candidate round of the plan search over the whole field
// (generate_config.tpp:554-625): samples at start[a] + k * step[a] along each
// axis, cnt[a] of them, visited in the reference's nesting for `order`.
struct PlanSampleDesc {
    uint32_t D[3];
    uint32_t start[3], step[3], cnt[3];
    uint32_t d, order;   // round axis, traversal order (0 "120", 1 "021", 2 "012")
    uint32_t h, pad;     // half stride along d
    uint64_t dist;       // h * prod_{a > d} D[a]
    uint64_t off;        // first sample in the concatenated QB / error buffers
    uint64_t count;      // cnt[0] * cnt[1] * cnt[2]
    float t, eb, neg2;   // best t, round eb, 0.5f / eb
    float pad2;
};

void launch_plan_sample(const PlanSampleDesc* descs, uint32_t ncand, uint64_t max_count, const float* x, int32_t* qb,
                        float* err, hipStream_t st);
void launch_plan_minmax(const PlanSampleDesc* descs, uint32_t ncand, uint64_t max_count, const int32_t* qb,
                        int32_t* out2, hipStream_t st);
// out[c] = the float FMA chain sum of err[k]^2 over candidate c's samples, in order.
void launch_sq_chain(const PlanSampleDesc* descs, uint32_t ncand, const float* err, float* out, hipStream_t st);

}  // namespace asz
