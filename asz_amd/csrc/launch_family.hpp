// Launchers of the predictor-family probe kernels (probe.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace asz {

// One probe job: a context box [halo_begin, context_end) of the field copied
// into its own local arrays (predictor_family.tpp:38-59), pivots quantized on
// the lattice inside [context_begin, context_end) (predictor_family.tpp:152-159).
struct ProbeBox {
    uint32_t D[3];        // field dims
    uint32_t LD[3];       // local dims
    uint32_t halo[3];     // halo_begin (global)
    uint32_t cb[3];       // context_begin (global)
    uint32_t ps[3];       // pivot stride
    uint32_t pad;
    uint64_t base;        // first element of the box in the local arrays
    uint64_t n;           // LD[0] * LD[1] * LD[2]
    double inv_half_pe;   // 0.5 / (double) seed_eb
    double two_pe;        // 2 * (double) seed_eb
};

// Key sampling of one block of one job for its round's t selection
// (compress.tpp:96-116 on the local arrays).
struct ProbeKeys {
    uint64_t base;        // box base
    uint32_t LD[3];
    uint32_t lo[3], step[3], cnt[3];
    uint64_t dist;
    uint64_t key_off;     // first key pair of this unit
    float neg2;           // 0.5f / stat_eb
    uint32_t pad;
};

// One block of one job in one round (compress.tpp:51-88 on the local arrays).
struct ProbeUnit {
    uint64_t base;
    uint32_t LD[3];
    uint32_t start[3], stride[3], n[3];  // targets: start + j * stride, j < n (get_layer_qb_shape)
    uint32_t d, order, h, right;         // round axis, QB order, half stride, block end along d
    uint32_t quintic, diag;              // 1D6R interior; target block (records + QBs, out-of-line quantizer)
    uint64_t dist;
    uint64_t qb_off, rec_off;            // QB and record slots of a target block
    float t, pad;
    double inv_half_eb, eb_d, neg_two_eb;
};

void launch_probe_init(const ProbeBox* boxes, uint32_t nbox, uint64_t max_n, const float* field, float* orig,
                       float* recon, hipStream_t st);
void launch_probe_keys(const ProbeKeys* units, uint32_t nunit, const float* recon, int64_t* keys, hipStream_t st);
// rec: (pred, recon, orig) float triplets of the target blocks' targets in loop order
void launch_probe_round(const ProbeUnit* units, uint32_t nunit, const float* orig, float* recon, int32_t* qb,
                        float* rec, hipStream_t st);

}  // namespace asz
