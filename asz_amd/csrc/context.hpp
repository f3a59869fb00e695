// The library context (asz_ctx) and the device / pinned buffer pools shared
// by the host translation units of the GPU core.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "internal.hpp"
#include "launch.hpp"
#include "pool.hpp"

#define HIPCHECK(x)                                                                               \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess)                                                                     \
            ::asz::fail(ASZ_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));            \
    } while (0)

namespace asz {

// ------------------------------------------------------------------ device buffers

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    template <typename T>
    T* get(size_t n) {
        size_t bytes = std::max<size_t>(n * sizeof(T), 16);
        if (bytes > cap) {
            if (p) HIPCHECK(hipFree(p));
            p = nullptr;
            cap = 0;
            hipError_t e = hipMalloc(&p, bytes);
            if (e != hipSuccess) fail(ASZ_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
            cap = bytes;
        }
        return static_cast<T*>(p);
    }
};

// Page-locked host memory for the large host <-> device transfers.
struct PinnedBuf {
    void* p = nullptr;
    size_t cap = 0;
    ~PinnedBuf() { if (p) (void)hipHostFree(p); }
    template <typename T>
    T* get(size_t n) {
        size_t bytes = std::max<size_t>(n * sizeof(T), 64);
        if (bytes > cap) {
            if (p) HIPCHECK(hipHostFree(p));
            p = nullptr;
            cap = 0;
            bytes = std::max(bytes, (size_t)1 << 20);
            if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess)
                fail(ASZ_ERR_NOMEM, "pinned host allocation of " + std::to_string(bytes) + " bytes failed");
            cap = bytes;
        }
        return static_cast<T*>(p);
    }
};

// A plan with every entry past its `layers` rounds (and the padding) zeroed,
// so that two plans compare bytewise equal exactly when their rounds match
// (a caller may leave the unused entries of its struct uninitialised).
inline asz_plan canonical_plan(const asz_plan& p) {
    asz_plan c;
    std::memset(&c, 0, sizeof(c));
    c.layers = p.layers;
    const int n = std::max(0, std::min<int>(p.layers, ASZ_MAX_LAYERS));
    for (int r = 0; r < n; r++) {
        c.axis[r] = p.axis[r];
        c.order[r] = p.order[r];
    }
    return c;
}
inline bool same_plan(const asz_plan& a, const asz_plan& b) {
    const asz_plan x = canonical_plan(a), y = canonical_plan(b);
    return std::memcmp(&x, &y, sizeof(x)) == 0;
}

// Inputs that shape a compression plan and its device tables.  Compared
// bytewise (the struct is zeroed before it is filled, the plan canonical).
struct PlanKey {
    uint64_t dims[3];
    asz_plan plan;
    uint64_t bpr, h1;
    uint32_t coarse;  // most leading rounds run as block rounds (kRoundBlock)
    uint32_t pad;
    RoundTune tune[ASZ_MAX_LAYERS];  // the autotuner's tile shapes (all zero: the defaults)
};

// The autotuner's choices for one (dims, plan, block pivot ratio, sampling
// count): a tile shape per round.
struct TuneEntry {
    uint64_t dims[3];
    asz_plan plan;  // canonical
    uint64_t bpr, h1;
    RoundTune tune[ASZ_MAX_LAYERS];
};

struct PlanCache {
    bool valid = false;
    uint64_t used = 0;             // context tick of the last call that used it
    PlanKey key;
    DevBuf meta;                   // the device tables (layout in up)
    Plan P;
    std::vector<FoldDesc> folds;   // per round (make_fold_descs)
    uint64_t fpart_total = 0;      // fold scratch Partials over all rounds
    struct {
        uint64_t tabs, ss, segs, folds, bytes;
    } up{};
};

// A captured device sequence (multi-field compression) and its instance.
struct GraphEntry {
    std::vector<uint64_t> sig;
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    uint64_t used = 0;
    void release() {
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
        x = nullptr;
        g = nullptr;
    }
};
// The epoch the kernels of a graph publish with (calls without a graph never use it).
constexpr uint32_t kGraphEpoch = 0xFFFFFFFFu;

}  // namespace asz

struct asz_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    asz_stats stats{};
    asz::DevBuf work, sym, tblk, hist, arrays, chunk_arr, words, keys, key_off, misc, dec_up, start, exit0, exit1,
        count, out_pos, partials, fparts, ghist, xfer, hmeta, scratch,
        chunk_bits, chunk_off, pg_meta, pg_qb, pg_err, fdesc, stage, enc_stage, hmeta_early, tune_work, hdr_bytes, hdr_meta;
    asz::PinnedBuf pin_meta, pin_xfer, pin_huff, pin_stream, pin_fdesc, pin_range, pin_early, pin_hdr, pin_hdr_bytes, pin_dec;
    // exact histograms of arrays wider than the global window, and the t and
    // pivot arrays' histograms, which no round waits for
    hipStream_t side = nullptr;
    // ordering: main stream -> side, side -> main stream, and the encoder's
    // early table upload (side) -> its pack pass (main stream)
    hipEvent_t side_ev[3] = {nullptr, nullptr, nullptr};
    hipEvent_t order_ev = nullptr;  // asz_stream_wait / batch workers: cross-stream ordering
    // host result slots (pinned, written by the GPU): kHostSlotWords per array
    uint32_t* hres_host = nullptr;
    uint32_t* hres_dev = nullptr;
    uint64_t hres_arrays = 0;
    // after the slots: the device-range verdict (min bits, max bits, status)
    // that k_field_desc writes for asz_compress_field
    uint32_t* hrange_host = nullptr;
    uint32_t* hrange_dev = nullptr;
    uint32_t epoch = 0;
    std::vector<asz_ctx*> kids;  // batch workers: one context (and HIP stream) each
    // plans of the last kPlanSlots (dims, plan, pivot ratio, sampling count)
    // keys, least recently used replaced first
    static constexpr int kPlanSlots = 4;
    asz::PlanCache pcs[kPlanSlots];
    uint64_t pc_tick = 0;
    bool stream_on_device = false;
    uint32_t batch_fields = 0;  // ASZ_OPT_BATCH_FIELDS
    bool wide_codes = false;  // ASZ_OPT_WIDE_CODES
    int coarse_max = -1;      // ASZ_OPT_COARSE_ROUNDS (-1: every leading round that qualifies)
    bool graphs = true;       // ASZ_OPT_GRAPHS
    std::vector<asz::TuneEntry> tunes;  // autotuned tile shapes (asz_autotune, asz_tune_load)
    static constexpr size_t kGraphSlots = 8;
    std::vector<asz::GraphEntry> graph_cache;  // captured multi-field device sequences, least recently used first out
    float dev_mn = 0, dev_mx = 0;  // range of the last asz_compress_field
    uint8_t* out_stream = nullptr;  // last compressed stream (pin_stream, or device memory)
    uint64_t out_bytes = 0;
    // multi-field calls: each field's stream (in words or pin_stream) and statistics
    std::vector<std::pair<uint8_t*, uint64_t>> fouts;
    std::vector<asz_stats> fstats;
    hipEvent_t ev[8];
    // per-kernel event pairs (ASZ_OPT_KERNEL_EVENTS)
    bool kev_on = false;
    int kev_round = -1;  // >= 0: events around that round's kernel only (ASZ_OPT_TIMED_ROUND)
    std::vector<hipEvent_t> kev;
    std::vector<std::pair<int, int>> kev_tag;
    int kev_n = 0;
    std::unique_ptr<asz::TaskPool> pool;  // host workers of plan generation / family selection (lazy)
};

namespace asz {

// The side stream and its ordering events, created on first use (so that a
// context normally owns one HIP stream).  Not thread-safe: callers create it
// before their worker threads can.  The events only order the two streams of
// one device (no host waits on them), so they skip the system-scope fence.
inline void ensure_side(asz_ctx* c) {
    if (!c->side) HIPCHECK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    for (hipEvent_t& e : c->side_ev)
        if (!e) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence));
}

inline TaskPool& host_pool(asz_ctx* c) {
    if (!c->pool) c->pool.reset(new TaskPool(TaskPool::default_workers()));
    return *c->pool;
}

// The range of one field from its launch_range result words (codec.cpp).
void range_result(asz_ctx* c, const uint32_t* w, const float* d_field, float* mn, float* mx, const std::string& who);
// The autotuned tile shapes for (dims, plan, block pivot ratio, sampling count), if any (codec.cpp).
const TuneEntry* find_tune(const asz_ctx* c, const uint64_t dims[3], const asz_plan& plan, uint64_t bpr, uint64_t h1);
// One compression timing one round's kernel (codec.cpp; the autotuner's trial).
double compress_timed_round(asz_ctx* c, float* d_work, const uint64_t dims[3], const asz_plan& plan,
                            const asz_params& p, float mn, float mx, int timed_round);
// Resolves, once per device (current), the address of the allocation-free
// range scan's scratch array (noalloc.hip).
void noalloc_setup(int device);

inline float abs_eb_of(float rel_eb, float mn, float mx) {
    float range = daz(daz(mx) - daz(mn));
    return daz(daz(rel_eb) * range);
}

}  // namespace asz

template <typename F>
inline int guarded(asz_ctx* c, F&& f) {
    try {
        f();
        return ASZ_OK;
    } catch (const asz::Error& e) {
        if (c) c->err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        if (c) c->err = "host allocation failed";
        return ASZ_ERR_NOMEM;
    } catch (const std::exception& e) {
        if (c) c->err = e.what();
        return ASZ_ERR_DATA;
    }
}

