// Host orchestration of the GPU compression core and the C ABI (asz_gpu.h).
//
// compress (param_c::compress, compress.tpp:146-299 without Zstd):
//   pivots -> for each round: t (GPU selection / exact host scoring / injected)
//   -> round kernel (QBs, reconstruction, per-workgroup error / QB range /
//   histogram window) -> fold of the round's partials and histograms (the
//   small coarse rounds in one batch).  The kernel that finishes an array
//   writes its QB range and used histogram bins straight into pinned host
//   memory and raises a flag; host worker threads build the reference-tree
//   codebooks as the flags come up, while this thread is still launching and
//   the GPU runs the later rounds.  Then one upload of headers and codebooks
//   -> persistent chunk encoder -> per-array scan of chunk bit counts ->
//   bit-exact compaction into a stream buffer that already holds every
//   header.  The stream stays in HBM (ASZ_OPT_STREAM_ON_DEVICE) or is copied
//   once to pinned memory.  Plan tables are uploaded once and reused while
//   the plan inputs repeat.
// decompress (param_c::decompress, decompress.tpp:88-196): host parses the
//   headers and builds decode tables; the GPU decodes every array and runs
//   pivots and rounds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <atomic>
#include <immintrin.h>
#include <memory>
#include <mutex>
#include <thread>

#include "internal.hpp"
#include "launch.hpp"
#include "launch_misc.hpp"
#include "scoring.hpp"

#include "context.hpp"

namespace asz {

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}
float ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

enum KevKind { KEV_PIVOT, KEV_TSEL, KEV_ROUND, KEV_HIST, KEV_ENCODE, KEV_DECODE, KEV_COARSE, KEV_RANGE };

// s: the stream the kernel runs on (nullptr: the context's stream)
int kev_begin(asz_ctx* c, int kind, int r, hipStream_t s = nullptr) {
    if (!c->kev_on || c->kev_n * 2 + 2 > (int)c->kev.size()) return -1;
    if (c->kev_round >= 0 && !((kind == KEV_ROUND || kind == KEV_COARSE) && r == c->kev_round)) return -1;
    int i = c->kev_n++;
    c->kev_tag[i] = {kind, r};
    (void)hipEventRecord(c->kev[2 * i], s ? s : c->stream);
    return i;
}
void kev_end(asz_ctx* c, int i, hipStream_t s = nullptr) {
    if (i >= 0) (void)hipEventRecord(c->kev[2 * i + 1], s ? s : c->stream);
}
// After the stream is synchronised: fold the recorded pairs into the stats.
void kev_collect(asz_ctx* c) {
    asz_stats& S = c->stats;
    for (int i = 0; i < c->kev_n; i++) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->kev[2 * i], c->kev[2 * i + 1]);
        int r = c->kev_tag[i].second;
        switch (c->kev_tag[i].first) {
            case KEV_PIVOT: S.pivot_ms += ms; break;
            case KEV_TSEL: S.tsel_round_ms[r] += ms; break;
            case KEV_ROUND: S.round_ms[r] += ms; break;
            case KEV_HIST: S.hist_ms += ms; break;
            case KEV_ENCODE: S.encode_ms += ms; break;
            case KEV_RANGE: S.range_ms += ms; break;
            case KEV_COARSE:
                S.coarse_ms += ms;
                S.round_ms[r] += ms;
                break;
            default: S.decode_ms += ms; break;
        }
    }
    c->kev_n = 0;
}

uint64_t get_le64(const uint8_t* p) {
    uint64_t v = 0;
    for (int k = 0; k < 8; k++) v |= (uint64_t)p[k] << (8 * k);
    return v;
}
uint64_t get_be64(const uint8_t* p) {
    uint64_t v = 0;
    for (int k = 0; k < 8; k++) v = (v << 8) | p[k];
    return v;
}
// the Huffman payload's bit length is stored xor this (huffman.hpp)
constexpr uint64_t kPayloadLenMask = 0x1234abcdull;

void validate_params(const asz_params& p) {
    if (!(p.rel_eb > 0 && std::isfinite(p.rel_eb))) fail(ASZ_ERR_ARG, "relative error bound must be positive and finite");
    if (!(p.pivot_eb_ratio > 0 && p.pivot_eb_ratio <= 1)) fail(ASZ_ERR_ARG, "pivot_eb_ratio must be in (0,1]");
    if (!(p.eb_ratio > 0 && p.eb_ratio <= 1)) fail(ASZ_ERR_ARG, "eb_ratio must be in (0,1]");
    if (!(p.slope_ratio > 0 && p.slope_ratio <= 1)) fail(ASZ_ERR_ARG, "slope_ratio must be in (0,1]");
    if (p.block_pivot_ratio == 0) fail(ASZ_ERR_ARG, "block_pivot_ratio must be positive");
    if (p.t_precision < 16 || p.t_precision % 16 != 0) fail(ASZ_ERR_ARG, "t_precision must be a positive multiple of 16");
}

PivotDesc pivot_desc(const Plan& P) {
    PivotDesc D{};
    for (int a = 0; a < 3; a++) {
        D.D[a] = (uint32_t)P.g.D[a];
        D.ps[a] = (uint32_t)P.g.pivot_stride[a];
        D.pn[a] = P.pn[a];
    }
    D.count = P.pivot_count;
    D.inv_half_pe = 0.5 / (double)P.pivot_eb;
    D.two_pe = (double)P.pivot_eb * 2.0;
    return D;
}

// Largest per-axis block target counts of round R (the QB order move's grid).
void round_cmax(const Plan& P, const RoundDesc& R, uint32_t cmax[3]) {
    for (int a = 0; a < 3; a++) {
        cmax[a] = 0;
        for (uint32_t b = 0; b < R.nb[a]; b++) cmax[a] = std::max(cmax[a], P.tables[R.tab_off[a] + b]);
    }
}

// The grid dimensions the round kernels use (sweep: y = a lattice axis, z =
// segments x fields; row, t selection: y = fields) must fit the launch limits.
void check_grid(const RoundDesc& R, uint32_t nf) {
    if (R.n[0] > 65535 || R.n[1] > 65535 || (uint64_t)R.nseg * nf > 65535 || nf > 65535)
        fail(ASZ_ERR_ARG, "field dimensions exceed the launch grid");
}


// Layout of the Huffman-coded arrays: array 0 = t array (int64 semantics),
// arrays 1..L+1 = pivot array and round arrays.
struct Layout {
    std::vector<HuffArray> arrays;
    std::vector<HuffArrayDev> dev;
    uint64_t total_bytes = 0;
    int max_len = 0;
};

}  // namespace

// ------------------------------------------------------------------ range

// Preset result words of nf range scans (launch_range).
static std::vector<uint32_t> range_init(uint32_t nf) {
    std::vector<uint32_t> init(kRangeWords * (uint64_t)nf, 0u);
    for (uint32_t f = 0; f < nf; f++) {
        init[kRangeWords * f] = 0xFFFFFFFFu;      // min key
        init[kRangeWords * f + 4] = 0xFFFFFFFFu;  // first zero index (uint64 ~0)
        init[kRangeWords * f + 5] = 0xFFFFFFFFu;
    }
    return init;
}

static float key_to_float(uint32_t k) {
    uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

void range_result(asz_ctx* c, const uint32_t* w, const float* d_field, float* mn, float* mx, const std::string& who);

// Range scans of nf fields: one launch (one per field when the sizes
// differ), one copy back, one synchronisation.  names (may be null): the
// fields' indices, named in the error messages.
static void range_scan_fields(asz_ctx* c, uint32_t nf, const float* const* d_fields, const uint64_t* n, float* mn,
                              float* mx, const uint32_t* names = nullptr) {
    // device: per-field result words, then the field pointers of a one-launch scan
    const uint64_t nw = kRangeWords * (uint64_t)nf;
    uint32_t* r = c->misc.get<uint32_t>(nw + 2 * (uint64_t)nf);
    std::vector<uint32_t> init = range_init(nf);
    bool same_n = nf > 1;
    for (uint32_t f = 1; f < nf; f++) same_n = same_n && n[f] == n[0];
    if (same_n) {
        init.resize(nw + 2 * (uint64_t)nf);
        std::memcpy(init.data() + nw, d_fields, nf * sizeof(float*));
    }
    HIPCHECK(hipMemcpyAsync(r, init.data(), init.size() * 4, hipMemcpyHostToDevice, c->stream));
    if (same_n)
        launch_range(d_fields[0], n[0], r, c->stream, reinterpret_cast<const float* const*>(r + nw), nf);
    else
        for (uint32_t f = 0; f < nf; f++) launch_range(d_fields[f], n[f], r + kRangeWords * f, c->stream);
    std::vector<uint32_t> res(nw);
    HIPCHECK(hipMemcpyAsync(res.data(), r, nw * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    for (uint32_t f = 0; f < nf; f++) {
        const std::string who = names ? "field " + std::to_string(names[f]) + ": " : std::string();
        range_result(c, res.data() + kRangeWords * f, d_fields[f], mn + f, mx + f, who);
    }
}

// The range of one field from its kRangeWords result words (launch_range),
// as load_input finds it (s1_compress.cpp:128-137); who prefixes the error
// messages.
void range_result(asz_ctx* c, const uint32_t* w, const float* d_field, float* mn, float* mx, const std::string& who) {
    if (w[2]) fail(ASZ_ERR_DATA, who + "input contains a nonfinite value");
    float vmin = key_to_float(w[0]), vmax = key_to_float(w[1]);
    // ties among zeros / denormals under DAZ: minmax_element keeps the first
    // smallest and the last largest element, with their own bits
    uint64_t zi[2];
    std::memcpy(zi, w + 4, 16);
    if ((vmin == 0.0f || vmax == 0.0f) && zi[0] != ~0ull) {
        if (vmin == 0.0f) HIPCHECK(hipMemcpyAsync(&vmin, d_field + zi[0], 4, hipMemcpyDeviceToHost, c->stream));
        if (vmax == 0.0f) HIPCHECK(hipMemcpyAsync(&vmax, d_field + zi[1], 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
    }
    if (daz(vmin) == daz(vmax)) fail(ASZ_ERR_DATA, who + "max_val==min_val");
    *mn = vmin;
    *mx = vmax;
}

static void range_scan(asz_ctx* c, const float* d_field, uint64_t n, float* mn, float* mx) {
    range_scan_fields(c, 1, &d_field, &n, mn, mx);
}

// ------------------------------------------------------------------ default slope

// Default slopes of nf fields of shape dims: every field's AB-key samples in
// one buffer (one launch per kSlopeMaxFields fields), one copy back and one
// synchronisation, then the fields' host scoring in parallel on the
// context's pool.  gpu_hold, when given, is released once the GPU part is
// done (the batch's workers take turns at it).
static void default_slope_fields(asz_ctx* c, uint32_t nf, const float* const* d_fields, const uint64_t dims[3],
                                 const asz_params& p, const float* mn, const float* mx, SlopeResult* out,
                                 std::unique_lock<std::mutex>* gpu_hold = nullptr) {
    SlopeFieldsDesc B{};
    uint64_t off[4] = {0, 0, 0, 0};  // key offsets of the three axes in one field's slice
    for (int d = 0; d < 3; d++) {
        SlopeSampleDesc S{};
        uint64_t lo[3] = {0, 0, 0}, hi[3] = {dims[0], dims[1], dims[2]}, st[3] = {1, 1, 1};
        lo[d] = 3;
        hi[d] = dims[d] - 3;
        st[d] = 2;
        uint64_t dist = 1;
        for (int a = d + 1; a < 3; a++) dist *= dims[a];
        uint64_t vol = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
        uint64_t den = st[2] * (st[0] * st[1] * p.h1_block_sampling_num);
        uint64_t ss = (uint64_t)std::cbrt((double)((float)vol / (float)den));
        if (ss == 0) ss = 1;
        for (int a = 0; a < 3; a++) {
            S.D[a] = (uint32_t)dims[a];
            S.lo[a] = (uint32_t)lo[a];
            S.hi[a] = (uint32_t)hi[a];
            S.step[a] = (uint32_t)(st[a] * ss);
            S.cnt[a] = (uint32_t)((hi[a] - lo[a] + S.step[a] - 1) / S.step[a]);
        }
        S.dist = dist;
        B.S[d] = S;
        B.off[d] = off[d];
        off[d + 1] = off[d] + 2 * (uint64_t)S.cnt[0] * S.cnt[1] * S.cnt[2];
    }
    const uint64_t per = off[3];
    B.per = per;
    int64_t* dk = c->keys.get<int64_t>(per * nf + 2);
    for (uint32_t f0 = 0; f0 < nf; f0 += kSlopeMaxFields) {
        B.nf = std::min<uint32_t>(kSlopeMaxFields, nf - f0);
        for (uint32_t j = 0; j < B.nf; j++) {
            B.f[j] = d_fields[f0 + j];
            B.scale[j] = 1.0f / (abs_eb_of(p.rel_eb, mn[f0 + j], mx[f0 + j]) * p.stat_eb_ratio);
        }
        launch_slope_sample_fields(B, dk + f0 * per, c->stream);
    }
    int64_t* hk = c->pin_xfer.get<int64_t>(per * nf + 2);
    HIPCHECK(hipMemcpyAsync(hk, dk, per * nf * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (gpu_hold) gpu_hold->unlock();
    host_pool(c).run(nf, [&](size_t f) {
        std::vector<std::vector<int64_t>> keys(3);
        for (int d = 0; d < 3; d++) keys[d].assign(hk + f * per + off[d], hk + f * per + off[d + 1]);
        out[f] = default_slope_from_keys(keys, abs_eb_of(p.rel_eb, mn[f], mx[f]), p.stat_eb_ratio,
                                         p.default_slope_eb_span, p.default_slope_invalid_ratio_threshold,
                                         p.default_slope_fallback, p.t_min, p.t_max, p.t_precision, mn[f], mx[f]);
    });
}

static SlopeResult default_slope(asz_ctx* c, const float* d_field, const uint64_t dims[3], const asz_params& p,
                                 float mn, float mx) {
    SlopeResult r;
    default_slope_fields(c, 1, &d_field, dims, p, &mn, &mx, &r);
    return r;
}

// ------------------------------------------------------------------ compress

namespace {

// Device layout of a plan's tables, uploaded once per plan (offsets in bytes).
using PlanUpload = decltype(PlanCache::up);

PlanUpload plan_upload_layout(const Plan& P) {
    auto al = [](uint64_t v) { return (v + 255) / 256 * 256; };
    PlanUpload u{};
    u.tabs = 0;
    u.ss = al(u.tabs + P.tables.size() * 4);
    u.segs = al(u.ss + P.ss.size() * 4);
    u.folds = al(u.segs + P.segs.size() * sizeof(SweepSeg));
    u.bytes = al(u.folds + P.rounds.size() * sizeof(FoldDesc)) + 256;
    return u;
}

}  // namespace

// Pinned host slots the GPU writes each array's histogram results into
// (coherent, mapped: written by kernels while they run, read by the host),
// followed by the device-range verdict words.
static void ensure_host_slots(asz_ctx* c, uint64_t narr) {
    if (c->hres_arrays >= narr) return;
    if (c->hres_host) HIPCHECK(hipHostFree(c->hres_host));
    c->hres_host = nullptr;
    c->hres_dev = nullptr;
    c->hrange_host = nullptr;
    c->hrange_dev = nullptr;
    c->hres_arrays = 0;
    const uint64_t n = std::max<uint64_t>(narr, 16);
    const uint64_t bytes = n * kHostSlotWords * 4 + 64;
    void* h = nullptr;
    if (hipHostMalloc(&h, bytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
        fail(ASZ_ERR_NOMEM, "pinned host allocation of the histogram slots failed");
    std::memset(h, 0, bytes);
    void* d = nullptr;
    HIPCHECK(hipHostGetDevicePointer(&d, h, 0));
    c->hres_host = static_cast<uint32_t*>(h);
    c->hres_dev = static_cast<uint32_t*>(d);
    c->hrange_host = c->hres_host + n * kHostSlotWords;
    c->hrange_dev = c->hres_dev + n * kHostSlotWords;
    c->hres_arrays = n;
    c->epoch = 0;
}

// The autotuned tile shapes for (dims, plan, block pivot ratio, sampling count), if any.
const TuneEntry* find_tune(const asz_ctx* c, const uint64_t dims[3], const asz_plan& plan, uint64_t bpr, uint64_t h1) {
    const asz_plan cp = canonical_plan(plan);
    for (const TuneEntry& e : c->tunes)
        if (std::memcmp(e.dims, dims, sizeof(e.dims)) == 0 && e.bpr == bpr && e.h1 == h1 &&
            std::memcmp(&e.plan, &cp, sizeof(cp)) == 0)
            return &e;
    return nullptr;
}

// Runs a device sequence as a HIP graph: captured from `st` (with the side
// stream it forks to) the first time a signature is seen, instantiated, kept
// among the context's last kGraphSlots graphs, then launched on `st`.
template <typename F>
static void graph_run(asz_ctx* c, hipStream_t st, const std::vector<uint64_t>& sig, F&& fn) {
    GraphEntry* e = nullptr;
    for (GraphEntry& g : c->graph_cache)
        if (g.sig == sig) e = &g;
    if (!e) {
        HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        try {
            fn();
        } catch (...) {
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCHECK(hipStreamEndCapture(st, &g));
        hipGraphExec_t x = nullptr;
        const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(g);
            fail(ASZ_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        }
        if (c->graph_cache.size() >= asz_ctx::kGraphSlots) {
            auto lru = c->graph_cache.begin();
            for (auto it = c->graph_cache.begin(); it != c->graph_cache.end(); ++it)
                if (it->used < lru->used) lru = it;
            lru->release();
            c->graph_cache.erase(lru);
        }
        c->graph_cache.push_back(GraphEntry{sig, g, x, 0});
        e = &c->graph_cache.back();
        c->stats.graph_captured = 1;
    }
    e->used = ++c->pc_tick;
    HIPCHECK(hipGraphLaunch(e->x, st));
    c->stats.graph_launched = 1;
}

// One field of a compress call: its array (rewritten with the
// reconstruction), value range and default slope.
struct FieldIn {
    float* work;
    float mn, mx;
    float default_slope;
};

// Compresses nf fields of shape dims with one plan.  nf == 1 is the plain
// drop-in (c->out_stream / c->stats); with nf > 1 every per-round kernel
// covers all the fields (the field index is a grid dimension) and each
// field's stream and statistics land in c->fouts / c->fstats.  Multi-field
// calls take the GPU t selection (ASZ_TSEL_GPU or ASZ_TSEL_GRAD) and none of
// the optional outputs.
// stop_round >= 0 (one field): run the pivots and rounds [0, stop_round) only
// (no Huffman coding), leaving the field in the state the round stop_round
// reads.
// dev_range (one field, GPU t selection, no optional outputs): fin's min /
// max are not used; the range scan runs on the context's stream ahead of the
// compression and the error bounds follow from it on the device
// (k_field_desc), so the host never waits between the two.  The range comes
// back with the rest (c->dev_mn / c->dev_mx), and its errors (non-finite
// input, max == min) are reported once the rounds have run.
static void compress_fields(asz_ctx* c, uint32_t nf, const FieldIn* fin_in, const uint64_t dims[3],
                            const asz_plan& plan, const asz_params& p, int tsel_mode, const int64_t* t_inject,
                            int32_t* qb_out, int64_t* t_out, int stop_round = -1, bool dev_range = false) {
    auto wall0 = Clock::now();
    validate_params(p);
    if (tsel_mode < ASZ_TSEL_GPU || tsel_mode > ASZ_TSEL_GRAD) fail(ASZ_ERR_ARG, "unknown t selection mode");
    if (nf == 0) fail(ASZ_ERR_ARG, "no field to compress");
    const bool multi = nf > 1;
    if ((multi || dev_range) && ((tsel_mode != ASZ_TSEL_GPU && tsel_mode != ASZ_TSEL_GRAD) || t_inject || qb_out ||
                                 t_out || stop_round >= 0))
        fail(ASZ_ERR_ARG, "multi-field and device-range calls take the GPU t selection and no optional outputs");
    if (dev_range && multi) fail(ASZ_ERR_ARG, "device-range calls take one field");
    hipStream_t st = c->stream;
    std::vector<FieldIn> fin(fin_in, fin_in + nf);  // min / max filled in later with dev_range
    float* d_work = fin[0].work;
    const float mn = fin[0].mn, mx = fin[0].mx;
    // the plan tables do not depend on the bounds (any positive one will do)
    const float aeb = dev_range ? 1.0f : abs_eb_of(p.rel_eb, mn, mx);

    // plan and its device tables, kept per (dims, plan, block pivot ratio,
    // sampling count) for the last few keys (the error bounds and slopes are
    // set per call)
    PlanKey key;
    std::memset(&key, 0, sizeof(key));
    for (int a = 0; a < 3; a++) key.dims[a] = dims[a];
    key.plan = canonical_plan(plan);
    key.bpr = p.block_pivot_ratio;
    key.h1 = p.h1_block_sampling_num;
    // leading rounds that may run as block rounds (t selection and targets in
    // one launch per round)
    key.coarse = c->coarse_max < 0 ? (uint32_t)ASZ_MAX_LAYERS : (uint32_t)c->coarse_max;
    if (const TuneEntry* te = find_tune(c, dims, plan, p.block_pivot_ratio, p.h1_block_sampling_num))
        std::memcpy(key.tune, te->tune, sizeof(key.tune));
    PlanCache* pslot = nullptr;
    for (PlanCache& e : c->pcs)
        if (e.valid && std::memcmp(&key, &e.key, sizeof(key)) == 0) pslot = &e;
    const bool hit = pslot != nullptr;
    if (!hit) {
        pslot = &c->pcs[0];
        for (PlanCache& e : c->pcs)
            if ((e.valid ? e.used : 0) < (pslot->valid ? pslot->used : 0)) pslot = &e;
    }
    PlanCache& pc = *pslot;
    pc.used = ++c->pc_tick;
    if (!hit) {
        pc.valid = false;
        Geometry g0 = make_geometry(dims, plan, p.block_pivot_ratio);
        pc.P = make_plan(g0, aeb, p.pivot_eb_ratio, p.eb_ratio, p.slope_ratio, fin[0].default_slope,
                         p.h1_block_sampling_num, key.coarse, key.tune);
        const Plan& P0 = pc.P;
        const PlanUpload u = plan_upload_layout(P0);
        uint8_t* hm = c->pin_meta.get<uint8_t>(u.bytes);
        std::memcpy(hm + u.tabs, P0.tables.data(), P0.tables.size() * 4);
        std::memcpy(hm + u.ss, P0.ss.data(), P0.ss.size() * 4);
        std::memcpy(hm + u.segs, P0.segs.data(), P0.segs.size() * sizeof(SweepSeg));
        make_fold_descs(P0.rounds, pc.folds, pc.fpart_total);
        std::memcpy(hm + u.folds, pc.folds.data(), pc.folds.size() * sizeof(FoldDesc));
        uint8_t* dm = pc.meta.get<uint8_t>(u.bytes);
        HIPCHECK(hipMemcpyAsync(dm, hm, u.bytes, hipMemcpyHostToDevice, st));
        pc.up = u;
        pc.key = key;
        pc.valid = true;
    }
    // per-field scalars (field 0 last: the plan keeps field 0's bounds)
    FieldDesc* hfd = nullptr;
    if (multi) hfd = c->pin_fdesc.get<FieldDesc>(nf);
    for (uint32_t f = nf; f-- > 0 && !dev_range;) {
        set_plan_bounds(pc.P, abs_eb_of(p.rel_eb, fin[f].mn, fin[f].mx), p.pivot_eb_ratio, p.eb_ratio,
                        p.slope_ratio, fin[f].default_slope);
        if (!multi) break;
        FieldDesc& F = hfd[f];
        std::memset(&F, 0, sizeof(F));
        F.work = fin[f].work;
        F.inv_half_pe = 0.5 / (double)pc.P.pivot_eb;
        F.two_pe = (double)pc.P.pivot_eb * 2.0;
        F.range = daz(fin[f].mx) - daz(fin[f].mn);
        for (int r = 0; r < pc.P.g.L; r++) {
            const RoundDesc& R = pc.P.rounds[r];
            F.r[r] = FieldRound{R.inv_half_eb, R.eb_d, R.neg_two_eb, R.eb, R.layer_slope};
        }
    }
    const Plan& P = pc.P;
    const Geometry& g = P.g;
    uint8_t* dm = static_cast<uint8_t*>(pc.meta.p);
    const uint32_t* tabs = reinterpret_cast<const uint32_t*>(dm + pc.up.tabs);
    const uint32_t* ss = reinterpret_cast<const uint32_t*>(dm + pc.up.ss);
    const SweepSeg* segs = reinterpret_cast<const SweepSeg*>(dm + pc.up.segs);
    const FoldDesc* dfolds = reinterpret_cast<const FoldDesc*>(dm + pc.up.folds);

    const uint64_t N = g.N(), nbl = g.nblocks(), L = (uint64_t)g.L, nt = nbl * L;
    asz_stats& S = c->stats;
    S = asz_stats{};
    c->kev_n = 0;
    S.layers = L;
    S.t_count = nt;
    S.plan_cache_hit = hit ? 1 : 0;

    // Per-call buffers, one slice per field (multiples of 64 elements apart).
    // symbol buffer: t array (nt), then the pivot and round QB arrays, each
    // starting on a kSymAlign boundary
    auto al64 = [](uint64_t v) { return (v + 63) / 64 * 64; };
    const uint64_t npp = (P.pivot_count + 255) / 256;  // pivot kernel workgroups
    const uint64_t narr = L + 2;                       // t array, pivots, rounds
    FieldSet fs{};
    const uint64_t ntp = sym_align(nt);                // the QB arrays start after the padded t array
    fs.sym = al64(ntp + P.qb_total + 16);
    fs.tblk = al64(nt + 1);
    fs.part = al64(P.total_partials + 1);
    fs.phist = al64(P.total_hist_words + 1);
    fs.fpart = al64(npp + pc.fpart_total + 1);
    fs.ghist = narr * kHistStride;  // global histogram windows, one kHistStride slot per array
    fs.hres = narr * kHostSlotWords;
    fs.aligned16 = 1;
    for (uint32_t f = 0; f < nf; f++)
        if (reinterpret_cast<uintptr_t>(fin[f].work) & 15u) fs.aligned16 = 0;
    int32_t* sym0 = c->sym.get<int32_t>(fs.sym * nf);
    float* tblk = c->tblk.get<float>(fs.tblk * nf);
    Partial* part = c->partials.get<Partial>(fs.part * nf);
    Partial* ppart = c->fparts.get<Partial>(fs.fpart * nf);
    Partial* fpart = ppart + npp;
    // the global windows, then the rounds' histogram copies: one fill clears both
    uint32_t* ghist = c->ghist.get<uint32_t>((fs.ghist + fs.phist) * nf);
    uint32_t* phist = ghist + fs.ghist * nf;
    DevAcc* acc = c->xfer.get<DevAcc>(nf);
    FieldDesc* dfd_multi = nullptr;  // multi-field: the device copy of hfd (uploaded with the device sequence)
    if (multi) {
        dfd_multi = c->fdesc.get<FieldDesc>(nf);
        fs.fd = dfd_multi;
    }
    // before anything is launched: a reallocation moves the pinned slots
    ensure_host_slots(c, narr * nf);
    uint32_t* range_res = nullptr;  // dev_range: min bits, max bits, status (pinned)
    if (dev_range) {
        uint32_t* rng = c->misc.get<uint32_t>(kRangeWords + 4);
        const std::vector<uint32_t> init = range_init(1);
        HIPCHECK(hipMemcpyAsync(rng, init.data(), kRangeWords * 4, hipMemcpyHostToDevice, st));
        const int kg = kev_begin(c, KEV_RANGE, 0);
        launch_range(d_work, N, rng, st);
        kev_end(c, kg);
        RangeSeed seed{};
        seed.work = d_work;
        seed.rel_eb = p.rel_eb;
        seed.pivot_eb_ratio = p.pivot_eb_ratio;
        seed.L = (uint32_t)L;
        // as set_plan_bounds (compress.tpp:194-195)
        for (uint64_t r = 0; r < L; r++) {
            seed.eb_factor[r] = std::pow((double)p.eb_ratio, (double)(L - r - 1));
            seed.layer_slope[r] =
                (float)((double)fin[0].default_slope * std::pow((double)p.slope_ratio, (double)(L - r - 1)));
        }
        FieldDesc* dfd = c->fdesc.get<FieldDesc>(1);
        // the verdict goes straight into the pinned words (no copy, no event):
        // the host reads it only after a later kernel's published histogram
        // (or the call's final synchronisation), so it is there by then
        range_res = c->hrange_host;
        launch_field_desc(rng, d_work, seed, dfd, c->hrange_dev, st);
        fs.fd = dfd;
    }
    // dev_range: the range's verdict (from any thread: it is read after an
    // array's histogram arrived, and k_field_desc ran before that array's round)
    auto check_range = [&] {
        if (!dev_range) return;
        if (range_res[2] == 1) fail(ASZ_ERR_DATA, "input contains a nonfinite value");
        if (range_res[2] == 2) fail(ASZ_ERR_DATA, "max_val==min_val");
    };
    int32_t* sym = sym0;  // field 0
    int32_t* qb = sym + ntp;
    // staging array of the rounds whose QBs are written in "012" order first
    uint64_t stage_max = 0;
    for (uint64_t r = 0; r < L; r++)
        if (round_needs_staging(P.rounds[r])) stage_max = std::max(stage_max, P.rounds[r].count);
    const uint64_t stage_stride = al64(stage_max + 1);
    int32_t* stage = stage_max ? c->stage.get<int32_t>(stage_stride * nf) : nullptr;
    auto sym_of = [&](uint32_t f) { return sym0 + f * fs.sym; };
    auto gwin = [&](uint32_t f, uint64_t k) { return ghist + f * fs.ghist + k * kHistStride; };
    // Multi-field calls with the GPU t selection run their device sequence
    // (field descriptors, pivots, every round, folds) as a HIP graph, captured
    // once per (plan, field count, buffers) and replayed (ASZ_OPT_GRAPHS).
    const bool use_graph = c->graphs && multi && (tsel_mode == ASZ_TSEL_GPU || tsel_mode == ASZ_TSEL_GRAD) &&
                           stop_round < 0 && !c->kev_on;
    // The epoch the kernels publish with: a fresh one per call (never 0, the
    // slots start zeroed; never kGraphEpoch), or kGraphEpoch for a graph, whose
    // kernels keep the value they were captured with (the flags are cleared
    // before every graph launch instead).
    uint32_t epoch = kGraphEpoch;
    if (!use_graph) {
        do {
            epoch = ++c->epoch;
        } while (epoch == 0 || epoch == kGraphEpoch);
    } else {
        for (uint64_t i = 0; i < narr * nf; i++)
            __atomic_store_n(c->hres_host + i * kHostSlotWords, 0u, __ATOMIC_RELAXED);
    }
    std::vector<float> tf_inj;
    std::vector<int32_t> ti_inj;
    if (tsel_mode == ASZ_TSEL_INJECT) {
        if (!t_inject) fail(ASZ_ERR_ARG, "t_inject is required with ASZ_TSEL_INJECT");
        std::vector<float> tf(nt);
        std::vector<int32_t> ti(nt);
        for (uint64_t b = 0; b < nbl; b++)
            for (uint64_t r = 0; r < L; r++) {
                int64_t v = t_inject[b * L + r];
                tf[r * nbl + b] = (float)v / (float)p.t_precision;
                ti[b * L + r] = (int32_t)v;
            }
        tf_inj.swap(tf);
        ti_inj.swap(ti);
    }

    // Rounds [0, batch_end) are folded together after the last of them (one
    // launch for the small coarse rounds, whose codebooks are then built while
    // the six finest rounds run); every later round is folded on its own.
    // The block rounds [0, ncoarse) are always folded together.
    const uint64_t Lrun = stop_round >= 0 ? std::min<uint64_t>(L, (uint64_t)stop_round) : L;
    const uint64_t ncoarse = std::min<uint64_t>(P.coarse, Lrun);
    const uint64_t batch_end = std::max<uint64_t>(L >= 7 ? L - 6 : 0, ncoarse);
    // Arrays in the order the GPU finishes them: pivots, the batch, the later
    // rounds, with the t array (final after the last t selection) before the
    // last round.
    std::vector<uint64_t> order{1};
    for (uint64_t r = 0; r < L; r++) {
        if (r + 1 == L) order.push_back(0);
        if (r + 1 == batch_end)
            for (uint64_t q = 0; q <= r; q++) order.push_back(2 + q);
        else if (r + 1 > batch_end)
            order.push_back(2 + r);
    }

    // ---------------------------------------------------------------- Huffman codebooks
    // Host worker threads build the reference-tree codebook of each array as
    // soon as the GPU has written that array's range and bins into its host
    // slot (flag == epoch), concurrently with this thread's launches and with
    // the later rounds on the GPU.
    std::vector<Layout> lays(nf);
    for (Layout& lay : lays) {
        lay.arrays.resize(narr);
        lay.dev.resize(narr);
        for (uint64_t k = 0; k < narr; k++) {
            lay.arrays[k].count = k == 0 ? nt : (k == 1 ? P.pivot_count : P.rounds[k - 2].count);
            lay.arrays[k].is_int64 = k == 0;
        }
    }
    std::atomic<bool> enqueued{false}, abort_all{false};
    std::mutex side_mu;  // the exact-histogram path uses the side stream and shared scratch
    std::vector<double> host_ms_w;
    auto slot = [&](uint32_t f, uint64_t k) { return c->hres_host + (f * narr + k) * kHostSlotWords; };
    // spin: never sleep (the calling thread has nothing else to do)
    auto wait_slot = [&](uint32_t f, uint64_t k, bool spin) {
        volatile uint32_t* flag = slot(f, k);
        uint64_t spins = 0;
        while (*flag != epoch) {
            if (abort_all.load(std::memory_order_relaxed)) fail(ASZ_ERR_HIP, "compression aborted");
            if ((++spins & 0xFFFF) == 0) {
                const hipError_t q = hipStreamQuery(st);
                if (q != hipSuccess && q != hipErrorNotReady)
                    fail(ASZ_ERR_HIP, std::string("GPU stream failed: ") + hipGetErrorString(q));
                if (q == hipSuccess && enqueued.load() && *flag != epoch) {
                    std::this_thread::sleep_for(std::chrono::milliseconds(1));
                    if (*flag != epoch) fail(ASZ_ERR_HIP, "histogram results never arrived from the GPU");
                }
            }
            // spin briefly, then give the core away (batches run many of these waits at once)
            if (spin || spins < 4096) _mm_pause();
            else if (spins < 16384) std::this_thread::yield();
            else std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
        std::atomic_thread_fence(std::memory_order_acquire);
    };
    auto build = [&](uint32_t f, uint64_t k) {
        check_range();
        HuffArray& A = lays[f].arrays[k];
        if (A.count == 0) return;
        const uint32_t* hs = slot(f, k);
        HistTail tail;
        std::memcpy(&tail, hs + kHostSlotHead, sizeof(tail));
        const int32_t q0 = tail.qmin, q1 = tail.qmax;
        A.offset = q0;
        A.range = (uint64_t)((int64_t)q1 - (int64_t)q0 + 1);
        if (q0 > q1 || A.range > (1ull << 24)) fail(ASZ_ERR_DATA, "QB alphabet wider than 2^24 symbols");
        if (q0 >= -kHistCentre && q1 < kHistCentre) {
            build_codebook(A, hs + kHostSlotBins);
            if (!A.single) pack_book(A);
            return;
        }
        // the array's range leaves the global window: bin it again over its
        // exact range on the side stream (its symbols are final: the flag is up)
        std::lock_guard<std::mutex> lk(side_mu);
        HIPCHECK(hipSetDevice(c->device));
        HuffArrayDev D{};
        D.src_off = k == 0 ? 0 : (k == 1 ? ntp : ntp + P.rounds[k - 2].qb_off);
        D.count = A.count;
        D.offset = A.offset;
        D.range = A.range;
        std::vector<uint32_t> carr((A.count + kHistChunk - 1) / kHistChunk, 0);
        HuffArrayDev* ded = c->arrays.get<HuffArrayDev>(1);
        uint32_t* dcarr = c->chunk_arr.get<uint32_t>(carr.size() + 1);
        uint32_t* dh = c->hist.get<uint32_t>(A.range + 1);
        HIPCHECK(hipMemcpyAsync(ded, &D, sizeof(D), hipMemcpyHostToDevice, c->side));
        HIPCHECK(hipMemcpyAsync(dcarr, carr.data(), carr.size() * 4, hipMemcpyHostToDevice, c->side));
        launch_fill_i32(reinterpret_cast<int32_t*>(dh), A.range, 0, c->side);
        launch_hist(sym_of(f), ded, dcarr, carr.size(), dh, c->side);
        std::vector<uint32_t> hexact(A.range);
        HIPCHECK(hipMemcpyAsync(hexact.data(), dh, A.range * 4, hipMemcpyDeviceToHost, c->side));
        HIPCHECK(hipStreamSynchronize(c->side));
        build_codebook(A, hexact.data());
        if (!A.single) pack_book(A);
    };
    // Codebook workers: the host pool's threads take the arrays in `order`
    // (every field's copy of an array in turn) as their results arrive,
    // concurrently with the launches below and with the later rounds on the GPU.
    // With one field, this thread builds the last round's codebook itself once
    // its launches are done (no hand-over to a sleeping worker at the end).
    const uint64_t ntask_all = order.size() * nf, ntask = ntask_all - (nf == 1 ? 1 : 0);
    host_ms_w.assign(ntask_all, 0.0);
    auto task = [&](size_t i, bool spin) {
        const uint64_t k = order[i / nf];
        const uint32_t f = (uint32_t)(i % nf);
        if (lays[f].arrays[k].count == 0) return;  // an empty round has no slot
        wait_slot(f, k, spin);
        auto h0 = Clock::now();
        build(f, k);
        host_ms_w[i] = ms_since(h0);
    };
    TaskPool& pool = host_pool(c);
    bool workers_started = false;
    auto start_workers = [&]() {
        if (stop_round >= 0 || workers_started) return;
        pool.start(ntask, [&](size_t i) { task(i, false); });
        workers_started = true;
    };
    // stops and joins the workers on every exit path below
    struct Joiner {
        TaskPool& pool;
        std::atomic<bool>& ab;
        bool armed = true;
        ~Joiner() {
            if (!armed) return;
            ab = true;
            try {
                pool.wait();
            } catch (...) {
            }
        }
    } joiner{pool, abort_all};

    double tsel_host_ms = 0;
    // ASZ_TSEL_EXACT: every round's block t (per round) and the block-major t
    // array, alive until the call returns (their uploads are asynchronous)
    std::vector<float> tf_all(tsel_mode == ASZ_TSEL_EXACT ? L * nbl : 0);
    std::vector<int32_t> t_host(tsel_mode == ASZ_TSEL_EXACT ? nt : 0);
    std::vector<int64_t> hkeys;
    std::vector<uint64_t> koff;
    // With the GPU t selection, a round's fold runs in the same launch as the
    // next round's t selection (its time is then part of that round's
    // tsel_round_ms); fold_r0 / fold_nr: the rounds waiting for their fold.
    const bool gpu_tsel = tsel_mode == ASZ_TSEL_GPU || tsel_mode == ASZ_TSEL_GRAD;
    uint64_t fold_r0 = 0, fold_nr = 0;
    auto fold_groups = [&](uint64_t r0, uint64_t nr) {
        uint32_t mg = 0;
        for (uint64_t q = r0; q < r0 + nr; q++) mg = std::max(mg, pc.folds[q].groups);
        return mg;
    };
    // every t is final: the t array's range, histogram and tail, on the side
    // stream beside the last round
    auto t_array_side = [&]() {
        std::lock_guard<std::mutex> lk(side_mu);
        HIPCHECK(hipEventRecord(c->side_ev[0], st));
        HIPCHECK(hipStreamWaitEvent(c->side, c->side_ev[0], 0));
        launch_minmax_i32(sym, nt, (int32_t*)&acc[0].arr[0], c->side, nf, fs.sym, sizeof(DevAcc) / 4);
        launch_hist_small(sym, nt, nullptr, 0, gwin(0, 0), nullptr, c->side, nf, fs.sym, fs.ghist);
        launch_hist_tail(acc, 0, nullptr, 0, gwin(0, 0), c->hres_dev, epoch, c->side, nf, 0, fs.ghist, fs.hres);
        HIPCHECK(hipEventRecord(c->side_ev[1], c->side));
    };
    // after round r: its fold (with the next round's t selection when there
    // is one), or nothing while the coarse batch is still running
    auto after_round = [&](uint64_t r) {
        if (r + 1 < batch_end) return;
        const uint64_t f0 = r + 1 == batch_end ? 0 : r;
        if (gpu_tsel && r + 1 < Lrun) {
            fold_r0 = f0;  // with the next round's t selection
            fold_nr = r + 1 - f0;
        } else {
            int kh = kev_begin(c, KEV_HIST, (int)r);
            launch_fold_rounds(dfolds, (uint32_t)f0, (uint32_t)(r + 1 - f0), fold_groups(f0, r + 1 - f0), part, phist,
                               fpart, acc, ghist, c->hres_dev, epoch, st, fs, nf);
            kev_end(c, kh);
        }
    };
    S.coarse_rounds = ncoarse;
    // The device sequence of the call: field descriptors, accumulators, the
    // injected t, pivots (their histogram on the side stream), every round
    // with its t selection and folds, and the join of the side stream.
    auto device_sequence = [&]() {
        if (multi) HIPCHECK(hipMemcpyAsync(dfd_multi, hfd, nf * sizeof(FieldDesc), hipMemcpyHostToDevice, st));
        launch_init_call(acc, nf, ghist, nf * (fs.ghist + fs.phist), st);
        if (tsel_mode == ASZ_TSEL_INJECT) {
            HIPCHECK(hipMemcpyAsync(tblk, tf_inj.data(), nt * 4, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(sym, ti_inj.data(), nt * 4, hipMemcpyHostToDevice, st));
        }
        PivotDesc PD = pivot_desc(P);
        int kp = kev_begin(c, KEV_PIVOT, 0);
        launch_pivots(true, d_work, qb, PD, ppart, st, fs, nf);
        kev_end(c, kp);
        // the pivot array's histogram on the side stream: no round waits for it
        HIPCHECK(hipEventRecord(c->side_ev[0], st));
        HIPCHECK(hipStreamWaitEvent(c->side, c->side_ev[0], 0));
        launch_hist_small(nullptr, 0, sym + ntp, P.pivot_count, nullptr, gwin(0, 1), c->side, nf, fs.sym, fs.ghist);
        launch_hist_tail(acc, 1, ppart, npp, gwin(0, 1), c->hres_dev, epoch, c->side, nf, fs.fpart, fs.ghist,
                         fs.hres);
        HIPCHECK(hipEventRecord(c->side_ev[1], c->side));
        // codebook workers from here on (a graph starts them after its launch)
        if (!use_graph) start_workers();
        for (uint64_t r = 0; r < Lrun; r++) {
            const RoundDesc& R = P.rounds[r];
            // a block round (r < ncoarse) selects its t inside its own launch
            const bool block_round = r < ncoarse;
            if (!block_round) check_grid(R, nf);
            float* tb = tblk + r * nbl;
            if (gpu_tsel && !block_round) {
                TselDesc T{R, (uint32_t)L, (uint32_t)r, p.stat_eb_ratio, daz(mx) - daz(mn), (float)p.t_precision,
                           p.t_min, p.t_max};
                int kt = kev_begin(c, KEV_TSEL, (int)r);
                if (fold_nr) {
                    launch_fold_tsel(T, d_work, ss + r * nbl, tb, sym, acc, nbl, tsel_mode == ASZ_TSEL_GRAD, dfolds,
                                     (uint32_t)fold_r0, (uint32_t)fold_nr, fold_groups(fold_r0, fold_nr), part, phist,
                                     fpart, ghist, c->hres_dev, epoch, st, fs, nf);
                    fold_nr = 0;
                } else {
                    launch_tsel(T, d_work, ss + r * nbl, tb, sym, acc, nbl, tsel_mode == ASZ_TSEL_GRAD, st, fs, nf);
                }
                kev_end(c, kt);
            } else if (tsel_mode == ASZ_TSEL_EXACT) {
                auto h0 = Clock::now();
                TselDesc T{R, (uint32_t)L, (uint32_t)r, p.stat_eb_ratio, daz(mx) - daz(mn), (float)p.t_precision,
                           p.t_min, p.t_max};
                koff.assign(nbl + 1, 0);
                for (uint64_t b = 0; b < nbl; b++) koff[b + 1] = koff[b] + P.ns[r * nbl + b];
                uint64_t* dko = c->key_off.get<uint64_t>(nbl + 1);
                int64_t* dk = c->keys.get<int64_t>(2 * koff[nbl] + 2);
                HIPCHECK(hipMemcpyAsync(dko, koff.data(), (nbl + 1) * 8, hipMemcpyHostToDevice, st));
                launch_tsel_keys(T, d_work, ss + r * nbl, dko, dk, nbl, st);
                hkeys.resize(2 * koff[nbl] + 2);
                HIPCHECK(hipMemcpyAsync(hkeys.data(), dk, 2 * koff[nbl] * 8, hipMemcpyDeviceToHost, st));
                HIPCHECK(hipStreamSynchronize(st));
                std::vector<float> tf(nbl);
                std::vector<int32_t> ti(nbl);
                const float stat_eb = p.stat_eb_ratio * R.eb;
                for (uint64_t b = 0; b < nbl; b++) {
                    uint64_t n = koff[b + 1] - koff[b];
                    float t = 0.25f;
                    if (P.ss[r * nbl + b] == 0 || n == 0) S.t_fallbacks++;
                    else t = select_t_exact(hkeys.data() + 2 * koff[b], n, R.eb, stat_eb, R.layer_slope, mn, mx);
                    tf[b] = t;
                    float prod = t * (float)p.t_precision;
                    ti[b] = (int32_t)(prod + std::copysign(0.49999997f, prod));
                }
                // this round's t for the round kernel; the block-major t array
                // (compress.tpp:201) is kept on the host and goes up once, after
                // the last round's selection (the only host wait of a round is the
                // key download above)
                std::memcpy(tf_all.data() + r * nbl, tf.data(), nbl * 4);
                HIPCHECK(hipMemcpyAsync(tb, tf_all.data() + r * nbl, nbl * 4, hipMemcpyHostToDevice, st));
                for (uint64_t b = 0; b < nbl; b++) t_host[b * L + r] = ti[b];
                if (r + 1 == Lrun) HIPCHECK(hipMemcpyAsync(sym, t_host.data(), nt * 4, hipMemcpyHostToDevice, st));
                tsel_host_ms += ms_since(h0);
            }
            if (block_round) {
                // one launch: each block's t (selected here with the GPU modes, from
                // tblk otherwise) and its targets
                TselDesc T{R, (uint32_t)L, (uint32_t)r, p.stat_eb_ratio, daz(mx) - daz(mn), (float)p.t_precision,
                           p.t_min, p.t_max};
                const uint32_t tsel = !gpu_tsel ? 0u : (tsel_mode == ASZ_TSEL_GRAD ? 2u : 1u);
                int kc = kev_begin(c, KEV_COARSE, (int)r);
                launch_round_block(T, tsel, p.quintic != 0, d_work, sym, qb, tabs, ss + r * nbl, tb, part, phist,
                                   gwin(0, 2 + r), acc, st, fs, nf);
                kev_end(c, kc);
                // the last round's t values are written by its own launch
                if (r + 1 == L) t_array_side();
                after_round(r);
                continue;
            }
            if (r + 1 == L) t_array_side();
            int kr = kev_begin(c, KEV_ROUND, (int)r);
            if (round_needs_staging(R)) {
                RoundDesc Rs = R;
                Rs.order = 2;
                Rs.qb_off = 0;
                FieldSet fss = fs;
                fss.sym = stage_stride;
                launch_round_compress(Rs, p.quintic != 0, d_work, stage, tabs, tb, segs, part, phist, gwin(0, 2 + r), st,
                                      fss, nf);
                uint32_t cmax[3];
                round_cmax(P, R, cmax);
                launch_qb_order(R, tabs, cmax, stage, qb + R.qb_off, stage_stride, fs.sym, nf, true, st);
            } else {
                launch_round_compress(R, p.quintic != 0, d_work, qb, tabs, tb, segs, part, phist, gwin(0, 2 + r), st,
                                      fs, nf);
            }
            kev_end(c, kr);
            after_round(r);
        }
        HIPCHECK(hipGetLastError());
        // the main stream resumes after the side stream's histograms (its last work)
        HIPCHECK(hipStreamWaitEvent(st, c->side_ev[1], 0));
    };
    // what the captured launches depend on besides the field descriptors:
    // the plan, the mode, every buffer address and stride, the parameters
    // in their arguments (a multi-field call's kernels take each field's
    // array and layer bounds from its descriptor, never from their own
    // arguments, so field 0's pointer and bounds baked into them do not count)
    auto graph_signature = [&]() {
        std::vector<uint64_t> sig;
        const uint64_t* kw = reinterpret_cast<const uint64_t*>(&pc.key);
        sig.assign(kw, kw + sizeof(PlanKey) / 8);
        auto bits = [](float v) { uint32_t u; std::memcpy(&u, &v, 4); return (uint64_t)u; };
        const uint64_t vals[] = {nf, (uint64_t)tsel_mode, (uint64_t)p.quintic, ncoarse, Lrun, batch_end, epoch,
                                 bits(p.stat_eb_ratio), (uint64_t)p.t_precision, bits(p.t_min), bits(p.t_max),
                                 (uint64_t)(uintptr_t)dfd_multi, (uint64_t)(uintptr_t)hfd,
                                 (uint64_t)(uintptr_t)sym0, (uint64_t)(uintptr_t)tblk, (uint64_t)(uintptr_t)part,
                                 (uint64_t)(uintptr_t)ppart, (uint64_t)(uintptr_t)phist, (uint64_t)(uintptr_t)ghist,
                                 (uint64_t)(uintptr_t)acc, (uint64_t)(uintptr_t)stage, stage_stride,
                                 (uint64_t)(uintptr_t)pc.meta.p, (uint64_t)(uintptr_t)c->hres_dev, fs.sym, fs.tblk,
                                 fs.part, fs.phist, fs.fpart, fs.ghist, fs.hres, fs.aligned16};
        sig.insert(sig.end(), std::begin(vals), std::end(vals));
        return sig;
    };
    ensure_side(c);
    HIPCHECK(hipEventRecord(c->ev[0], st));
    if (use_graph) {
        graph_run(c, st, graph_signature(), device_sequence);
        start_workers();
    } else {
        device_sequence();
    }
    if (stop_round >= 0) {
        HIPCHECK(hipStreamSynchronize(st));
        S.wall_ms = ms_since(wall0);
        return;
    }
    HIPCHECK(hipEventRecord(c->ev[1], st));
    enqueued = true;
    joiner.armed = false;
    const auto enq0 = Clock::now();
    // One field: once every codebook but the last round's is built (the pool's
    // tasks), the pack pass codes those arrays while this thread builds the
    // last codebook.  Their chunks come first in the stream's chunk order, so
    // early_chunks chunks are packed here and the rest after the last codebook.
    uint64_t early_chunks = 0;
    auto early_pack = [&]() {
        const Layout& lay = lays[0];
        std::vector<HuffArrayDev> da;
        uint64_t book_a = 0;
        int len_a = 0;
        for (uint64_t k = 0; k + 1 < narr; k++) {
            const HuffArray& A = lay.arrays[k];
            if (!A.count || A.single) continue;
            HuffArrayDev D;
            std::memset(&D, 0, sizeof(D));
            D.src_off = k == 0 ? 0 : (k == 1 ? ntp : ntp + P.rounds[k - 2].qb_off);
            D.count = A.count;
            D.offset = A.offset;
            D.range = A.range;
            D.book_off = book_a;
            D.chunk0 = early_chunks;
            D.nchunk = (A.count + kEncChunk - 1) / kEncChunk;
            D.lds_book = A.range <= kLdsBook ? 1 : 0;
            early_chunks += D.nchunk;
            book_a += A.range;
            len_a = std::max(len_a, A.max_len);
            da.push_back(D);
        }
        if (!early_chunks || len_a > 32) {
            early_chunks = 0;
            return;
        }
        // every later allocation of the chunk counts and the staging slots fits
        // in these (the last array adds at most its count / kEncChunk chunks)
        const uint64_t last = lay.arrays[narr - 1].count;
        const uint64_t max_chunks = early_chunks + (last + kEncChunk - 1) / kEncChunk;
        uint32_t* cbits = c->chunk_bits.get<uint32_t>(max_chunks + 1);
        uint32_t* estage = c->enc_stage.get<uint32_t>(max_chunks * kEncChunk);
        auto al8 = [](uint64_t v) { return (v + 7) / 8 * 8; };
        const uint64_t o_code = 0, o_len = al8(book_a * 4), o_arr = al8(o_len + book_a),
                       bytes = o_arr + da.size() * sizeof(HuffArrayDev);
        uint8_t* hu = c->pin_early.get<uint8_t>(bytes);
        {
            size_t i = 0;
            for (uint64_t k = 0; k + 1 < narr; k++) {
                const HuffArray& A = lay.arrays[k];
                if (!A.count || A.single) continue;
                const uint64_t bo = da[i++].book_off;
                std::memcpy(hu + o_code + bo * 4, A.code32.data(), A.range * 4);
                std::memcpy(hu + o_len + bo, A.len.data(), A.range);
            }
        }
        std::memcpy(hu + o_arr, da.data(), da.size() * sizeof(HuffArrayDev));
        // The upload goes on the side stream, so that it overlaps the last
        // round; the pack pass follows the last round on the main stream (run
        // beside it, the two only share the HBM bandwidth).
        std::lock_guard<std::mutex> lk(side_mu);
        uint8_t* du = c->hmeta_early.get<uint8_t>(bytes);
        HIPCHECK(hipMemcpyAsync(du, hu, bytes, hipMemcpyHostToDevice, c->side));
        HIPCHECK(hipEventRecord(c->side_ev[2], c->side));
        HIPCHECK(hipStreamWaitEvent(st, c->side_ev[2], 0));
        const int kq = kev_begin(c, KEV_ENCODE, 1);
        launch_encode_pack(sym0, reinterpret_cast<const HuffArrayDev*>(du + o_arr), (uint32_t)da.size(), 0,
                           early_chunks, reinterpret_cast<const uint32_t*>(du + o_code), du + o_len, cbits, estage,
                           st);
        kev_end(c, kq);
    };
    try {
        if (nf == 1 && ntask < ntask_all && !c->wide_codes && stop_round < 0) {
            pool.wait();
            early_pack();
        }
        for (size_t i = ntask; i < ntask_all; i++) task(i, true);
    } catch (...) {
        abort_all = true;
        try {
            pool.wait();
        } catch (...) {
        }
        throw;
    }
    pool.wait();
    S.codebook_wait_ms = ms_since(enq0);
    double host_ms = 0;
    for (double v : host_ms_w) {
        host_ms += v;
        S.codebook_max_ms = std::max(S.codebook_max_ms, v);
    }
    host_ms /= std::max(1u, pool.workers());
    // errors and fallbacks from the slots
    c->fstats.assign(nf, asz_stats{});
    for (uint32_t f = 0; f < nf; f++) {
        asz_stats& Sf = c->fstats[f];
        Sf.layers = L;
        Sf.t_count = nt;
        for (uint64_t k = 0; k < narr; k++) {
            if (lays[f].arrays[k].count == 0) continue;
            HistTail tail;
            std::memcpy(&tail, slot(f, k) + kHostSlotHead, sizeof(tail));
            if (k == 0) {
                if (tsel_mode == ASZ_TSEL_GPU || tsel_mode == ASZ_TSEL_GRAD) Sf.t_fallbacks = tail.fallbacks;
            } else {
                Sf.max_err = std::max(Sf.max_err, tail.max_err);
                Sf.sq_err_sum += tail.sq;
            }
        }
    }
    check_range();
    if (dev_range) {
        std::memcpy(&fin[0].mn, range_res, 4);
        std::memcpy(&fin[0].mx, range_res + 1, 4);
        c->dev_mn = fin[0].mn;
        c->dev_mx = fin[0].mx;
    }
    S.max_err = c->fstats[0].max_err;
    S.sq_err_sum = c->fstats[0].sq_err_sum;
    if (tsel_mode == ASZ_TSEL_GPU || tsel_mode == ASZ_TSEL_GRAD) S.t_fallbacks = c->fstats[0].t_fallbacks;
    if (tsel_mode == ASZ_TSEL_INJECT) S.t_fallbacks = 0;
    auto hh0 = Clock::now();

    // codebooks and stream layout (compress.tpp:243-291), per field.  Only the
    // header bytes travel to the device (a compact blob scattered into the
    // zeroed stream buffer); payload bits are ORed in by the encoder.  Every
    // field's segments, header bytes, codebooks and array descriptors go up in
    // one upload.
    // Every field's header segments, header bytes, codebooks and array
    // descriptors in one upload, with offsets into the combined symbol
    // buffer, codebook table and stream buffer (field f's stream starts at
    // word woff[f]), so that one launch of each encoder kernel codes them all.
    std::vector<uint8_t> blob;
    std::vector<ByteSeg> bsegs;
    std::vector<HuffArrayDev> darrs;
    std::vector<uint64_t> woff(nf);
    uint64_t words_total = 0, book_total = 0, chunks = 0;
    int max_len = 0;
    for (uint32_t f = 0; f < nf; f++) {
        Layout& lay = lays[f];
        woff[f] = words_total;
        const uint64_t sbase = woff[f] * 4;  // the field's first stream byte
        uint64_t spos = 0;
        auto emit = [&](const uint8_t* q, uint64_t n) {
            if (!bsegs.empty() && bsegs.back().dst + bsegs.back().len == sbase + spos) bsegs.back().len += n;
            else bsegs.push_back(ByteSeg{sbase + spos, blob.size(), n});
            blob.insert(blob.end(), q, q + n);
            spos += n;
        };
        {
            uint8_t pre[9 + 2 * ASZ_MAX_LAYERS];
            std::memcpy(pre, &fin[f].mn, 4);
            std::memcpy(pre + 4, &fin[f].mx, 4);
            pre[8] = (uint8_t)(L | 0x40 | (p.quintic ? 0x80 : 0));
            for (uint64_t r = 0; r < L; r++) {
                pre[9 + 2 * r] = (uint8_t)('0' + g.axis[r]);
                pre[10 + 2 * r] = (uint8_t)('0' + g.order[r]);
            }
            emit(pre, 9 + 2 * L);
        }
        for (uint64_t k = 0; k < narr; k++) {
            HuffArray& A = lay.arrays[k];
            HuffArrayDev& D = lay.dev[k];
            std::memset(&D, 0, sizeof(D));
            uint8_t cnt8[8];
            for (int b = 0; b < 8; b++) cnt8[b] = (uint8_t)(A.count >> (8 * b));
            emit(cnt8, 8);
            if (!A.count) continue;
            emit(A.header.data(), A.header.size());
            if (A.single) continue;
            D.src_off = f * fs.sym + (k == 0 ? 0 : (k == 1 ? ntp : ntp + P.rounds[k - 2].qb_off));
            D.count = A.count;
            D.offset = A.offset;
            D.range = A.range;
            D.book_off = book_total;
            D.chunk0 = chunks;
            D.nchunk = (A.count + kEncChunk - 1) / kEncChunk;
            chunks += D.nchunk;
            D.payload_bit = (sbase + spos) * 8;
            spos += (A.payload_bits + 7) >> 3;
            D.lds_book = A.range <= kLdsBook ? 1 : 0;
            lay.max_len = std::max(lay.max_len, A.max_len);
            book_total += A.range;
            darrs.push_back(D);
        }
        lay.total_bytes = spos;
        // 128 spare words: the decoder's padding, so that a stream left here
        // is decompressed in place
        words_total += ((lay.total_bytes + 3) / 4 + 128 + 63) / 64 * 64;
        max_len = std::max(max_len, lay.max_len);
        c->fstats[f].stream_bytes = lay.total_bytes;
        for (uint64_t k = 0; k <= L; k++) c->fstats[f].qb_counts[k] = lay.arrays[k + 1].count;
    }
    S.stream_bytes = lays[0].total_bytes;
    for (uint64_t k = 0; k <= L; k++) S.qb_counts[k] = lays[0].arrays[k + 1].count;
    auto al8 = [](uint64_t v) { return (v + 7) / 8 * 8; };
    const int enc_len = c->wide_codes ? std::max(max_len, 33) : max_len;  // > 32: the wide-code path
    // With the early pack (every array but the last already packed), only the
    // last array's codes go up, with a descriptor of its own for its pack
    // pass; otherwise every array's (and the 64-bit book for the wide-code
    // path).  One upload carries them with the header segments and bytes and
    // the array descriptors.
    const bool late_only = early_chunks && enc_len <= 32, need_book = enc_len > 32;
    const HuffArray& AL = lays[0].arrays[narr - 1];
    const bool late = late_only && AL.count && !AL.single;
    const uint64_t late_chunks = late ? (AL.count + kEncChunk - 1) / kEncChunk : 0;
    if (late_only && chunks != early_chunks + late_chunks) fail(ASZ_ERR_DATA, "encoder chunk plan mismatch");
    const uint64_t tab_total = late_only ? (late ? AL.range : 0) : book_total;
    // codebooks as parallel tables (64-bit entries, 32-bit codes, byte lengths)
    const uint64_t o_seg = 0, o_blob = al8(bsegs.size() * sizeof(ByteSeg)), o_book = al8(o_blob + blob.size()),
                   o_code = al8(o_book + (need_book ? tab_total * 8 : 0)), o_len = al8(o_code + tab_total * 4),
                   o_arr = al8(o_len + tab_total), o_late = o_arr + darrs.size() * sizeof(HuffArrayDev),
                   up_bytes = o_late + sizeof(HuffArrayDev) + 8;
    uint8_t* hu = c->pin_huff.get<uint8_t>(up_bytes);
    if (late) {
        std::memcpy(hu + o_code, AL.code32.data(), AL.range * 4);
        std::memcpy(hu + o_len, AL.len.data(), AL.range);
        HuffArrayDev D = darrs.back();  // the last array: its chunks follow the early ones
        D.book_off = 0;
        std::memcpy(hu + o_late, &D, sizeof(D));
    }
    std::memcpy(hu + o_seg, bsegs.data(), bsegs.size() * sizeof(ByteSeg));
    std::memcpy(hu + o_blob, blob.data(), blob.size());
    if (!late_only) {
        size_t i = 0;
        for (uint32_t f = 0; f < nf; f++)
            for (uint64_t k = 0; k < narr; k++) {
                const HuffArray& A = lays[f].arrays[k];
                if (!A.count || A.single) continue;
                const uint64_t bo = darrs[i++].book_off;
                if (need_book) std::memcpy(hu + o_book + bo * 8, A.book.data(), A.range * 8);
                std::memcpy(hu + o_code + bo * 4, A.code32.data(), A.range * 4);
                std::memcpy(hu + o_len + bo, A.len.data(), A.range);
            }
    }
    std::memcpy(hu + o_arr, darrs.data(), darrs.size() * sizeof(HuffArrayDev));
    uint8_t* du = c->hmeta.get<uint8_t>(up_bytes);
    HIPCHECK(hipMemcpyAsync(du, hu, up_bytes, hipMemcpyHostToDevice, st));
    S.layout_ms = ms_since(hh0);
    S.huffman_host_ms = host_ms + S.layout_ms;

    // the stream buffer: the encoder stores whole payload words and ORs the
    // words it shares (which its scan zeroes first); the header bytes are
    // scattered in afterwards (they never share a byte with payload bits)
    uint32_t* dwords = c->words.get<uint32_t>(words_total);
    uint32_t* cbits = c->chunk_bits.get<uint32_t>(chunks + 1);
    uint64_t* coffs = c->chunk_off.get<uint64_t>(chunks + 1);
    HIPCHECK(hipEventRecord(c->ev[4], st));
    int ke = kev_begin(c, KEV_ENCODE, 0);
    const HuffArrayDev* darr = reinterpret_cast<const HuffArrayDev*>(du + o_arr);
    uint32_t* estage = enc_len <= 32 && chunks ? c->enc_stage.get<uint32_t>(chunks * kEncChunk) : nullptr;
    if (late_only) {
        // the last array's pack pass, then the scan and place passes over every chunk
        if (late) {
            const int kq = kev_begin(c, KEV_ENCODE, 2);
            launch_encode_pack(sym0, reinterpret_cast<const HuffArrayDev*>(du + o_late), 1, early_chunks, chunks,
                               reinterpret_cast<const uint32_t*>(du + o_code), du + o_len, cbits, estage, st);
            kev_end(c, kq);
        }
        launch_encode_place(darr, (uint32_t)darrs.size(), chunks, cbits, coffs, estage, dwords, st);
    } else {
        launch_encode(sym0, darr, (uint32_t)darrs.size(), chunks, reinterpret_cast<const uint64_t*>(du + o_book),
                      reinterpret_cast<const uint32_t*>(du + o_code), du + o_len, cbits, coffs, dwords, enc_len,
                      estage, st);
    }
    launch_scatter_bytes(reinterpret_cast<const ByteSeg*>(du + o_seg), (uint32_t)bsegs.size(), du + o_blob,
                         reinterpret_cast<uint8_t*>(dwords), st);
    kev_end(c, ke);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(c->ev[5], st));
    c->fouts.assign(nf, {nullptr, 0});
    if (c->stream_on_device) {
        for (uint32_t f = 0; f < nf; f++)
            c->fouts[f] = {reinterpret_cast<uint8_t*>(dwords + woff[f]), lays[f].total_bytes};
    } else {
        uint8_t* hs = c->pin_stream.get<uint8_t>(words_total * 4 + 1);
        for (uint32_t f = 0; f < nf; f++) {
            uint8_t* dst = hs + woff[f] * 4;
            HIPCHECK(hipMemcpyAsync(dst, dwords + woff[f], lays[f].total_bytes, hipMemcpyDeviceToHost, st));
            c->fouts[f] = {dst, lays[f].total_bytes};
        }
    }
    c->out_stream = c->fouts[0].first;
    c->out_bytes = c->fouts[0].second;
    if (qb_out) {  // the pivot and round arrays, contiguous on the host
        HIPCHECK(hipMemcpyAsync(qb_out, qb, P.pivot_count * 4, hipMemcpyDeviceToHost, st));
        uint64_t o = P.pivot_count;
        for (uint64_t r = 0; r < L; r++) {
            const RoundDesc& R = P.rounds[r];
            if (R.count) HIPCHECK(hipMemcpyAsync(qb_out + o, qb + R.qb_off, R.count * 4, hipMemcpyDeviceToHost, st));
            o += R.count;
        }
    }
    std::vector<int32_t> t32;
    if (t_out) {
        t32.resize(nt);
        HIPCHECK(hipMemcpyAsync(t32.data(), sym, nt * 4, hipMemcpyDeviceToHost, st));
    }
    HIPCHECK(hipStreamSynchronize(st));
    if (t_out)
        for (uint64_t i = 0; i < nt; i++) t_out[i] = t32[i];
    kev_collect(c);
    S.rounds_ms = ev_ms(c->ev[0], c->ev[1]);
    S.huffman_ms = ev_ms(c->ev[4], c->ev[5]);
    S.tsel_ms = tsel_host_ms;
    S.total_gpu_ms = S.rounds_ms + S.huffman_ms;
    S.wall_ms = ms_since(wall0);
}

// One compression of a device field for the autotuner: GPU t selection, the
// stream left in device memory, an event pair around round `timed_round`'s
// kernel only; returns that kernel's device time (ms).
double compress_timed_round(asz_ctx* c, float* d_work, const uint64_t dims[3], const asz_plan& plan,
                            const asz_params& p, float mn, float mx, int timed_round) {
    const bool on_dev = c->stream_on_device, kev = c->kev_on;
    const int kr = c->kev_round;
    c->stream_on_device = true;
    c->kev_on = true;
    c->kev_round = timed_round;
    if (c->kev.empty()) {
        c->kev.resize(2 * 256);
        c->kev_tag.resize(256);
        for (auto& e : c->kev) HIPCHECK(hipEventCreate(&e));
    }
    struct Restore {
        asz_ctx* c;
        bool on_dev, kev;
        int kr;
        ~Restore() {
            c->stream_on_device = on_dev;
            c->kev_on = kev;
            c->kev_round = kr;
        }
    } restore{c, on_dev, kev, kr};
    const FieldIn f{d_work, mn, mx, p.default_slope};
    compress_fields(c, 1, &f, dims, plan, p, ASZ_TSEL_GPU, nullptr, nullptr, nullptr);
    return c->stats.round_ms[timed_round];
}

// One field (the plain drop-in).
static void compress_impl(asz_ctx* c, float* d_work, const uint64_t dims[3], const asz_plan& plan,
                          const asz_params& p, float mn, float mx, int tsel_mode, const int64_t* t_inject,
                          int32_t* qb_out, int64_t* t_out, int stop_round = -1) {
    const FieldIn f{d_work, mn, mx, p.default_slope};
    compress_fields(c, 1, &f, dims, plan, p, tsel_mode, t_inject, qb_out, t_out, stop_round);
}

// ------------------------------------------------------------------ decompress

struct Parsed {
    float mn, mx;
    asz_plan plan{};
    bool quintic;
    uint64_t nt;
    std::vector<DecodeInfo> arrays;   // t array then L+1 QB arrays
    std::vector<uint64_t> counts;     // symbols per array
};

// beside: host work that needs only the preamble (min, max, plan), run on
// this thread while the trees are parsed on the others; its errors are
// reported after the parse's.
static Parsed parse_stream(const uint8_t* s, uint64_t n, const asz_params& p,
                           const std::function<void(const Parsed&)>& beside) {
    Parsed P;
    if (n < 9) fail(ASZ_ERR_DATA, "compressed stream is truncated");
    std::memcpy(&P.mn, s, 4);
    std::memcpy(&P.mx, s + 4, 4);
    uint8_t f = s[8];
    uint64_t L;
    if (f & 0x40) {
        P.quintic = (f & 0x80) != 0;
        L = f & 0x3F;
    } else {
        P.quintic = p.quintic != 0;
        L = f;
    }
    uint64_t off = 9;
    if (L == 0 || L > ASZ_MAX_LAYERS || off + 2 * L + 8 > n) fail(ASZ_ERR_DATA, "compressed stream is truncated");
    P.plan.layers = (int32_t)L;
    for (uint64_t r = 0; r < L; r++) {
        uint8_t a = s[off], o = s[off + 1];
        if (a < '0' || a > '2' || o < '0' || o > '2') fail(ASZ_ERR_DATA, "compressed stream contains an invalid predictor");
        P.plan.axis[r] = (uint8_t)(a - '0');
        P.plan.order[r] = (uint8_t)(o - '0');
        off += 2;
    }
    // Every array's header offset follows from the fixed-size fields (a tree
    // of nleaf leaves takes 2 nleaf - 1 + nleaf * width bits, whole bytes;
    // k_stream_regions walks the same way), so the trees of a stream with
    // many leaves are parsed and tabled on several threads.  A stream whose
    // fields do not add up takes the sequential parse, which reports the
    // first error in stream order.
    std::vector<uint64_t> hoff, cnts;
    uint64_t leaves = 0;
    bool walk_ok = true;
    {
        uint64_t o = off;
        for (uint64_t k = 0; k < L + 2 && walk_ok; k++) {
            if (o + 8 > n) { walk_ok = false; break; }
            const uint64_t cnt = get_le64(s + o);
            o += 8;
            cnts.push_back(cnt);
            hoff.push_back(o);
            if (cnt == 0 && k > 0) continue;
            if (o + 25 > n) { walk_ok = false; break; }
            const uint64_t width = s[o] & 0x3F, nleaf = get_be64(s + o + 9), maxval = get_be64(s + o + 17);
            if (nleaf == 0 || nleaf > n * 8) { walk_ok = false; break; }
            if (maxval == 1 || nleaf == 1) { o += 33; continue; }
            const uint64_t tree = (2 * nleaf - 1 + nleaf * width + 7) / 8;
            if (width > 31 || tree > n || o + 33 + tree > n) { walk_ok = false; break; }
            const uint64_t bits = get_be64(s + o + 25 + tree) ^ kPayloadLenMask;
            if (bits > n * 8 || (bits + 7) / 8 > n - (o + 33 + tree)) { walk_ok = false; break; }
            o += 33 + tree + (bits + 7) / 8;
            leaves += nleaf;
        }
    }
    const uint64_t na = L + 2;
    if (walk_ok && leaves >= 4096) {
        P.counts = cnts;
        P.nt = cnts[0];
        P.arrays.assign(na, DecodeInfo{});
        std::vector<std::exception_ptr> errs(na);
        std::atomic<uint64_t> next{0};
        auto work = [&] {
            for (uint64_t k; (k = next.fetch_add(1)) < na;) {
                if (cnts[k] == 0 && k > 0) continue;
                try {
                    P.arrays[k] = parse_huffman(s, n, hoff[k], cnts[k]);
                } catch (...) {
                    errs[k] = std::current_exception();
                }
            }
        };
        const unsigned nth = std::min<unsigned>(4, std::max(1u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (unsigned t = 1; t < nth; t++) th.emplace_back(work);
        std::exception_ptr berr;
        try {
            beside(P);
        } catch (...) {
            berr = std::current_exception();
        }
        work();
        for (auto& t : th) t.join();
        for (uint64_t k = 0; k < na; k++) {
            if (errs[k]) std::rethrow_exception(errs[k]);
            const bool empty = cnts[k] == 0 && k > 0;
            const uint64_t end = empty ? hoff[k] : P.arrays[k].end_off;
            if (k + 1 < na && end + 8 != hoff[k + 1]) fail(ASZ_ERR_DATA, "malformed Huffman tree");
        }
        if (berr) std::rethrow_exception(berr);
        return P;
    }
    for (uint64_t k = 0; k < L + 2; k++) {
        if (off + 8 > n) fail(ASZ_ERR_DATA, "compressed stream is truncated");
        uint64_t cnt = get_le64(s + off);
        off += 8;
        P.counts.push_back(cnt);
        if (k == 0) P.nt = cnt;
        if (cnt == 0 && k > 0) { P.arrays.push_back(DecodeInfo{}); continue; }
        DecodeInfo di = parse_huffman(s, n, off, cnt);
        off = di.end_off;
        P.arrays.push_back(std::move(di));
    }
    beside(P);
    return P;
}

// A stream in device memory: its header bytes (the ranges k_stream_regions
// lists) copied to the host, at their stream offsets in image, the only bytes
// the host parser reads.  A stream the walk finds malformed comes back whole,
// for the parser to report.
static const uint8_t* stream_headers(asz_ctx* c, const uint8_t* s, uint64_t n, hipStream_t st,
                                     std::unique_ptr<uint8_t[]>& image) {
    const uint64_t mwords = 3 + 2 * kStreamRegions;
    uint64_t cap = 1ull << 20;
    for (int attempt = 0; attempt < 2; attempt++) {
        uint8_t* dout = c->hdr_bytes.get<uint8_t>(cap);
        uint64_t* dmeta = c->hdr_meta.get<uint64_t>(mwords);
        uint64_t* hmeta = c->pin_hdr.get<uint64_t>(mwords);
        launch_stream_regions(s, n, dout, cap, dmeta, st);
        HIPCHECK(hipMemcpyAsync(hmeta, dmeta, mwords * 8, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        const uint64_t status = hmeta[0], need = hmeta[1], nreg = hmeta[2];
        if (status == 2) {
            cap = need;
            continue;
        }
        image.reset(new uint8_t[n]);
        if (status != 0 || nreg > kStreamRegions) {
            HIPCHECK(hipMemcpy(image.get(), s, n, hipMemcpyDeviceToHost));
            return image.get();
        }
        uint8_t* hb = c->pin_hdr_bytes.get<uint8_t>(need + 1);
        HIPCHECK(hipMemcpyAsync(hb, dout, need, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        uint64_t o = 0;
        for (uint64_t r = 0; r < nreg; r++) {
            std::memcpy(image.get() + hmeta[3 + 2 * r], hb + o, hmeta[4 + 2 * r]);
            o += hmeta[4 + 2 * r];
        }
        return image.get();
    }
    fail(ASZ_ERR_DATA, "compressed stream headers could not be gathered");
}

// dev_stream: s is in device memory (decompressed where it lies when it is
// the context's own stream buffer, copied into it otherwise).
static void decompress_impl(asz_ctx* c, const uint8_t* s_in, uint64_t n, const uint64_t dims[3], const asz_params& p,
                            float* d_out, int32_t* qb_out, int64_t* t_out, bool dev_stream = false) {
    const uint8_t* s = s_in;
    auto wall0 = Clock::now();
    validate_params(p);
    hipStream_t st = c->stream;
    asz_stats& S = c->stats;
    S = asz_stats{};
    // the stream goes up while the host parses its headers and builds the
    // decode tables
    // zero words after the stream: a write-pass thread parses up to 31
    // codewords of at most 64 bits past its own symbols
    const uint64_t nwords = (n + 3) / 4 + 128;
    // a device stream in the context's own stream buffer is decoded in place
    // (the buffer grown first, keeping the stream, when it lacks the padding)
    bool in_place = false;
    if (dev_stream) {
        const uint8_t* b = static_cast<const uint8_t*>(c->words.p);
        if (b && s < b + c->words.cap && b < s + n) {
            if (s != b) fail(ASZ_ERR_ARG, "a stream in the context's stream buffer must start it");
            if (c->words.cap < nwords * 4) {
                asz::DevBuf grown;
                uint32_t* nw = grown.get<uint32_t>(nwords);
                HIPCHECK(hipMemcpyAsync(nw, s, n, hipMemcpyDeviceToDevice, st));
                HIPCHECK(hipStreamSynchronize(st));
                std::swap(c->words.p, grown.p);
                std::swap(c->words.cap, grown.cap);
                s = static_cast<const uint8_t*>(c->words.p);
            }
            in_place = true;
        }
    }
    uint32_t* dwords = c->words.get<uint32_t>(nwords);
    // zero words past the stream (in place: past its last word, whose bytes
    // after the stream the encoder left zero; otherwise from its last partial
    // word on, which the copy below then fills)
    const uint64_t z0 = in_place ? (n + 3) / 4 : n / 4;
    launch_fill_i32(reinterpret_cast<int32_t*>(dwords + z0), nwords - z0, 0, st);
    std::unique_ptr<uint8_t[]> image;
    const uint8_t* hs = s;
    if (dev_stream) {
        // the headers first (the walk reads the caller's buffer), then the
        // copy, which runs while the host parses them
        hs = stream_headers(c, s, n, st, image);
        if (!in_place) HIPCHECK(hipMemcpyAsync(dwords, s, n, hipMemcpyDeviceToDevice, st));
    } else {
        HIPCHECK(hipMemcpyAsync(dwords, s, n, hipMemcpyHostToDevice, st));
    }
    Geometry g;
    Plan P;
    const Parsed Q = parse_stream(hs, n, p, [&](const Parsed& head) {
        g = make_geometry(dims, head.plan, p.block_pivot_ratio);
        const float aeb = abs_eb_of(p.rel_eb, head.mn, head.mx);
        P = make_plan(g, aeb, p.pivot_eb_ratio, p.eb_ratio, p.slope_ratio, 1.0f, p.h1_block_sampling_num, 0,
                      nullptr, false);
    });
    const uint64_t N = g.N(), nbl = g.nblocks(), L = (uint64_t)g.L, nt = nbl * L;
    if (Q.nt < nt) fail(ASZ_ERR_DATA, "compressed stream contains too few t values");
    if (Q.counts[1] != P.pivot_count) fail(ASZ_ERR_DATA, "compressed stream has a wrong pivot count");
    for (uint64_t r = 0; r < L; r++)
        if (Q.counts[r + 2] < P.rounds[r].count) fail(ASZ_ERR_DATA, "compressed stream has too few QBs");
    S.layers = L;
    S.t_count = Q.nt;
    S.stream_bytes = n;

    // symbols: t array (Q.nt) then the L+1 QB arrays as stored, each array
    // starting on a kSymAlign boundary (the decoder stores whole cache lines)
    uint64_t total_sym = 0;
    std::vector<uint64_t> out_off(L + 2);
    for (uint64_t k = 0; k < L + 2; k++) {
        out_off[k] = total_sym;
        total_sym = sym_align(total_sym + Q.counts[k]);
    }
    int32_t* sym = c->sym.get<int32_t>(total_sym + 16);
    // decode tables and descriptors
    std::vector<uint32_t> tables;
    std::vector<DecArrayDev> da;
    std::vector<uint32_t> wg_arr;  // decoder workgroup -> array (256 subsequences each)
    for (uint64_t k = 0; k < L + 2; k++) {
        const DecodeInfo& di = Q.arrays[k];
        if (Q.counts[k] == 0) continue;
        if (di.single) {
            if (di.single_count != Q.counts[k])
                fail(ASZ_ERR_DATA, "Huffman payload length does not match its value count");
            launch_fill_i32(sym + out_off[k], Q.counts[k], (int32_t)di.offset, st);
            continue;
        }
        if (di.max_len > 64) fail(ASZ_ERR_DATA, "Huffman code longer than 64 bits");
        DecArrayDev D{};
        D.payload_bit = di.payload_off * 8;
        D.bits = di.payload_bits;
        D.count = Q.counts[k];
        D.out_off = out_off[k];
        D.table_off = tables.size();
        D.offset = di.offset;
        D.root_bits = (uint32_t)std::min(di.max_len, kRootBits);
        D.max_len = (uint32_t)di.max_len;
        D.lead = sync_lead(di.payload_bits, Q.counts[k]);
        D.pair_off = kNoPairTable;
        D.nsub = (di.payload_bits + kSubBits - 1) / kSubBits;
        D.wg0 = wg_arr.size();
        D.nwg = (D.nsub + 255) / 256;
        D.sub0 = D.wg0 * 256;
        tables.insert(tables.end(), di.table.begin(), di.table.end());
        if (di.max_len <= 32) {
            // the sync pass's root table: two codes per lookup where both fit
            const uint32_t rb = D.root_bits, mask = (1u << rb) - 1;
            D.pair_off = (uint32_t)tables.size();
            for (uint32_t q = 0; q <= mask; q++) {
                const uint32_t e1 = di.table[q];
                uint32_t e = e1;
                if (!(e1 & 0x80000000u)) {
                    const uint32_t l1 = e1 >> 24, e2 = di.table[(q >> l1) & mask];
                    if (!(e2 & 0x80000000u) && (e2 >> 24) <= rb - l1) e = kPairFlag | l1 << 8 | (l1 + (e2 >> 24));
                }
                tables.push_back(e);
            }
        }
        for (uint64_t i = 0; i < D.nwg; i++) wg_arr.push_back((uint32_t)da.size());
        da.push_back(D);
    }
    const uint64_t nwg = wg_arr.size(), nsub = nwg * 256;
    // decode tables, array descriptors, workgroup map, the plan's tables and
    // sweep segments: one upload from pinned memory (a pageable copy waits
    // on the host for its staging)
    auto al = [](uint64_t x) { return (x + 255) & ~uint64_t(255); };
    const uint64_t o_da = al(tables.size() * 4), o_wg = al(o_da + da.size() * sizeof(DecArrayDev)),
                   o_tabs = al(o_wg + nwg * 4), o_segs = al(o_tabs + P.tables.size() * 4),
                   up_bytes = al(o_segs + P.segs.size() * sizeof(SweepSeg));
    uint8_t* hup = c->pin_dec.get<uint8_t>(up_bytes);
    std::memcpy(hup, tables.data(), tables.size() * 4);
    std::memcpy(hup + o_da, da.data(), da.size() * sizeof(DecArrayDev));
    std::memcpy(hup + o_wg, wg_arr.data(), nwg * 4);
    std::memcpy(hup + o_tabs, P.tables.data(), P.tables.size() * 4);
    if (!P.segs.empty()) std::memcpy(hup + o_segs, P.segs.data(), P.segs.size() * sizeof(SweepSeg));
    uint8_t* dup = c->dec_up.get<uint8_t>(up_bytes);
    HIPCHECK(hipMemcpyAsync(dup, hup, up_bytes, hipMemcpyHostToDevice, st));
    const uint32_t* dtab = reinterpret_cast<const uint32_t*>(dup);
    const DecArrayDev* dda = reinterpret_cast<const DecArrayDev*>(dup + o_da);
    const uint32_t* dwg = reinterpret_cast<const uint32_t*>(dup + o_wg);
    const uint32_t* tabs = reinterpret_cast<const uint32_t*>(dup + o_tabs);
    const SweepSeg* dsegs = reinterpret_cast<const SweepSeg*>(dup + o_segs);
    uint64_t* dstart = c->start.get<uint64_t>(nsub + 1);
    uint64_t* dex0 = c->exit0.get<uint64_t>(nsub + 1);
    uint64_t* dex1 = c->exit1.get<uint64_t>(nsub + 1);
    uint32_t* dcnt = c->count.get<uint32_t>(nsub + 1);
    uint64_t* dwgs = c->out_pos.get<uint64_t>(2 * nwg + 2);
    // flags[0]: error bits of the write pass; flags[8 + i]: "changed" of fix pass i of a group
    constexpr uint32_t kFixGroup = 3;  // a pass settles chains inside workgroups; the next, their boundaries
    unsigned int* flags = c->misc.get<unsigned int>(8 + kFixGroup);
    float* tblk = c->tblk.get<float>(nt + 1);
    uint64_t stage_max = 0;
    for (uint64_t r = 0; r < L; r++)
        if (round_needs_staging(P.rounds[r])) stage_max = std::max(stage_max, P.rounds[r].count);
    int32_t* stage = stage_max ? c->stage.get<int32_t>(stage_max + 1) : nullptr;
    PivotDesc PD = pivot_desc(P);
    // host slots: the last fix pass's "changed" flag and the decoder's error bits
    uint32_t* hflag = c->pin_range.get<uint32_t>(2);
    launch_fill_i32(reinterpret_cast<int32_t*>(flags), 8 + kFixGroup, 0, st);
    HIPCHECK(hipEventRecord(c->ev[0], st));
    launch_dec_sync(dwords, dda, dwg, nwg, dtab, dstart, dex0, dcnt, st);
    // The whole pipeline is enqueued behind one group of fix passes and the
    // host waits once, at the end: parses from different starts merge within
    // a few codewords, so a group nearly always reaches the fixed point (a
    // pass after convergence changes nothing and its workgroups skip the table
    // load).  Every parse stays inside its subsequence's bits, so a pipeline
    // run before convergence reads and writes in bounds; if the group's last
    // pass still changed an exit, further groups run (one host wait each)
    // until none does, and then the pipeline runs again.
    auto fix_group = [&] {
        launch_fill_i32(reinterpret_cast<int32_t*>(flags + 8), kFixGroup, 0, st);
        for (uint32_t i = 0; i < kFixGroup; i++) {
            launch_dec_fix(dwords, dda, dwg, nwg, dtab, dstart, dex0, dex1, dcnt, flags + 8 + i, st);
            std::swap(dex0, dex1);
        }
        HIPCHECK(hipMemcpyAsync(hflag, flags + 8 + kFixGroup - 1, 4, hipMemcpyDeviceToHost, st));
    };
    auto pipeline = [&] {
        c->kev_n = 0;
        int kd = kev_begin(c, KEV_DECODE, 0);
        launch_fill_i32(reinterpret_cast<int32_t*>(flags), 1, 0, st);
        launch_dec_offsets(dda, (uint32_t)da.size(), dwg, nwg, dcnt, dwgs, dwgs + nwg + 1, flags, st);
        launch_dec_write(dwords, dda, dwg, nwg, dtab, dstart, dcnt, dwgs + nwg + 1, sym, flags, st);
        kev_end(c, kd);
        HIPCHECK(hipMemcpyAsync(hflag + 1, flags, 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipEventRecord(c->ev[1], st));
        // reconstruction: t floats, pivots, rounds (block-major QB order within each round array)
        HIPCHECK(hipEventRecord(c->ev[2], st));
        launch_t_float(sym, nbl, (uint32_t)L, (float)p.t_precision, tblk, st);
        int kp = kev_begin(c, KEV_PIVOT, 0);
        launch_pivots(false, d_out, sym + out_off[1], PD, nullptr, st);
        kev_end(c, kp);
        for (uint64_t r = 0; r < L; r++) {
            RoundDesc R = P.rounds[r];
            check_grid(R, 1);
            R.qb_off = 0;
            int kr = kev_begin(c, KEV_ROUND, (int)r);
            if (round_needs_staging(R)) {
                // the round's QBs back into "012" order, then the round reads them there
                uint32_t cmax[3];
                round_cmax(P, R, cmax);
                launch_qb_order(R, tabs, cmax, stage, sym + out_off[r + 2], 0, 0, 1, false, st);
                R.order = 2;
                launch_round_decompress(R, Q.quintic, d_out, stage, tabs, tblk + r * nbl, dsegs, st);
            } else {
                launch_round_decompress(R, Q.quintic, d_out, sym + out_off[r + 2], tabs, tblk + r * nbl, dsegs, st);
            }
            kev_end(c, kr);
        }
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(c->ev[3], st));
        HIPCHECK(hipStreamSynchronize(st));
    };
    fix_group();
    pipeline();
    if (hflag[0]) {
        uint64_t iters = kFixGroup;
        do {
            iters += kFixGroup;
            if (iters > nsub + 2) fail(ASZ_ERR_DATA, "Huffman decoder did not synchronise");
            fix_group();
            HIPCHECK(hipStreamSynchronize(st));
        } while (hflag[0]);
        pipeline();
    }
    const unsigned int derr = hflag[1];
    if (derr & 1) fail(ASZ_ERR_DATA, "Huffman payload length does not match its value count");
    if (derr & 2) fail(ASZ_ERR_DATA, "Huffman payload contains too many values");
    if (derr & 4) fail(ASZ_ERR_DATA, "Huffman payload ends inside a code");
    if (qb_out) {  // the pivot and round arrays, contiguous on the host
        uint64_t o = 0;
        for (uint64_t k = 1; k < L + 2; k++) {
            if (Q.counts[k]) HIPCHECK(hipMemcpyAsync(qb_out + o, sym + out_off[k], Q.counts[k] * 4,
                                                     hipMemcpyDeviceToHost, st));
            o += Q.counts[k];
        }
    }
    std::vector<int32_t> t32;
    if (t_out) {
        t32.resize(Q.nt);
        HIPCHECK(hipMemcpyAsync(t32.data(), sym, Q.nt * 4, hipMemcpyDeviceToHost, st));
    }
    HIPCHECK(hipStreamSynchronize(st));
    if (t_out)
        for (uint64_t i = 0; i < Q.nt; i++) t_out[i] = t32[i];
    for (uint64_t k = 0; k <= L; k++) S.qb_counts[k] = Q.counts[k + 1];
    kev_collect(c);
    S.huffman_ms = ev_ms(c->ev[0], c->ev[1]);
    if (S.decode_ms == 0) S.decode_ms = S.huffman_ms;
    S.rounds_ms = ev_ms(c->ev[2], c->ev[3]);
    S.total_gpu_ms = S.huffman_ms + S.rounds_ms;
    S.wall_ms = ms_since(wall0);
    (void)N;
}

}  // namespace asz

// ================================================================== C ABI

using asz::fail;

extern "C" {

void asz_params_default(asz_params* p) {
    p->rel_eb = 1e-3f;
    p->pivot_eb_ratio = 0.05f;
    p->eb_ratio = 0.76f;
    p->slope_ratio = 0.54f;
    p->block_pivot_ratio = 4;
    p->t_precision = 32;
    p->stat_eb_ratio = 0.2f;
    p->h1_block_sampling_num = 256;
    p->t_min = 1.0f / 16;
    p->t_max = 6.0f / 16;
    p->default_slope_eb_span = 1.5f;
    p->default_slope_invalid_ratio_threshold = 1.0f;
    p->default_slope_fallback = 1000.0f;
    p->default_slope = 1000.0f;
    p->quintic = 0;
}

int asz_ctx_create(int device, asz_ctx** out) {
    *out = nullptr;
    asz_ctx* c = new asz_ctx();
    int rc = guarded(c, [&] {
        int n = 0;
        HIPCHECK(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) fail(ASZ_ERR_ARG, "no such HIP device");
        HIPCHECK(hipSetDevice(device));
        c->device = device;
        asz::noalloc_setup(device);
        HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (auto& e : c->ev) HIPCHECK(hipEventCreate(&e));
    });
    if (rc != ASZ_OK) {
        std::string msg = c->err;
        delete c;
        fprintf(stderr, "asz_ctx_create: %s\n", msg.c_str());
        return rc;
    }
    *out = c;
    return ASZ_OK;
}

void asz_ctx_destroy(asz_ctx* c) {
    if (!c) return;
    for (asz_ctx* k : c->kids) asz_ctx_destroy(k);
    c->kids.clear();
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    for (auto& e : c->ev) (void)hipEventDestroy(e);
    if (c->hres_host) (void)hipHostFree(c->hres_host);
    for (hipEvent_t e : c->side_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->side) (void)hipStreamDestroy(c->side);
    for (auto& e : c->kev) (void)hipEventDestroy(e);
    for (auto& g : c->graph_cache) g.release();
    c->graph_cache.clear();
    if (c->order_ev) (void)hipEventDestroy(c->order_ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* asz_last_error(asz_ctx* c) { return c ? c->err.c_str() : "no context"; }

int asz_set_option(asz_ctx* c, int option, int64_t value) {
    return guarded(c, [&] {
        if (option == ASZ_OPT_STREAM_ON_DEVICE) {
            c->stream_on_device = value != 0;
            return;
        }
        if (option == ASZ_OPT_WIDE_CODES) {
            c->wide_codes = value != 0;
            return;
        }
        if (option == ASZ_OPT_GRAPHS) {
            c->graphs = value != 0;
            return;
        }
        if (option == ASZ_OPT_COARSE_ROUNDS) {
            if (value < -1 || value > ASZ_MAX_LAYERS) fail(ASZ_ERR_ARG, "coarse rounds must be in [-1, 63]");
            c->coarse_max = (int)value;
            return;
        }
        if (option == ASZ_OPT_BATCH_FIELDS) {
            if (value < 0 || value > 65535) fail(ASZ_ERR_ARG, "fields per chunk must be in [0, 65535]");
            c->batch_fields = (uint32_t)value;
            return;
        }
        if (option != ASZ_OPT_KERNEL_EVENTS && option != ASZ_OPT_TIMED_ROUND) fail(ASZ_ERR_ARG, "unknown option");
        HIPCHECK(hipSetDevice(c->device));
        if (option == ASZ_OPT_TIMED_ROUND) {
            if (value >= ASZ_MAX_LAYERS) fail(ASZ_ERR_ARG, "no such round");
            c->kev_round = value < 0 ? -1 : (int)value;
            value = value >= 0 ? 1 : 0;
        } else {
            c->kev_round = -1;
        }
        c->kev_on = value != 0;
        if (c->kev_on && c->kev.empty()) {
            c->kev.resize(2 * 256);
            c->kev_tag.resize(256);
            for (auto& e : c->kev) HIPCHECK(hipEventCreate(&e));
        }
    });
}

int asz_stream_wait(asz_ctx* c, void* stream) {
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        if (!c->order_ev) HIPCHECK(hipEventCreateWithFlags(&c->order_ev, hipEventDisableTiming));
        HIPCHECK(hipEventRecord(c->order_ev, static_cast<hipStream_t>(stream)));
        HIPCHECK(hipStreamWaitEvent(c->stream, c->order_ev, 0));
    });
}

int asz_get_stats(asz_ctx* c, asz_stats* out) {
    *out = c->stats;
    return ASZ_OK;
}

int asz_device_alloc(asz_ctx* c, uint64_t bytes, void** d_ptr) {
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        if (hipMalloc(d_ptr, bytes ? bytes : 1) != hipSuccess)
            fail(ASZ_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    });
}

int asz_device_free(asz_ctx* c, void* d_ptr) {
    return guarded(c, [&] { HIPCHECK(hipFree(d_ptr)); });
}

int asz_copy_to_device(asz_ctx* c, void* d_dst, const void* src, uint64_t bytes) {
    return guarded(c, [&] {
        HIPCHECK(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

int asz_copy_to_host(asz_ctx* c, void* dst, const void* d_src, uint64_t bytes) {
    return guarded(c, [&] {
        HIPCHECK(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

int asz_geometry(const uint64_t dims[3], const asz_plan* plan, uint64_t bpr, uint64_t ps[3], uint64_t bs[3],
                 uint64_t bn[3]) {
    return guarded(nullptr, [&] {
        asz::Geometry g = asz::make_geometry(dims, *plan, bpr);
        for (int a = 0; a < 3; a++) { ps[a] = g.pivot_stride[a]; bs[a] = g.block_shape[a]; bn[a] = g.block_num[a]; }
    });
}

int asz_range_f32(asz_ctx* c, const float* d_field, uint64_t n, float* mn, float* mx) {
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        if (n == 0) fail(ASZ_ERR_ARG, "empty field");
        auto t0 = asz::Clock::now();
        asz::range_scan(c, d_field, n, mn, mx);
        c->stats.range_ms = asz::ms_since(t0);
    });
}

int asz_default_slope(asz_ctx* c, const float* d_field, const uint64_t dims[3], const asz_params* p, float mn,
                      float mx, float* slope, float* inv, int* fb, int* reason) {
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        auto t0 = asz::Clock::now();
        asz::SlopeResult r = asz::default_slope(c, d_field, dims, *p, mn, mx);
        for (const auto& w : r.warnings)
            printf("Warning: for rel_eb %f, only %zu/%zu valid default-slope cases were found\n", (double)w.rel,
                   (size_t)w.valid, (size_t)w.total);
        *slope = r.slope;
        *inv = r.max_invalid_ratio;
        *fb = r.fallback ? 1 : 0;
        *reason = r.reason;
        c->stats.preprocess_ms = asz::ms_since(t0);
    });
}

int asz_compress(asz_ctx* c, float* d_work, const uint64_t dims[3], const asz_plan* plan, const asz_params* p,
                 float mn, float mx, int tsel_mode, const int64_t* t_inject, const uint8_t** stream,
                 uint64_t* stream_bytes, int32_t* qb_out, int64_t* t_out) {
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        asz::compress_impl(c, d_work, dims, *plan, *p, mn, mx, tsel_mode, t_inject, qb_out, t_out);
        *stream = c->out_stream;
        *stream_bytes = c->out_bytes;
    });
}

int asz_compress_field(asz_ctx* c, float* d_work, const uint64_t dims[3], const asz_plan* plan, const asz_params* p,
                       int tsel_mode, const uint8_t** stream, uint64_t* stream_bytes, float* mn, float* mx) {
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        const uint64_t N = dims[0] * dims[1] * dims[2];
        if (N == 0) fail(ASZ_ERR_ARG, "empty field");
        const asz::FieldIn f{d_work, 0.0f, 0.0f, p->default_slope};
        asz::compress_fields(c, 1, &f, dims, *plan, *p, tsel_mode, nullptr, nullptr, nullptr, -1, true);
        *stream = c->out_stream;
        *stream_bytes = c->out_bytes;
        *mn = c->dev_mn;
        *mx = c->dev_mx;
    });
}

int asz_compress_host(asz_ctx* c, const float* field, const uint64_t dims[3], const asz_plan* plan,
                      const asz_params* p, int tsel_mode, const int64_t* t_inject, const uint8_t** stream,
                      uint64_t* stream_bytes, float* recon, int32_t* qb_out, int64_t* t_out) {
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        const uint64_t N = dims[0] * dims[1] * dims[2];
        float* w = c->work.get<float>(N);
        HIPCHECK(hipMemcpyAsync(w, field, N * 4, hipMemcpyHostToDevice, c->stream));
        float mn, mx;
        asz::range_scan(c, w, N, &mn, &mx);
        const bool on_dev = c->stream_on_device;
        c->stream_on_device = false;  // this entry point returns a host stream
        try {
            asz::compress_impl(c, w, dims, *plan, *p, mn, mx, tsel_mode, t_inject, qb_out, t_out);
        } catch (...) {
            c->stream_on_device = on_dev;
            throw;
        }
        c->stream_on_device = on_dev;
        if (recon) {
            HIPCHECK(hipMemcpyAsync(recon, w, N * 4, hipMemcpyDeviceToHost, c->stream));
            HIPCHECK(hipStreamSynchronize(c->stream));
        }
        *stream = c->out_stream;
        *stream_bytes = c->out_bytes;
    });
}

int asz_prepare_round(asz_ctx* c, float* d_work, const uint64_t dims[3], const asz_plan* plan, const asz_params* p,
                      float mn, float mx, int round, const int64_t* t_int) {
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        if (round < 0 || round >= plan->layers) fail(ASZ_ERR_ARG, "no such round");
        if (!t_int) fail(ASZ_ERR_ARG, "t values are required");
        asz::compress_impl(c, d_work, dims, *plan, *p, mn, mx, ASZ_TSEL_INJECT, t_int, nullptr, nullptr, round);
    });
}

static int round_energy(asz_ctx* c, const float* d_work, const uint64_t dims[3], const asz_plan* plan,
                        const asz_params* p, int round, int quintic, const float* d_theta, const float* d_gE,
                        double* d_E, float* d_g) {
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        if (round < 0 || round >= plan->layers) fail(ASZ_ERR_ARG, "no such round");
        asz::Geometry g = asz::make_geometry(dims, *plan, p->block_pivot_ratio);
        asz::Plan P = asz::make_plan(g, 1.0f, p->pivot_eb_ratio, p->eb_ratio, p->slope_ratio, 1.0f,
                                     p->h1_block_sampling_num);
        asz::launch_energy(P.rounds[round], g.nblocks(), quintic != 0, d_gE != nullptr, d_work, d_theta, d_gE, d_E,
                           d_g, c->stream);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

int asz_round_energy(asz_ctx* c, const float* d_work, const uint64_t dims[3], const asz_plan* plan,
                     const asz_params* p, int round, int quintic, const float* d_theta, double* d_E) {
    return round_energy(c, d_work, dims, plan, p, round, quintic, d_theta, nullptr, d_E, nullptr);
}

int asz_round_energy_adjoint(asz_ctx* c, const float* d_work, const uint64_t dims[3], const asz_plan* plan,
                             const asz_params* p, int round, int quintic, const float* d_theta, const float* d_gE,
                             float* d_gtheta) {
    if (!d_gE) {
        if (c) c->err = "the upstream gradient is required";
        return ASZ_ERR_ARG;
    }
    return round_energy(c, d_work, dims, plan, p, round, quintic, d_theta, d_gE, nullptr, d_gtheta);
}

int asz_compress_batch(asz_ctx* c, uint32_t nfields, float* const* d_fields, const uint64_t* dims,
                       const asz_plan* plans, const asz_params* p, int tsel_mode, uint32_t concurrency,
                       uint8_t** streams, uint64_t* stream_bytes, asz_batch_field* info) {
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        if (nfields == 0) return;
        if (tsel_mode == ASZ_TSEL_INJECT) fail(ASZ_ERR_ARG, "t injection is not available in batches");
        for (uint32_t i = 0; i < nfields; i++) streams[i] = nullptr;
        const uint32_t T = std::max<uint32_t>(1, std::min<uint32_t>(concurrency ? concurrency : 6, nfields));
        // Fields of one shape and plan (time steps of a series, same-shape
        // fields) are compressed together in chunks: every per-round kernel of
        // a chunk covers all of its fields.  Chunks hold about a T-th of their
        // group, at most kMaxChunk (or ASZ_OPT_BATCH_FIELDS) fields and
        // kMaxChunkPoints points.
        const uint32_t kMaxChunk = c->batch_fields ? c->batch_fields : 16;
        constexpr uint64_t kMaxChunkPoints = 1ull << 30;
        const bool can_group = tsel_mode == ASZ_TSEL_GPU || tsel_mode == ASZ_TSEL_GRAD;
        std::vector<std::vector<uint32_t>> groups;
        for (uint32_t i = 0; i < nfields; i++) {
            bool placed = false;
            for (auto& gr : groups) {
                const uint32_t j = gr[0];
                if (can_group && std::memcmp(dims + 3 * (uint64_t)i, dims + 3 * (uint64_t)j, 24) == 0 &&
                    asz::same_plan(plans[i], plans[j])) {
                    gr.push_back(i);
                    placed = true;
                    break;
                }
            }
            if (!placed) groups.push_back({i});
        }
        std::vector<std::vector<uint32_t>> chunks;
        for (const auto& gr : groups) {
            const uint64_t* d = dims + 3 * (uint64_t)gr[0];
            const uint64_t n = std::max<uint64_t>(1, d[0] * d[1] * d[2]);
            uint64_t cs = (gr.size() + T - 1) / T;
            cs = std::max<uint64_t>(1, std::min<uint64_t>({cs, kMaxChunk, kMaxChunkPoints / n}));
            for (size_t i0 = 0; i0 < gr.size(); i0 += cs)
                chunks.emplace_back(gr.begin() + i0, gr.begin() + std::min<size_t>(gr.size(), i0 + cs));
        }
        auto points = [&](const std::vector<uint32_t>& ch) {
            const uint64_t* d = dims + 3 * (uint64_t)ch[0];
            return d[0] * d[1] * d[2] * ch.size();
        };
        std::stable_sort(chunks.begin(), chunks.end(),
                         [&](const auto& x, const auto& y) { return points(x) > points(y); });
        // largest, smallest, second largest, second smallest, ...: large and
        // small chunks interleave, so the small chunks' latency-bound rounds
        // overlap the large chunks' bandwidth-bound ones; the middle-sized
        // chunks come last
        {
            std::vector<std::vector<uint32_t>> mixed;
            for (size_t lo = 0, hi = chunks.size(); lo < hi;) {
                mixed.push_back(std::move(chunks[lo++]));
                if (lo < hi) mixed.push_back(std::move(chunks[--hi]));
            }
            chunks.swap(mixed);
        }
        const uint32_t W = std::min<uint32_t>(T, (uint32_t)chunks.size());
        while (c->kids.size() < W) {
            asz_ctx* k = nullptr;
            const int rc = asz_ctx_create(c->device, &k);
            if (rc != ASZ_OK) fail(rc, "cannot create a batch worker context");
            c->kids.push_back(k);
        }
        const bool on_device = c->stream_on_device, wide = c->wide_codes;
        const int coarse_max = c->coarse_max;
        const bool graphs = c->graphs;
        // the workers' streams start after everything enqueued on this context's
        // stream (asz_stream_wait orders that after the caller's)
        if (!c->order_ev) HIPCHECK(hipEventCreateWithFlags(&c->order_ev, hipEventDisableTiming));
        HIPCHECK(hipEventRecord(c->order_ev, c->stream));
        for (uint32_t t = 0; t < W; t++) HIPCHECK(hipStreamWaitEvent(c->kids[t]->stream, c->order_ev, 0));
        std::atomic<uint32_t> next{0};
        std::atomic<uint64_t> graph_launches{0}, graph_captures{0};
        std::mutex err_mu, pre_mu;
        std::string first_err;
        int first_code = ASZ_OK;
        const uint32_t nchunks = (uint32_t)chunks.size();
        auto worker = [&](asz_ctx* k) {
            (void)hipSetDevice(k->device);
            for (;;) {
                const uint32_t ci = next.fetch_add(1);
                if (ci >= nchunks) return;
                const std::vector<uint32_t>& ch = chunks[ci];
                const uint32_t nf = (uint32_t)ch.size();
                const int rc = guarded(k, [&] {
                    const uint64_t* d = dims + 3 * (uint64_t)ch[0];
                    const uint64_t n = d[0] * d[1] * d[2];
                    std::vector<const float*> ptrs(nf);
                    std::vector<uint64_t> ns(nf, n);
                    std::vector<float> mn(nf), mx(nf);
                    for (uint32_t j = 0; j < nf; j++) ptrs[j] = d_fields[ch[j]];
                    // the workers take turns at the range scan and the slope
                    // samples, so the first chunk's rounds start after its own
                    // scan instead of after every worker's
                    std::unique_lock<std::mutex> hold(pre_mu);
                    asz::range_scan_fields(k, nf, ptrs.data(), ns.data(), mn.data(), mx.data(), ch.data());
                    std::vector<asz::SlopeResult> sr(nf);
                    asz::default_slope_fields(k, nf, ptrs.data(), d, *p, mn.data(), mx.data(), sr.data(), &hold);
                    std::vector<asz::FieldIn> fin(nf);
                    for (uint32_t j = 0; j < nf; j++) fin[j] = asz::FieldIn{d_fields[ch[j]], mn[j], mx[j], sr[j].slope};
                    k->stream_on_device = on_device;
                    k->wide_codes = wide;
                    k->coarse_max = coarse_max;
                    k->graphs = graphs;
                    k->tunes = c->tunes;
                    asz::compress_fields(k, nf, fin.data(), d, plans[ch[0]], *p, tsel_mode, nullptr, nullptr,
                                         nullptr);
                    graph_launches += k->stats.graph_launched;
                    graph_captures += k->stats.graph_captured;
                    for (uint32_t j = 0; j < nf; j++) {
                        const uint32_t i = ch[j];
                        const auto& o = k->fouts[j];
                        uint8_t* out = nullptr;
                        if (on_device) {
                            // a device buffer of the stream's size, filled device to device
                            if (hipMalloc(&out, o.second ? o.second : 1) != hipSuccess)
                                fail(ASZ_ERR_NOMEM, "device allocation failed");
                            streams[i] = out;
                            HIPCHECK(hipMemcpyAsync(out, o.first, o.second, hipMemcpyDeviceToDevice, k->stream));
                        } else {
                            out = static_cast<uint8_t*>(std::malloc(o.second ? o.second : 1));
                            if (!out) fail(ASZ_ERR_NOMEM, "host allocation failed");
                            streams[i] = out;
                            std::memcpy(out, o.first, o.second);
                        }
                        stream_bytes[i] = o.second;
                        if (info) {
                            const asz_stats& Sf = k->fstats[j];
                            info[i].min_val = mn[j];
                            info[i].max_val = mx[j];
                            info[i].default_slope = sr[j].slope;
                            info[i].pad = 0;
                            info[i].max_err = Sf.max_err;
                            info[i].sq_err_sum = Sf.sq_err_sum;
                            info[i].t_fallbacks = Sf.t_fallbacks;
                        }
                    }
                    if (on_device) HIPCHECK(hipStreamSynchronize(k->stream));
                });
                if (rc != ASZ_OK) {
                    std::lock_guard<std::mutex> lk(err_mu);
                    if (first_code == ASZ_OK) {
                        first_code = rc;
                        // errors that concern one field name it; others name the chunk
                        if (k->err.rfind("field ", 0) == 0) {
                            first_err = k->err;
                        } else {
                            first_err = nf == 1 ? "field " : "fields ";
                            for (uint32_t j = 0; j < nf; j++) first_err += (j ? ", " : "") + std::to_string(ch[j]);
                            first_err += ": " + k->err;
                        }
                    }
                    next = nchunks;  // stop handing out chunks
                    return;
                }
            }
        };
        std::vector<std::thread> ts;
        for (uint32_t t = 0; t < W; t++) ts.emplace_back(worker, c->kids[t]);
        for (auto& t : ts) t.join();
        // chunks whose device sequence ran as a graph, and how many were captured
        c->stats.graph_launched = graph_launches;
        c->stats.graph_captured = graph_captures;
        if (first_code != ASZ_OK) {
            for (uint32_t i = 0; i < nfields; i++) {
                if (on_device) (void)hipFree(streams[i]);
                else std::free(streams[i]);
                streams[i] = nullptr;
            }
            fail(first_code, first_err);
        }
    });
}

int asz_decompress(asz_ctx* c, const uint8_t* stream, uint64_t n, const uint64_t dims[3], const asz_params* p,
                   float* d_out, int32_t* qb_out, int64_t* t_out) {
    const int rc = guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        asz::decompress_impl(c, stream, n, dims, *p, d_out, qb_out, t_out);
    });
    // an error can leave the stream's upload running: it reads the caller's buffer
    if (rc != ASZ_OK) (void)hipStreamSynchronize(c->stream);
    return rc;
}

int asz_decompress_device(asz_ctx* c, const uint8_t* d_stream, uint64_t n, const uint64_t dims[3],
                          const asz_params* p, float* d_out) {
    const int rc = guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        asz::decompress_impl(c, d_stream, n, dims, *p, d_out, nullptr, nullptr, true);
    });
    if (rc != ASZ_OK) (void)hipStreamSynchronize(c->stream);
    return rc;
}

int asz_decompress_host(asz_ctx* c, const uint8_t* stream, uint64_t n, const uint64_t dims[3], const asz_params* p,
                        float* out) {
    const int rc = guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        const uint64_t N = dims[0] * dims[1] * dims[2];
        float* w = c->work.get<float>(N);
        asz::decompress_impl(c, stream, n, dims, *p, w, nullptr, nullptr);
        HIPCHECK(hipMemcpyAsync(out, w, N * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
    if (rc != ASZ_OK) (void)hipStreamSynchronize(c->stream);
    return rc;
}

}  // extern "C"
