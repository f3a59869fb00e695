// Host scoring of GPU-sampled AB keys in the reference's order.
//
// get_default_slope (get_default_slope.tpp:5-83) and the exact per-block t
// scoring (compress.tpp:90-144) both build an unordered_map of (A_key, B_key)
// counts and sum float terms in its iteration order (get_h_psnr.tpp:5-31).
// The keys are computed on the GPU with the reference's arithmetic; the map
// and the float sums run here with libstdc++'s unordered_map, the hash the
// reference specialises (asz_param.hpp:22-40), and GCC's lowering of the
// scoring loop (fmaf where GCC contracts, double-rounded MSE steps), so the
// results are the reference's bit for bit.
#include <cfloat>
#include <algorithm>
#include <cmath>
#include <memory>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "scoring.hpp"

namespace asz {

namespace {
inline int64_t lround_f(float y) {
    float s = y + std::copysign(0.49999997f, y);
    if (!(s > -9223372036854775808.0f && s < 9223372036854775808.0f)) return INT64_MIN;
    return (int64_t)s;
}
}  // namespace

size_t KeyHash::operator()(const std::tuple<int64_t, int64_t>& k) const {
    uint64_t h = 0;
    h = h * 1145141919810ull + (uint64_t)std::get<0>(k);
    h = h * 1145141919810ull + (uint64_t)std::get<1>(k);
    return (size_t)h;
}

void build_ab_map(const int64_t* keys, uint64_t n, ABMap& m) {
    for (uint64_t i = 0; i < n; i++) m[std::make_tuple(keys[2 * i], keys[2 * i + 1])]++;
}

void flatten_ab(const ABMap& m, FlatAB& f) {
    f.a.clear();
    f.b.clear();
    f.n.clear();
    f.a.reserve(m.size());
    f.b.reserve(m.size());
    f.n.reserve(m.size());
    for (const auto& kv : m) {
        f.a.push_back((float)std::get<0>(kv.first));
        f.b.push_back((float)std::get<1>(kv.first));
        f.n.push_back(kv.second);
    }
}

namespace {
// Quantizer bins in [-kQbDirect, kQbDirect) are counted in a flat array (per
// thread, all zero between calls); the rest in a hash map.
constexpr int64_t kQbDirect = 4096;
struct QbScratch {
    std::vector<uint64_t> cnt = std::vector<uint64_t>(2 * kQbDirect, 0);
    std::vector<int64_t> order;  // distinct bins in first-occurrence order
};
QbScratch& qb_scratch() {
    thread_local QbScratch s;
    return s;
}

// Bump allocation for the per-call bin map (its nodes and bucket arrays),
// released all at once when the next call starts.  The map's layout does not
// depend on its allocator.
struct BinArena {
    std::vector<std::unique_ptr<char[]>> blocks;
    size_t used = 0, cap = 0;
    void reset() {
        if (blocks.size() > 1) {  // keep the largest block
            std::unique_ptr<char[]> last = std::move(blocks.back());
            blocks.clear();
            blocks.push_back(std::move(last));
        }
        used = 0;
    }
    void* get(size_t n, size_t align) {
        used = (used + align - 1) & ~(align - 1);
        if (blocks.empty() || used + n > cap) {
            cap = std::max<size_t>(2 * cap, std::max<size_t>(n + align, 1 << 16));
            blocks.emplace_back(new char[cap]);
            used = 0;
        }
        void* p = blocks.back().get() + used;
        used += n;
        return p;
    }
};
BinArena& bin_arena() {
    thread_local BinArena a;
    return a;
}
template <typename T>
struct BinAlloc {
    using value_type = T;
    BinArena* a;
    explicit BinAlloc(BinArena* x) : a(x) {}
    template <typename U>
    BinAlloc(const BinAlloc<U>& o) : a(o.a) {}
    T* allocate(size_t n) { return static_cast<T*>(a->get(n * sizeof(T), alignof(T))); }
    void deallocate(T*, size_t) {}
    template <typename U>
    bool operator==(const BinAlloc<U>& o) const { return a == o.a; }
    template <typename U>
    bool operator!=(const BinAlloc<U>& o) const { return a != o.a; }
};
using BinMap = std::unordered_map<int64_t, uint64_t, std::hash<int64_t>, std::equal_to<int64_t>,
                                  BinAlloc<std::pair<const int64_t, uint64_t>>>;
}  // namespace

void h1_psnr(const FlatAB& f, float t, float abs_eb, float stat_eb, float min_val, float max_val, float& h1,
             float& psnr) {
    const float neg_2_eb = 0.5f / abs_eb;
    const float two_t = t * 2.0f;
    const float two_eb = abs_eb * 2.0f;
    QbScratch& S = qb_scratch();
    S.order.clear();
    std::unordered_map<int64_t, uint64_t> far;
    float mse = 0.0f;
    uint64_t n = 0;
    const size_t m = f.n.size();
    for (size_t i = 0; i < m; i++) {
        const float err = std::fmaf(f.b[i], 2.0f, f.a[i] * two_t) * stat_eb;
        const int64_t q = lround_f(err * neg_2_eb);
        const uint64_t c = f.n[i];  // >= 1
        if (q >= -kQbDirect && q < kQbDirect) {
            uint64_t& slot = S.cnt[(size_t)(q + kQbDirect)];
            if (!slot) S.order.push_back(q);
            slot += c;
        } else {
            auto it = far.find(q);
            if (it == far.end()) {
                S.order.push_back(q);
                far.emplace(q, c);
            } else {
                it->second += c;
            }
        }
        const float de = std::fmaf(-(float)q, two_eb, err);
        const double ded = (double)de;
        mse = (float)std::fma((double)c, ded * ded, (double)mse);
        n += c;
    }
    // The reference's bin map (get_h_psnr.tpp:13-19) receives its distinct
    // bins in first-occurrence order; its layout, and so the order of the
    // entropy sum below, depends on that sequence only, not on the counts.
    BinArena& arena = bin_arena();
    arena.reset();
    BinMap qb_map{BinAlloc<std::pair<const int64_t, uint64_t>>(&arena)};
    for (const int64_t q : S.order) {
        uint64_t c;
        if (q >= -kQbDirect && q < kQbDirect) {
            c = S.cnt[(size_t)(q + kQbDirect)];
            S.cnt[(size_t)(q + kQbDirect)] = 0;
        } else {
            c = far[q];
        }
        qb_map[q] = c;
    }
    const float dn = (float)n;
    const float rec = 1.0f / dn;
    h1 = 0.0f;
    for (const auto& kv : qb_map) {
        float p = (float)kv.second * rec;
        h1 = std::fmaf(-p, log2f(p), h1);
    }
    const float range = max_val - min_val;
    const float r2 = range * range;
    psnr = log10f((r2 / mse) * dn) * 10.0f;
}

float select_t_exact(const int64_t* keys, uint64_t n, float layer_eb, float stat_eb, float layer_slope, float min_val,
                     float max_val) {
    static const float cand[5] = {1.0f / 16, 1.0f / 4, 9.0f / 32, 5.0f / 16, 11.0f / 32};
    float best_t = 4.0f / 16;
    ABMap m;
    build_ab_map(keys, n, m);
    if (m.empty()) return best_t;
    FlatAB fm;
    flatten_ab(m, fm);
    float best_h1 = 0, best_psnr = 0;
    h1_psnr(fm, best_t, layer_eb, stat_eb, min_val, max_val, best_h1, best_psnr);
    for (float t : cand) {
        if (t == best_t || best_psnr > FLT_MAX) continue;
        float h1, psnr;
        h1_psnr(fm, t, layer_eb, stat_eb, min_val, max_val, h1, psnr);
        float score = std::fmaf(best_h1 - h1, layer_slope, psnr - best_psnr);
        if (score > 0.0f || psnr > FLT_MAX) { best_t = t; best_h1 = h1; best_psnr = psnr; }
    }
    return best_t;
}

SlopeResult default_slope_from_keys(const std::vector<std::vector<int64_t>>& keys_per_dir, float abs_eb,
                                    float stat_eb_ratio, float eb_span, float invalid_threshold, float fallback,
                                    float t_min, float t_max, int t_precision, float min_val, float max_val) {
    SlopeResult res;
    const float ebs[2] = {abs_eb * eb_span, abs_eb / eb_span};
    const float base = abs_eb * stat_eb_ratio;
    const float stat_eb = base * 0.5f;
    std::vector<float> cand;
    int64_t tlo = lround_f(t_min * (float)t_precision), thi = lround_f(t_max * (float)t_precision);
    for (int64_t ti = tlo; ti <= thi; ti++) cand.push_back((float)ti / (float)t_precision);
    float avg_h1[2] = {0, 0}, avg_psnr[2] = {0, 0};
    uint64_t valid[2] = {0, 0};
    for (int d = 0; d < 3; d++) {
        ABMap m;
        build_ab_map(keys_per_dir[d].data(), keys_per_dir[d].size() / 2, m);
        if (m.empty())
            fail(ASZ_ERR_DATA, "In get_default_slope, direction " + std::to_string(d) + " AB_map is empty");
        FlatAB fm;
        flatten_ab(m, fm);
        for (int i = 0; i < 2; i++)
            for (float t : cand) {
                float h1, psnr;
                h1_psnr(fm, t, ebs[i], stat_eb, min_val, max_val, h1, psnr);
                if (psnr > FLT_MAX) continue;
                avg_h1[i] += h1;
                avg_psnr[i] += psnr;
                valid[i]++;
            }
    }
    const uint64_t total = 3 * cand.size();
    const float rec = 1.0f / (float)total;
    float max_invalid = 0.0f;
    for (int i = 0; i < 2; i++) {
        max_invalid = std::max(max_invalid, std::fmaf(-(float)valid[i], rec, 1.0f));
        if (valid[i] != total)
            res.warnings.push_back({ebs[i] / (max_val - min_val), valid[i], total});
    }
    res.max_invalid_ratio = max_invalid;
    if (max_invalid >= invalid_threshold) {
        res.slope = fallback;
        res.fallback = true;
        res.reason = 1;
        return res;
    }
    for (int i = 0; i < 2; i++) { avg_h1[i] /= (float)valid[i]; avg_psnr[i] /= (float)valid[i]; }
    const float dh = avg_h1[1] - avg_h1[0];
    const float dp = avg_psnr[1] - avg_psnr[0];
    const float scale = std::max(std::fabs(avg_h1[1]), std::fabs(avg_h1[0]));
    const float tol = scale > 1.0f ? scale * 7.62939453125e-6f : 7.62939453125e-6f;
    if (dh <= tol || dp <= 0.0f) {
        res.slope = fallback;
        res.fallback = true;
        res.reason = 2;
        return res;
    }
    res.slope = dp / dh;
    return res;
}

}  // namespace asz
