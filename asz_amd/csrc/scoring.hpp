// Host scoring of AB-key samples (see preprocess.cpp).
#pragma once

#include <cstdint>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "internal.hpp"

namespace asz {

struct KeyHash {
    size_t operator()(const std::tuple<int64_t, int64_t>& k) const;
};
using ABMap = std::unordered_map<std::tuple<int64_t, int64_t>, uint64_t, KeyHash>;

void build_ab_map(const int64_t* keys, uint64_t n, ABMap& m);
// A map's entries in its iteration order, keys as the scoring loop converts them.
struct FlatAB {
    std::vector<float> a, b;
    std::vector<uint64_t> n;
};
void flatten_ab(const ABMap& m, FlatAB& f);
void h1_psnr(const FlatAB& f, float t, float abs_eb, float stat_eb, float min_val, float max_val, float& h1,
             float& psnr);
float select_t_exact(const int64_t* keys, uint64_t n, float layer_eb, float stat_eb, float layer_slope, float min_val,
                     float max_val);

struct SlopeWarning {
    float rel;
    uint64_t valid, total;
};
struct SlopeResult {
    float slope = 0;
    float max_invalid_ratio = 0;
    bool fallback = false;
    int reason = 0;  // 0 none, 1 invalid_ratio, 2 invalid_difference
    std::vector<SlopeWarning> warnings;
};
SlopeResult default_slope_from_keys(const std::vector<std::vector<int64_t>>& keys_per_dir, float abs_eb,
                                    float stat_eb_ratio, float eb_span, float invalid_threshold, float fallback,
                                    float t_min, float t_max, int t_precision, float min_val, float max_val);

}  // namespace asz
