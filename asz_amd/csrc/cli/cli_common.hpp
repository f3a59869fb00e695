// Shared pieces of the drop-in executables: the reference's argument grammar,
// validation messages and printed configuration (s1_compress.cpp,
// s1_decompress.cpp, generate_config.tpp:5-49).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "asz_gpu.h"

struct CliState {
    asz_params p;
    std::string in_path, z_path, o_path;
    std::vector<uint64_t> dims;
    std::vector<std::string> plan;
    bool auto_family = true;
    bool t_min_set = false, t_max_set = false;
    bool report_t_histogram = false;
    uint64_t max_layer_num = 14, full_predictor_search_layer_num = 4, generate_config_sample_stride = 5;
    uint64_t pf_suffix_layers = 6, pf_sample_blocks = 8, pf_sample_point_limit = 32768;
    float pf_sample_fraction_limit = 0.02f;
    uint64_t pf_coarse_bpr = 3, pf_fine_bpr = 4;
    float pf_fine_eb_threshold = 1e-4f;
    int tsel = ASZ_TSEL_EXACT;  // the reference's t, block for block (compress.tpp:90-144)
    int device = 0;
    int zstd_level = 3;         // compress.tpp:297
    uint32_t zstd_threads = 1;
};

// s1_compress.cpp:5-28
inline std::vector<std::string> parse_predictor_plan(const std::string& text) {
    std::vector<std::string> out;
    size_t begin = 0;
    while (begin <= text.size()) {
        size_t end = text.find(',', begin);
        std::string pr = text.substr(begin, end == std::string::npos ? std::string::npos : end - begin);
        bool valid = false;
        for (int d = 0; d < 3; d++)
            for (const char* o : {"120", "021", "012"})
                if (pr == "cubic" + std::to_string(d) + "_" + o) valid = true;
        if (!valid) throw std::invalid_argument("invalid predictor in plan: " + pr);
        out.push_back(pr);
        if (end == std::string::npos) break;
        begin = end + 1;
    }
    if (out.empty()) throw std::invalid_argument("predictor plan cannot be empty");
    return out;
}

inline asz_plan to_plan(const std::vector<std::string>& v) {
    asz_plan pl{};
    pl.layers = (int32_t)v.size();
    for (size_t r = 0; r < v.size() && r < ASZ_MAX_LAYERS; r++) {
        pl.axis[r] = (uint8_t)(v[r][5] - '0');
        const std::string o = v[r].substr(7);
        pl.order[r] = o == "120" ? 0 : o == "021" ? 1 : 2;
    }
    return pl;
}

inline std::vector<std::string> from_plan(const asz_plan& pl) {
    std::vector<std::string> v;
    static const char* ord[3] = {"120", "021", "012"};
    for (int r = 0; r < pl.layers; r++) v.push_back("cubic" + std::to_string(pl.axis[r]) + "_" + ord[pl.order[r]]);
    return v;
}

inline int64_t lround_f32(float y) {
    float s = y + std::copysign(0.49999997f, y);
    return (int64_t)s;
}

// generate_config.tpp:5-49
inline void print_config(const CliState& s, const std::vector<std::string>& plan, const uint64_t ps[3],
                         const uint64_t bs[3], const uint64_t bn[3]) {
    const asz_params& p = s.p;
    printf("predictor_vec={");
    for (size_t i = 0; i < plan.size(); i++) printf(i + 1 == plan.size() ? "\"%s\"" : "\"%s\",", plan[i].c_str());
    printf("};\n");
    printf("pivot_eb_ratio=%.2f;\n", p.pivot_eb_ratio);
    printf("eb_ratio=%.2f;\n", p.eb_ratio);
    printf("slope_ratio=%.2f;\n", p.slope_ratio);
    printf("max_layer_num=%lu;\n", (unsigned long)s.max_layer_num);
    printf("full_predictor_search_layer_num=%lu;\n", (unsigned long)s.full_predictor_search_layer_num);
    printf("generate_config_sample_stride=%lu;\n", (unsigned long)s.generate_config_sample_stride);
    printf("h1_block_sampling_num=%lu;\n", (unsigned long)p.h1_block_sampling_num);
    printf("stat_eb_ratio=%.3f;\n", p.stat_eb_ratio);
    printf("predictor_family_suffix_layers=%lu;\n", (unsigned long)s.pf_suffix_layers);
    printf("predictor_family_sample_blocks=%lu;\n", (unsigned long)s.pf_sample_blocks);
    printf("predictor_family_sample_point_limit=%lu;\n", (unsigned long)s.pf_sample_point_limit);
    printf("predictor_family_sample_fraction_limit=%.6f;\n", s.pf_sample_fraction_limit);
    printf("predictor_family_coarse_probe_block_pivot_ratio=%lu;\n", (unsigned long)s.pf_coarse_bpr);
    printf("predictor_family_fine_probe_block_pivot_ratio=%lu;\n", (unsigned long)s.pf_fine_bpr);
    printf("predictor_family_fine_eb_threshold=%.12g;\n", s.pf_fine_eb_threshold);
    printf("default_slope_invalid_ratio_threshold=%.3f;\n", p.default_slope_invalid_ratio_threshold);
    printf("default_slope_fallback=%.6f;\n", p.default_slope_fallback);
    printf("default_slope_eb_span=%.6f;\n", p.default_slope_eb_span);
    printf("t_precision=%d;\n", p.t_precision);
    printf("t_range=[%.6f,%.6f];\n", p.t_min, p.t_max);
    printf("t_candidates={");
    int64_t lo = lround_f32(p.t_min * (float)p.t_precision), hi = lround_f32(p.t_max * (float)p.t_precision);
    for (int64_t i = lo; i <= hi; i++) printf("%s%.6f", i == lo ? "" : ",", (float)i / (float)p.t_precision);
    printf("};\n");
    printf("block_t_candidates={0.062500,0.250000,0.281250,0.312500,0.343750};\n");
    printf("block_pivot_ratio=%lu;\n", (unsigned long)p.block_pivot_ratio);
    printf("insufficient_sample_t=%.6f;\n", 0.25);
    printf("predictor_family=%s;\n", p.quintic ? "quintic" : "cubic");
    printf("pivot_stride={%lu,%lu,%lu};\n", (unsigned long)ps[0], (unsigned long)ps[1], (unsigned long)ps[2]);
    printf("block_shape={%lu,%lu,%lu};\n", (unsigned long)bs[0], (unsigned long)bs[1], (unsigned long)bs[2]);
    printf("block_num={%lu,%lu,%lu};\n", (unsigned long)bn[0], (unsigned long)bn[1], (unsigned long)bn[2]);
}

inline std::vector<uint8_t> read_file(const std::string& path, const char* what) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error(std::string("cannot open ") + what + ": " + path);
    f.seekg(0, std::ios::end);
    uint64_t n = (uint64_t)f.tellg();
    f.seekg(0, std::ios::beg);
    std::vector<uint8_t> v(n);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)n);
    if (!f) throw std::runtime_error(std::string("cannot read complete ") + what + ": " + path);
    return v;
}

inline void check(asz_ctx* c, int rc) {
    if (rc == ASZ_OK) return;
    std::string msg = asz_last_error(c);
    if (rc == ASZ_ERR_ARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}
