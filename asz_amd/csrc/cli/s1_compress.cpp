// Drop-in s1_compress (reference: c_codes/s1_compress.cpp).  Same command
// line, validation, stdout lines and .asz stream; the core runs on the GPU
// through libasz_gpu.so.  Additional flags (not in the reference):
//   --tsel exact|gpu|grad  per-block t from GPU samples scored on the host in
//                      the reference's order (default: the reference's t and
//                      stream), from the GPU scorer (faster; t may differ from
//                      the reference's where two candidates score within float
//                      rounding of each other), or from the gradient-tuned form
//   --zstd-level L     Zstd level of the .asz frame (default 3, the reference's)
//   --zstd-threads T   host threads for the Zstd frame (default 1; 0: automatic)
//   --device N         HIP device
#include <chrono>
#include <cstring>
#include <map>

#include "cli_common.hpp"

using namespace std;
using Clock = chrono::high_resolution_clock;

static double secs(Clock::time_point a, Clock::time_point b) { return chrono::duration<double>(b - a).count(); }

// s1_compress.cpp:49-99
static void validate(CliState& s) {
    asz_params& p = s.p;
    if (!(p.pivot_eb_ratio > 0 && p.pivot_eb_ratio <= 1)) throw invalid_argument("pivot_eb_ratio must be in (0,1]");
    if (!(p.eb_ratio > 0 && p.eb_ratio <= 1)) throw invalid_argument("eb_ratio must be in (0,1]");
    if (!(p.slope_ratio > 0 && p.slope_ratio <= 1)) throw invalid_argument("slope_ratio must be in (0,1]");
    if (p.block_pivot_ratio == 0) throw invalid_argument("block_pivot_ratio must be positive");
    if (s.max_layer_num == 0 || s.max_layer_num > 63) throw invalid_argument("max_layer_num must be in [1,63]");
    if (s.full_predictor_search_layer_num == 0 || s.full_predictor_search_layer_num > 63)
        throw invalid_argument("full_predictor_search_layer_num must be in [1,63]");
    if (s.generate_config_sample_stride == 0) throw invalid_argument("generate_config_sample_stride must be positive");
    if (p.h1_block_sampling_num == 0) throw invalid_argument("h1_block_sampling_num must be positive");
    if (!(p.stat_eb_ratio > 0 && p.stat_eb_ratio <= 1)) throw invalid_argument("stat_eb_ratio must be in (0,1]");
    if (s.pf_suffix_layers == 0) throw invalid_argument("predictor_family_suffix_layers must be positive");
    if (s.pf_sample_blocks == 0) throw invalid_argument("predictor_family_sample_blocks must be positive");
    if (s.pf_sample_point_limit == 0) throw invalid_argument("predictor_family_sample_point_limit must be positive");
    if (!(s.pf_sample_fraction_limit > 0 && s.pf_sample_fraction_limit <= 1))
        throw invalid_argument("predictor_family_sample_fraction_limit must be in (0,1]");
    if (s.pf_coarse_bpr == 0 || s.pf_fine_bpr == 0)
        throw invalid_argument("predictor family probe block pivot ratios must be positive");
    if (!(s.pf_fine_eb_threshold > 0 && isfinite(s.pf_fine_eb_threshold)))
        throw invalid_argument("predictor_family_fine_eb_threshold must be positive and finite");
    if (!(p.default_slope_invalid_ratio_threshold > 0 && p.default_slope_invalid_ratio_threshold <= 1))
        throw invalid_argument("default_slope_invalid_ratio_threshold must be in (0,1]");
    if (!(p.default_slope_fallback > 0 && isfinite(p.default_slope_fallback)))
        throw invalid_argument("default_slope_fallback must be positive and finite");
    if (!(p.default_slope_eb_span > 1 && isfinite(p.default_slope_eb_span)))
        throw invalid_argument("default_slope_eb_span must be greater than 1 and finite");
    if (p.t_precision < 16 || p.t_precision % 16 != 0) throw invalid_argument("t_precision must be a positive multiple of 16");
    if (s.t_min_set != s.t_max_set) throw invalid_argument("t_min and t_max must be supplied together");
    if (!(p.t_min >= 0 && p.t_min <= p.t_max && p.t_max <= 0.5))
        throw invalid_argument("t range must satisfy 0 <= t_min <= t_max <= 0.5");
    if (fabs(p.t_min * p.t_precision - round(p.t_min * p.t_precision)) > 1e-6 ||
        fabs(p.t_max * p.t_precision - round(p.t_max * p.t_precision)) > 1e-6)
        throw invalid_argument("t_min and t_max must be representable at t_precision");
    if (p.t_min > 4.0f / 16 || p.t_max < 5.0f / 16) throw invalid_argument("t range must include 4/16 and 5/16");
    // asz_param.hpp:194-197
    if (p.t_precision % 32 || p.t_min > 1.0f / 16 || p.t_max < 11.0f / 32)
        throw invalid_argument("block t candidates require t_precision divisible by 32 and t range covering [1/16,11/32]");
}

int main(int argc, char* argv[]) try {
    if (argc == 2 && string(argv[1]) == "--help") {
        printf("usage: s1_compress -i INPUT -z COMPRESSED -o RECONSTRUCTED -3 DIM0 DIM1 DIM2 -M REL REL_EB [options]\n");
        return 0;
    }
    CliState s;
    asz_params_default(&s.p);
    for (int i = 1; i < argc; i++) {
        string a = argv[i];
        auto next = [&]() -> string {
            if (i + 1 >= argc) throw invalid_argument("missing value for " + a);
            return argv[++i];
        };
        if (a == "-f") continue;
        else if (a == "-i") { s.in_path = next(); s.z_path = s.in_path + ".asz"; s.o_path = s.z_path + ".bin"; }
        else if (a == "-z") { s.z_path = next(); s.o_path = s.z_path + ".bin"; }
        else if (a == "-o") s.o_path = next();
        else if (a == "-3") { uint64_t d0 = stoul(next()), d1 = stoul(next()), d2 = stoul(next()); s.dims = {d0, d1, d2}; }
        else if (a == "-M") { i++; s.p.rel_eb = stof(next()); }
        else if (a == "--predictor-plan") s.plan = parse_predictor_plan(next());
        else if (a == "--predictor-family") {
            string f = next();
            s.auto_family = false;
            if (f == "cubic") s.p.quintic = 0;
            else if (f == "quintic") s.p.quintic = 1;
            else throw invalid_argument("predictor_family must be cubic or quintic");
        }
        else if (a == "--pivot-eb-ratio") s.p.pivot_eb_ratio = stof(next());
        else if (a == "--eb-ratio") s.p.eb_ratio = stof(next());
        else if (a == "--slope-ratio") s.p.slope_ratio = stof(next());
        else if (a == "--block-pivot-ratio") s.p.block_pivot_ratio = stoull(next());
        else if (a == "--max-layer-num") s.max_layer_num = stoull(next());
        else if (a == "--full-predictor-search-layer-num") s.full_predictor_search_layer_num = stoull(next());
        else if (a == "--generate-config-sample-stride") s.generate_config_sample_stride = stoull(next());
        else if (a == "--h1-block-sampling-num") s.p.h1_block_sampling_num = stoull(next());
        else if (a == "--stat-eb-ratio") s.p.stat_eb_ratio = stof(next());
        else if (a == "--predictor-family-suffix-layers") s.pf_suffix_layers = stoull(next());
        else if (a == "--predictor-family-sample-blocks") s.pf_sample_blocks = stoull(next());
        else if (a == "--predictor-family-sample-point-limit") s.pf_sample_point_limit = stoull(next());
        else if (a == "--predictor-family-sample-fraction-limit") s.pf_sample_fraction_limit = stof(next());
        else if (a == "--predictor-family-coarse-probe-block-pivot-ratio") s.pf_coarse_bpr = stoull(next());
        else if (a == "--predictor-family-fine-probe-block-pivot-ratio") s.pf_fine_bpr = stoull(next());
        else if (a == "--predictor-family-fine-eb-threshold") s.pf_fine_eb_threshold = stof(next());
        else if (a == "--default-slope-invalid-ratio-threshold") s.p.default_slope_invalid_ratio_threshold = stof(next());
        else if (a == "--default-slope-fallback") s.p.default_slope_fallback = stof(next());
        else if (a == "--default-slope-eb-span") s.p.default_slope_eb_span = stof(next());
        else if (a == "--t-precision") s.p.t_precision = stoi(next());
        else if (a == "--t-min") { s.p.t_min = stof(next()); s.t_min_set = true; }
        else if (a == "--t-max") { s.p.t_max = stof(next()); s.t_max_set = true; }
        else if (a == "--report-t-histogram") s.report_t_histogram = true;
        else if (a == "--tsel") {
            string m = next();
            if (m == "gpu") s.tsel = ASZ_TSEL_GPU;
            else if (m == "exact") s.tsel = ASZ_TSEL_EXACT;
            else if (m == "grad") s.tsel = ASZ_TSEL_GRAD;
            else throw invalid_argument("tsel must be gpu, exact or grad");
        }
        else if (a == "--device") s.device = stoi(next());
        else if (a == "--zstd-level") s.zstd_level = stoi(next());
        else if (a == "--zstd-threads") s.zstd_threads = (uint32_t)stoul(next());
        else throw invalid_argument("unknown argument: " + a);
    }
    validate(s);

    // load_input (s1_compress.cpp:101-140)
    if (s.in_path.empty() || s.z_path.empty()) throw invalid_argument("input and compressed-output paths are required");
    if (s.dims.size() != 3 || s.dims[0] < 7 || s.dims[1] < 7 || s.dims[2] < 7)
        throw invalid_argument("all three dimensions must be at least 7");
    if (!(s.p.rel_eb > 0 && isfinite(s.p.rel_eb))) throw invalid_argument("relative error bound must be positive and finite");
    const uint64_t dims[3] = {s.dims[0], s.dims[1], s.dims[2]};
    if (dims[0] > UINT64_MAX / dims[1] || dims[0] * dims[1] > UINT64_MAX / dims[2])
        throw invalid_argument("data dimensions overflow the element count");
    const uint64_t N = dims[0] * dims[1] * dims[2];
    vector<float> data(N);
    {
        ifstream in(s.in_path, ios::binary);
        if (!in) throw runtime_error("cannot open input: " + s.in_path);
        in.read(reinterpret_cast<char*>(data.data()), (streamsize)(N * 4));
        if (!in) throw runtime_error("cannot read complete input: " + s.in_path);
        if (in.peek() != char_traits<char>::eof()) throw runtime_error("input size does not match the dimensions: " + s.in_path);
    }
    asz_ctx* ctx = nullptr;
    if (asz_ctx_create(s.device, &ctx) != ASZ_OK) throw runtime_error("cannot initialise the HIP device");
    float* d_work = nullptr;
    if (asz_device_alloc(ctx, N * 4, (void**)&d_work) != ASZ_OK) throw runtime_error(asz_last_error(ctx));
    check(ctx, asz_copy_to_device(ctx, d_work, data.data(), N * 4));
    float mn, mx;
    auto r0 = Clock::now();
    int rc = asz_range_f32(ctx, d_work, N, &mn, &mx);
    auto r1 = Clock::now();
    if (rc != ASZ_OK) {
        string m = asz_last_error(ctx);
        if (m == "input contains a nonfinite value") throw runtime_error(m + ": " + s.in_path);
        throw runtime_error(m);
    }
    double data_range_time = secs(r0, r1);
    printf("Data range time: %f s\n", data_range_time);

    // prepare_compression_plan (s1_compress.cpp:30-47)
    float slope, inv;
    int fb, reason;
    auto d0 = Clock::now();
    check(ctx, asz_default_slope(ctx, d_work, dims, &s.p, mn, mx, &slope, &inv, &fb, &reason));
    auto d1 = Clock::now();
    static const char* reasons[3] = {"none", "invalid_ratio", "invalid_difference"};
    printf("Default slope statistics: max_invalid_ratio=%.6f, fallback=%s, reason=%s, slope=%.6f\n", inv,
           fb ? "true" : "false", reasons[reason], slope);
    printf("Get default slope time: %f s\n", secs(d0, d1));
    s.p.default_slope = slope;
    printf("default_slope=%f\n", slope);
    if (s.plan.empty()) {
        // generate_config (generate_config.tpp:443-678) on the GPU
        asz_plan gen{};
        asz_plan_report rep{};
        check(ctx, asz_generate_plan(ctx, d_work, dims, &s.p, mn, mx, (uint32_t)s.max_layer_num,
                                     (uint32_t)s.full_predictor_search_layer_num,
                                     (uint32_t)s.generate_config_sample_stride, &gen, &rep));
        static const char* ord[3] = {"120", "021", "012"};
        for (uint32_t k = 0; k < rep.searched; k++) {
            string name = rep.best_axis[k] < 0 ? string()
                                               : "cubic" + to_string(rep.best_axis[k]) + "_" + ord[rep.best_order[k]];
            printf("best_layer_predictor=%s, best_dist=%f\n", name.c_str(), rep.best_dist[k]);
        }
        printf("Generate config breakdown: statistics=%f s,sample prediction=%f s,lossless estimation=%f s\n",
               rep.statistics_s, rep.sample_prediction_s, rep.lossless_estimation_s);
        printf("Generate config time: %f s\n", rep.total_s);
        s.plan = from_plan(gen);
    }
    if (s.plan.size() > 63) throw invalid_argument("predictor plan cannot contain more than 63 layers");
    asz_plan plan = to_plan(s.plan);
    if (s.auto_family) {
        // select_predictor_family (predictor_family.tpp:217-251) on the GPU
        asz_family_opts fo{};
        fo.suffix_layers = s.pf_suffix_layers;
        fo.sample_blocks = s.pf_sample_blocks;
        fo.sample_point_limit = s.pf_sample_point_limit;
        fo.sample_fraction_limit = s.pf_sample_fraction_limit;
        fo.fine_eb_threshold = s.pf_fine_eb_threshold;
        fo.coarse_probe_bpr = s.pf_coarse_bpr;
        fo.fine_probe_bpr = s.pf_fine_bpr;
        asz_family_report fr{};
        check(ctx, asz_select_family(ctx, d_work, dims, &plan, &s.p, mn, mx, &fo, &fr));
        if (fr.blocks == 0)
            printf("Predictor family selection: selector=h1, suffix_layers=%lu, blocks=0, sampled_block_points=0, "
                   "sample_point_budget=%lu, probe_block_pivot_ratio=%lu, fallback=point_budget, family=cubic\n",
                   (unsigned long)fr.suffix_layers, (unsigned long)fr.sample_point_budget,
                   (unsigned long)fr.probe_block_pivot_ratio);
        else
            printf("Predictor family selection: selector=h1, suffix_layers=%lu, blocks=%lu, sampled_block_points=%lu, "
                   "sample_point_budget=%lu, probe_block_pivot_ratio=%lu, points=%lu, cubic_mse=%.12Le, "
                   "quintic_mse=%.12Le, cubic_psnr=%.12Le, quintic_psnr=%.12Le, cubic_h1=%.12Le, quintic_h1=%.12Le, "
                   "six_score_gain=%.12Le, family=%s\n",
                   (unsigned long)fr.suffix_layers, (unsigned long)fr.blocks, (unsigned long)fr.sampled_block_points,
                   (unsigned long)fr.sample_point_budget, (unsigned long)fr.probe_block_pivot_ratio,
                   (unsigned long)fr.points, fr.cubic_mse, fr.quintic_mse, fr.cubic_psnr, fr.quintic_psnr, fr.cubic_h1,
                   fr.quintic_h1, fr.score_gain, fr.quintic ? "quintic" : "cubic");
        printf("Predictor family selection time: %f s\n", fr.time_s);
        s.p.quintic = fr.quintic;
    }
    uint64_t ps[3], bs[3], bn[3];
    check(ctx, asz_geometry(dims, &plan, s.p.block_pivot_ratio, ps, bs, bn));
    print_config(s, s.plan, ps, bs, bn);

    // compress (compress.tpp:146-321)
    printf("min_val=%f max_val=%f\n", mn, mx);
    auto c0 = Clock::now();
    const uint8_t* hs = nullptr;
    uint64_t hn = 0;
    const uint64_t nt = bn[0] * bn[1] * bn[2] * (uint64_t)plan.layers;
    vector<int64_t> t_int(s.report_t_histogram ? nt : 0);
    check(ctx, asz_compress(ctx, d_work, dims, &plan, &s.p, mn, mx, s.tsel, nullptr, &hs, &hn, nullptr,
                            s.report_t_histogram ? t_int.data() : nullptr));
    auto c1 = Clock::now();
    asz_stats st;
    asz_get_stats(ctx, &st);
    uint64_t total_qb = 0;
    for (uint64_t k = 0; k <= st.layers; k++) total_qb += st.qb_counts[k];
    printf("total_qb_num=%lu\n", (unsigned long)total_qb);
    printf("t_int_vec.size()=%lu\n", (unsigned long)st.t_count);
    if (s.report_t_histogram) {
        map<int64_t, uint64_t> h;
        for (auto t : t_int) h[t]++;
        printf("T histogram:");
        bool first = true;
        for (auto& kv : h) { printf("%s%lld=%lu", first ? "" : ",", (long long)kv.first, (unsigned long)kv.second); first = false; }
        printf("\n");
        printf("T fallback count:%lu\n", (unsigned long)st.t_fallbacks);
    }
    uint8_t* zs = nullptr;
    uint64_t zn = 0;
    auto z0 = Clock::now();
    if (asz_zstd_compress_mt(hs, hn, s.zstd_level, s.zstd_threads, &zs, &zn, nullptr) != ASZ_OK)
        throw runtime_error("ZSTD_compress failed");
    auto z1 = Clock::now();
    double core = secs(c0, z1);
    printf("Huffman compression time: %f s\n", (st.huffman_ms + st.huffman_host_ms) / 1e3);
    printf("GPU kernel time: rounds=%f s, huffman=%f s, stream_bytes=%lu\n", st.rounds_ms / 1e3, st.huffman_ms / 1e3,
           (unsigned long)st.stream_bytes);
    printf("Zstd compression time: %f s\n", secs(z0, z1));
    printf("Core compression time: %f s\n", core);
    printf("Compression time: %f s\n", data_range_time + core);
    {
        ofstream out(s.z_path, ios::binary);
        if (!out) throw runtime_error("cannot open compressed output: " + s.z_path);
        out.write(reinterpret_cast<const char*>(zs), (streamsize)zn);
        if (!out) throw runtime_error("cannot write compressed output: " + s.z_path);
    }
    float abs_eb = s.p.rel_eb * (mx - mn);
    float cr = (float)N * 4 / (float)zn;
    float br = 32 / cr;
    double mse = st.sq_err_sum / (double)N;
    float psnr = (float)(10 * log10((double)(mx - mn) * (double)(mx - mn) / mse));
    printf("max_err= %f (%f%%)\n", st.max_err, st.max_err * 100 / abs_eb);
    printf("CR: %f, BR: %f, PSNR: %f\n", cr, br, psnr);
    printf("dist=%f\n", -slope * br + psnr);
    asz_free(zs);
    asz_device_free(ctx, d_work);
    asz_ctx_destroy(ctx);
    (void)c1;
    return 0;
} catch (const exception& e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
}
