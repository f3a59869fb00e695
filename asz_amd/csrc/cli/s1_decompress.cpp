// Drop-in s1_decompress (reference: c_codes/s1_decompress.cpp).  Same command
// line, validation and stdout lines; Huffman decoding and reconstruction run
// on the GPU through libasz_gpu.so, Zstd on the host.
#include <chrono>

#include "cli_common.hpp"

using namespace std;
using Clock = chrono::high_resolution_clock;

static double secs(Clock::time_point a, Clock::time_point b) { return chrono::duration<double>(b - a).count(); }

int main(int argc, char* argv[]) try {
    if (argc == 2 && string(argv[1]) == "--help") {
        printf("usage: s1_decompress -i INPUT -z COMPRESSED -o RECONSTRUCTED -3 DIM0 DIM1 DIM2 -M REL REL_EB [options]\n");
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
        else if (a == "-i") { s.in_path = next(); s.z_path = s.in_path + ".asz"; s.o_path = s.in_path + ".bin"; }
        else if (a == "-z") { s.z_path = next(); s.o_path = s.z_path + ".bin"; }
        else if (a == "-o") s.o_path = next();
        else if (a == "-3") { uint64_t d0 = stoul(next()), d1 = stoul(next()), d2 = stoul(next()); s.dims = {d0, d1, d2}; }
        else if (a == "-M") { i++; s.p.rel_eb = stof(next()); }
        else if (a == "--predictor-family") {
            string f = next();
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
        else if (a == "--t-precision") s.p.t_precision = stoi(next());
        else if (a == "--device") s.device = stoi(next());
        else throw invalid_argument("unknown argument: " + a);
    }
    // s1_decompress.cpp:97-125
    if (s.z_path.empty() || s.o_path.empty())
        throw invalid_argument("compressed-input and decompressed-output paths are required");
    if (s.dims.size() != 3 || s.dims[0] < 7 || s.dims[1] < 7 || s.dims[2] < 7)
        throw invalid_argument("all three dimensions must be at least 7");
    if (!(s.p.rel_eb > 0 && isfinite(s.p.rel_eb))) throw invalid_argument("relative error bound must be positive and finite");
    const uint64_t dims[3] = {s.dims[0], s.dims[1], s.dims[2]};
    if (dims[0] > UINT64_MAX / dims[1] || dims[0] * dims[1] > UINT64_MAX / dims[2])
        throw invalid_argument("data dimensions overflow the element count");
    if (!(s.p.pivot_eb_ratio > 0 && s.p.pivot_eb_ratio <= 1)) throw invalid_argument("pivot_eb_ratio must be in (0,1]");
    if (!(s.p.eb_ratio > 0 && s.p.eb_ratio <= 1)) throw invalid_argument("eb_ratio must be in (0,1]");
    if (!(s.p.slope_ratio > 0 && s.p.slope_ratio <= 1)) throw invalid_argument("slope_ratio must be in (0,1]");
    if (s.p.block_pivot_ratio == 0) throw invalid_argument("block_pivot_ratio must be positive");
    if (s.max_layer_num == 0 || s.max_layer_num > 63) throw invalid_argument("max_layer_num must be in [1,63]");
    if (s.full_predictor_search_layer_num == 0 || s.full_predictor_search_layer_num > 63)
        throw invalid_argument("full_predictor_search_layer_num must be in [1,63]");
    if (s.p.t_precision < 16 || s.p.t_precision % 16 != 0)
        throw invalid_argument("t_precision must be a positive multiple of 16");

    vector<uint8_t> z = read_file(s.z_path, "compressed input");
    if (z.empty()) throw runtime_error("compressed input is empty: " + s.z_path);
    const uint64_t N = dims[0] * dims[1] * dims[2];

    auto t0 = Clock::now();
    uint8_t* hs = nullptr;
    uint64_t hn = 0;
    if (asz_zstd_decompress(z.data(), z.size(), &hs, &hn) != ASZ_OK)
        throw runtime_error("ZSTD_decompress error in decompress()");
    auto t1 = Clock::now();
    asz_ctx* ctx = nullptr;
    if (asz_ctx_create(s.device, &ctx) != ASZ_OK) throw runtime_error("cannot initialise the HIP device");
    float* d_out = nullptr;
    if (asz_device_alloc(ctx, N * 4, (void**)&d_out) != ASZ_OK) throw runtime_error(asz_last_error(ctx));
    auto t2 = Clock::now();
    check(ctx, asz_decompress(ctx, hs, hn, dims, &s.p, d_out, nullptr, nullptr));
    vector<float> out(N);
    check(ctx, asz_copy_to_host(ctx, out.data(), d_out, N * 4));
    auto t3 = Clock::now();
    asz_stats st;
    asz_get_stats(ctx, &st);
    double huff = st.huffman_ms / 1e3, recon = st.rounds_ms / 1e3;
    double setup = secs(t2, t3) - huff - recon;
    printf("Zstd decompression time: %f s\n", secs(t0, t1));
    printf("Huffman decompression time: %f s\n", huff);
    printf("Reconstruction setup time: %f s\n", setup > 0 ? setup : 0.0);
    printf("Prediction reconstruction time: %f s\n", recon);
    printf("Decompression time: %f s\n", secs(t0, t1) + secs(t2, t3));
    float cr = (float)N * 4 / (float)z.size();
    float br = 32 / cr;
    printf("CR= %f\n", cr);
    printf("BR= %f\n", br);
    {
        ofstream f(s.o_path, ios::binary);
        if (!f) throw runtime_error("cannot open decompressed output: " + s.o_path);
        f.write(reinterpret_cast<const char*>(out.data()), (streamsize)(N * 4));
        if (!f) throw runtime_error("cannot write decompressed output: " + s.o_path);
    }
    asz_free(hs);
    asz_device_free(ctx, d_out);
    asz_ctx_destroy(ctx);
    return 0;
} catch (const exception& e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
}
