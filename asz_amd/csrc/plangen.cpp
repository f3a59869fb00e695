// Automatic plan generation on the GPU: param_c::generate_config
// (generate_config.tpp:443-678), the reverse greedy search that builds the
// round list from the finest round up.
//
// Per searched round (finest first), for each candidate predictor (all nine
// (axis, order) pairs while layer < full_predictor_search_layer_num, else the
// three "012" ones):
//   1. best t: the (A, B) key map of the candidate's axis over the original
//      field at the doubled pivot stride (get_AB_map.tpp:5-24; the keys are
//      sampled on the GPU by the default-slope kernel, the map and its H1-PSNR
//      scores run on the host in the reference's unordered_map order), scanned
//      over candidate_t from t = 1/4 (generate_config.tpp:498-548);
//   2. open-loop sample compression of the candidate round with that t
//      (k_plan_sample for all candidates in one launch, the float MSE chain
//      replayed by k_sq_chain);
//   3. the samples' Huffman stream (k_hist, host reference-tree codebooks,
//      the GPU chunk encoder) and its Zstd level-3 size on host threads
//      (generate_config.tpp:626-647);
//   4. the candidate with the largest psnr - slope * br is prepended to the
//      plan and its axis's pivot stride doubles (generate_config.tpp:649-671).
// The float arithmetic of every step follows GCC's lowering of the
// reference, so the plan and the printed best_dist values are the reference's.
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>

#include "context.hpp"
#include "launch_misc.hpp"
#include "launch_plangen.hpp"
#include "scoring.hpp"

extern "C" {
size_t ZSTD_compressBound(size_t srcSize);
size_t ZSTD_compress(void* dst, size_t dstCapacity, const void* src, size_t srcSize, int compressionLevel);
unsigned ZSTD_isError(size_t code);
}

namespace asz {

namespace {

using Clock = std::chrono::steady_clock;
double secs_since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

struct Candidate {
    int axis, order;  // order: 0 "120", 1 "021", 2 "012" (predictor[9] - '0')
};
// candidate_predictors (asz_param.hpp:147) and the reduced set (generate_config.tpp:492)
constexpr Candidate kAll[9] = {{0, 0}, {0, 1}, {0, 2}, {1, 0}, {1, 1}, {1, 2}, {2, 0}, {2, 1}, {2, 2}};
constexpr Candidate kReduced[3] = {{0, 2}, {1, 2}, {2, 2}};

// Best t of one axis (generate_config.tpp:527-548) from the (h1, psnr) of
// t = 1/4 (sc[0]) and of each candidate_t (sc[1 + k]): keep a candidate
// whose fmaf(-slope, h1, psnr) is larger.
float pick_t(const std::vector<std::pair<float, float>>& sc, float slope, const std::vector<float>& cand_t) {
    const float initial = 4.0f / 16;
    float best_t = initial;
    float best = std::fmaf(-slope, sc[0].first, sc[0].second);
    for (size_t k = 0; k < cand_t.size(); k++) {
        if (cand_t[k] == initial) continue;
        const float dist = std::fmaf(-slope, sc[1 + k].first, sc[1 + k].second);
        if (dist > best) {
            best = dist;
            best_t = cand_t[k];
        }
    }
    return best_t;
}

uint64_t al64(uint64_t v) { return (v + 63) / 64 * 64; }

}  // namespace

void generate_plan(asz_ctx* c, const float* d_field, const uint64_t dims[3], const asz_params& p, float mn, float mx,
                   uint32_t max_layers, uint32_t full_layers, uint32_t gcs, asz_plan& out, asz_plan_report& rep) {
    if (max_layers == 0 || max_layers > ASZ_MAX_LAYERS) fail(ASZ_ERR_ARG, "max_layer_num must be in [1,63]");
    if (full_layers == 0 || full_layers > ASZ_MAX_LAYERS)
        fail(ASZ_ERR_ARG, "full_predictor_search_layer_num must be in [1,63]");
    if (gcs == 0) fail(ASZ_ERR_ARG, "generate_config_sample_stride must be positive");
    if (p.h1_block_sampling_num == 0) fail(ASZ_ERR_ARG, "h1_block_sampling_num must be positive");
    for (int a = 0; a < 3; a++)
        if (dims[a] == 0 || dims[a] >= (1ull << 31)) fail(ASZ_ERR_ARG, "field dimensions out of range");
    std::memset(&rep, 0, sizeof(rep));
    const auto t_start = Clock::now();
    hipStream_t st = c->stream;
    const uint64_t D[3] = {dims[0], dims[1], dims[2]};
    const float aeb = abs_eb_of(p.rel_eb, mn, mx);
    const float range = daz(daz(mx) - daz(mn));
    const float r2 = daz(range * range);
    // candidate_t (asz_param.hpp:198-202)
    std::vector<float> cand_t;
    {
        const long tlo = std::lround(p.t_min * (float)p.t_precision), thi = std::lround(p.t_max * (float)p.t_precision);
        for (long ti = tlo; ti <= thi; ti++) cand_t.push_back((float)ti / (float)p.t_precision);
    }
    uint64_t ps[3] = {1, 1, 1};
    std::vector<Candidate> chosen;  // finest first
    double stat_s = 0, pred_s = 0, lossless_s = 0;

    for (uint32_t layer = 0; layer < max_layers; layer++) {
        // generate_config.tpp:483-487
        const float layer_eb = daz((float)((double)aeb * std::pow((double)p.eb_ratio, (double)layer)));
        const float neg2 = daz(0.5f / layer_eb);
        const float stat_eb = daz(layer_eb * p.stat_eb_ratio);
        const float neg2s = daz(0.5f / stat_eb);
        const float slope =
            (float)((double)p.default_slope * std::pow((double)p.slope_ratio, (double)layer));
        const Candidate* cands = layer < full_layers ? kAll : kReduced;
        const uint32_t ncand = layer < full_layers ? 9 : 3;

        // ---- 1. best t per axis
        auto ph = Clock::now();
        bool used[3] = {false, false, false};
        for (uint32_t k = 0; k < ncand; k++) used[cands[k].axis] = true;
        SlopeSampleDesc SD[3];
        bool has[3] = {false, false, false};
        uint64_t koff[4] = {0, 0, 0, 0};
        for (int d = 0; d < 3; d++) {
            koff[d + 1] = koff[d];
            if (!used[d] || !(D[d] > 3 * ps[d])) continue;  // generate_config.tpp:505-525
            uint64_t s2[3] = {ps[0], ps[1], ps[2]};
            s2[d] *= 2;
            uint64_t lo[3] = {0, 0, 0}, hi[3] = {D[0], D[1], D[2]};
            lo[d] = 3 * s2[d] / 2;
            hi[d] = D[d] - 3 * s2[d] / 2;
            if (hi[d] <= lo[d]) continue;  // the sampling loop runs no iteration
            const uint64_t vol = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
            const uint64_t den = s2[2] * (s2[0] * s2[1] * p.h1_block_sampling_num);
            uint64_t ss = (uint64_t)std::cbrt((double)((float)vol / (float)den));  // get_AB_map.tpp:9-12
            if (ss == 0) ss = 1;
            SlopeSampleDesc S{};
            for (int a = 0; a < 3; a++) {
                const uint64_t step = s2[a] * ss;
                S.D[a] = (uint32_t)D[a];
                S.lo[a] = (uint32_t)lo[a];
                S.hi[a] = (uint32_t)hi[a];
                S.step[a] = (uint32_t)step;
                S.cnt[a] = (uint32_t)((hi[a] - lo[a] + step - 1) / step);
            }
            uint64_t dist = s2[d] / 2;
            for (int a = d + 1; a < 3; a++) dist *= D[a];
            S.dist = dist;
            S.neg_2_stat_eb = neg2s;
            SD[d] = S;
            has[d] = true;
            koff[d + 1] = koff[d] + 2 * (uint64_t)S.cnt[0] * S.cnt[1] * S.cnt[2];
        }
        float best_t[3] = {0.25f, 0.25f, 0.25f};
        if (koff[3]) {
            int64_t* dk = c->keys.get<int64_t>(koff[3] + 2);
            for (int d = 0; d < 3; d++)
                if (has[d]) launch_slope_sample(SD[d], d_field, dk + koff[d], st);
            int64_t* hk = c->pin_xfer.get<int64_t>(koff[3] + 2);
            HIPCHECK(hipMemcpyAsync(hk, dk, koff[3] * sizeof(int64_t), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            // the (h1, psnr) scores of every (axis, t) on the host workers
            ABMap maps[3];
            FlatAB flats[3];
            for (int d = 0; d < 3; d++)
                if (has[d]) {
                    build_ab_map(hk + koff[d], (koff[d + 1] - koff[d]) / 2, maps[d]);
                    flatten_ab(maps[d], flats[d]);
                }
            const size_t nt = cand_t.size() + 1;
            std::vector<std::pair<float, float>> sc[3];
            std::vector<std::pair<int, size_t>> tasks;
            for (int d = 0; d < 3; d++) {
                if (!has[d] || maps[d].empty()) continue;
                sc[d].resize(nt);
                for (size_t k = 0; k < nt; k++)
                    if (k == 0 || cand_t[k - 1] != 0.25f) tasks.push_back({d, k});
            }
            host_pool(c).run(tasks.size(), [&](size_t i) {
                const int d = tasks[i].first;
                const size_t k = tasks[i].second;
                const float t = k == 0 ? 0.25f : cand_t[k - 1];
                h1_psnr(flats[d], t, layer_eb, stat_eb, mn, mx, sc[d][k].first, sc[d][k].second);
            });
            for (int d = 0; d < 3; d++)
                if (!sc[d].empty()) best_t[d] = pick_t(sc[d], slope, cand_t);
        }
        stat_s += secs_since(ph);

        // ---- 2. sample compression of every candidate
        ph = Clock::now();
        std::vector<PlanSampleDesc> S(ncand);
        uint64_t total = 0, max_count = 0;
        for (uint32_t k = 0; k < ncand; k++) {
            const int d = cands[k].axis, o = cands[k].order;
            const int full = o;  // "120": axis 0 unsampled, "021": axis 1, "012": axis 2
            uint64_t s2[3] = {ps[0], ps[1], ps[2]};
            s2[d] *= 2;
            PlanSampleDesc& Q = S[k];
            std::memset(&Q, 0, sizeof(Q));
            uint64_t cnt = 1;
            for (int a = 0; a < 3; a++) {
                const uint64_t step = s2[a] * (a == full ? 1 : gcs);
                const uint64_t start = a == d ? s2[d] / 2 : 0;
                const uint64_t n = start < D[a] ? (D[a] - start + step - 1) / step : 0;
                Q.D[a] = (uint32_t)D[a];
                Q.start[a] = (uint32_t)start;
                Q.step[a] = (uint32_t)std::min<uint64_t>(step, UINT32_MAX);
                Q.cnt[a] = (uint32_t)n;
                cnt *= n;
            }
            Q.d = (uint32_t)d;
            Q.order = (uint32_t)o;
            Q.h = (uint32_t)(s2[d] / 2);
            uint64_t dist = s2[d] / 2;
            for (int a = d + 1; a < 3; a++) dist *= D[a];
            Q.dist = dist;
            Q.off = total;
            Q.count = cnt;
            Q.t = best_t[d];
            Q.eb = layer_eb;
            Q.neg2 = neg2;
            total = al64(total + cnt);
            max_count = std::max(max_count, cnt);
        }
        const uint64_t o_desc = 0, o_mm = al64(ncand * sizeof(PlanSampleDesc)), o_mse = o_mm + 128,
                       meta_bytes = o_mse + 64;
        uint8_t* hm = c->pin_meta.get<uint8_t>(meta_bytes);
        std::memcpy(hm + o_desc, S.data(), ncand * sizeof(PlanSampleDesc));
        int32_t* mm = reinterpret_cast<int32_t*>(hm + o_mm);
        for (uint32_t k = 0; k < ncand; k++) { mm[2 * k] = INT32_MAX; mm[2 * k + 1] = INT32_MIN; }
        uint8_t* dm = c->pg_meta.get<uint8_t>(meta_bytes);
        const PlanSampleDesc* dS = reinterpret_cast<const PlanSampleDesc*>(dm + o_desc);
        int32_t* dq = c->pg_qb.get<int32_t>(total + 64);
        float* de = c->pg_err.get<float>(total + 64);
        HIPCHECK(hipMemcpyAsync(dm, hm, meta_bytes, hipMemcpyHostToDevice, st));
        launch_plan_sample(dS, ncand, max_count, d_field, dq, de, st);
        launch_plan_minmax(dS, ncand, max_count, dq, reinterpret_cast<int32_t*>(dm + o_mm), st);
        launch_sq_chain(dS, ncand, de, reinterpret_cast<float*>(dm + o_mse), st);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(hm + o_mm, dm + o_mm, meta_bytes - o_mm, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        int32_t qmm[18];
        float mse_sum[9];
        std::memcpy(qmm, hm + o_mm, 2 * ncand * sizeof(int32_t));
        std::memcpy(mse_sum, hm + o_mse, ncand * sizeof(float));
        pred_s += secs_since(ph);

        // ---- 3. Huffman stream of each candidate's samples and its Zstd size
        ph = Clock::now();
        std::vector<HuffArray> A(ncand);
        std::vector<HuffArrayDev> H(ncand);
        std::vector<uint32_t> carr;
        uint64_t hist_total = 0;
        for (uint32_t k = 0; k < ncand; k++) {
            std::memset(&H[k], 0, sizeof(HuffArrayDev));
            if (!S[k].count) continue;
            const int64_t lo = qmm[2 * k], hi = qmm[2 * k + 1];
            A[k].count = S[k].count;
            A[k].offset = lo;
            A[k].range = (uint64_t)(hi - lo + 1);
            if (A[k].range > (1ull << 24)) fail(ASZ_ERR_DATA, "QB alphabet wider than 2^24 symbols");
            H[k].src_off = S[k].off;
            H[k].count = S[k].count;
            H[k].offset = lo;
            H[k].range = A[k].range;
            H[k].hist_off = hist_total;
            H[k].hchunk0 = carr.size();
            hist_total += A[k].range;
            carr.insert(carr.end(), (S[k].count + kHistChunk - 1) / kHistChunk, k);
        }
        {
            const uint64_t o_arr = 0, o_carr = al64(ncand * sizeof(HuffArrayDev)),
                           up = o_carr + carr.size() * 4 + 64;
            uint8_t* hu = c->pin_huff.get<uint8_t>(up);
            std::memcpy(hu + o_arr, H.data(), ncand * sizeof(HuffArrayDev));
            std::memcpy(hu + o_carr, carr.data(), carr.size() * 4);
            uint8_t* du = c->hmeta.get<uint8_t>(up);
            uint32_t* dh = c->hist.get<uint32_t>(hist_total + 1);
            HIPCHECK(hipMemcpyAsync(du, hu, up, hipMemcpyHostToDevice, st));
            launch_fill_i32(reinterpret_cast<int32_t*>(dh), hist_total, 0, st);
            launch_hist(dq, reinterpret_cast<const HuffArrayDev*>(du + o_arr),
                        reinterpret_cast<const uint32_t*>(du + o_carr), carr.size(), dh, st);
            uint32_t* hh = c->pin_xfer.get<uint32_t>(hist_total + 1);
            HIPCHECK(hipMemcpyAsync(hh, dh, hist_total * 4, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            host_pool(c).run(ncand, [&](size_t k) {
                if (S[k].count) build_codebook(A[k], hh + H[k].hist_off);
            });
        }
        // every candidate's Huffman bytes (save + encode, huffman.hpp:322-381) back to back
        std::vector<uint64_t> boff(ncand + 1, 0);
        std::vector<uint8_t> blob;
        std::vector<ByteSeg> bsegs;
        uint64_t book_total = 0, chunks = 0;
        int max_len = 0;
        for (uint32_t k = 0; k < ncand; k++) {
            boff[k + 1] = boff[k];
            if (!S[k].count) continue;
            bsegs.push_back(ByteSeg{boff[k], blob.size(), A[k].header.size()});
            blob.insert(blob.end(), A[k].header.begin(), A[k].header.end());
            boff[k + 1] = boff[k] + A[k].header.size();
            if (A[k].single) continue;
            H[k].book_off = book_total;
            H[k].chunk0 = chunks;
            H[k].nchunk = (A[k].count + kEncChunk - 1) / kEncChunk;
            H[k].payload_bit = boff[k + 1] * 8;
            chunks += H[k].nchunk;
            boff[k + 1] += (A[k].payload_bits + 7) >> 3;
            book_total += A[k].range;
        }
        const uint64_t total_bytes = boff[ncand];
        std::vector<uint64_t> book(book_total + 1);
        for (uint32_t k = 0; k < ncand; k++) {
            if (!S[k].count || A[k].single) continue;
            int alen = 0;
            for (uint64_t s = 0; s < A[k].range; s++) {
                if (A[k].len[s] > 56) fail(ASZ_ERR_DATA, "Huffman code longer than 56 bits");
                book[H[k].book_off + s] = A[k].code[s] | ((uint64_t)A[k].len[s] << 56);
                alen = std::max<int>(alen, A[k].len[s]);
            }
            H[k].lds_book = A[k].range <= kLdsBook ? 1 : 0;
            max_len = std::max(max_len, alen);
        }
        uint8_t* hbytes = nullptr;
        if (total_bytes) {
            // codebooks as three parallel tables (64-bit entries, 32-bit codes, byte lengths)
            const uint64_t o_seg = 0, o_blob = al64(bsegs.size() * sizeof(ByteSeg)), o_book = al64(o_blob + blob.size()),
                           o_code = al64(o_book + book_total * 8), o_len = al64(o_code + book_total * 4),
                           o_arr = al64(o_len + book_total), up = o_arr + ncand * sizeof(HuffArrayDev);
            uint8_t* hu = c->pin_huff.get<uint8_t>(up);
            std::memcpy(hu + o_seg, bsegs.data(), bsegs.size() * sizeof(ByteSeg));
            std::memcpy(hu + o_blob, blob.data(), blob.size());
            std::memcpy(hu + o_book, book.data(), book_total * 8);
            for (uint64_t e = 0; e < book_total; e++) {
                reinterpret_cast<uint32_t*>(hu + o_code)[e] = (uint32_t)book[e];
                hu[o_len + e] = (uint8_t)(book[e] >> 56);
            }
            std::memcpy(hu + o_arr, H.data(), ncand * sizeof(HuffArrayDev));
            uint8_t* du = c->hmeta.get<uint8_t>(up);
            HIPCHECK(hipMemcpyAsync(du, hu, up, hipMemcpyHostToDevice, st));
            // payload words written (or zeroed and ORed) by the encoder, then the header bytes
            const uint64_t nwords = (total_bytes + 3) / 4 + 4;
            uint32_t* dwords = c->words.get<uint32_t>(nwords);
            if (chunks) {
                const HuffArrayDev* darr = reinterpret_cast<const HuffArrayDev*>(du + o_arr);
                const int enc_len = c->wide_codes ? std::max(max_len, 33) : max_len;  // > 32: the wide-code path
                uint32_t* estage = enc_len <= 32 ? c->enc_stage.get<uint32_t>(chunks * kEncChunk) : nullptr;
                launch_encode(dq, darr, ncand, chunks, reinterpret_cast<const uint64_t*>(du + o_book),
                              reinterpret_cast<const uint32_t*>(du + o_code), du + o_len,
                              c->chunk_bits.get<uint32_t>(chunks + 1), c->chunk_off.get<uint64_t>(chunks + 1), dwords,
                              enc_len, estage, st);
            }
            launch_scatter_bytes(reinterpret_cast<const ByteSeg*>(du + o_seg), (uint32_t)bsegs.size(), du + o_blob,
                                 reinterpret_cast<uint8_t*>(dwords), st);
            HIPCHECK(hipGetLastError());
            hbytes = c->pin_stream.get<uint8_t>(total_bytes + 1);
            HIPCHECK(hipMemcpyAsync(hbytes, dwords, total_bytes, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
        }
        // Zstd level 3 of each candidate's bytes (generate_config.tpp:639-641), one host thread each
        std::vector<uint64_t> zlen(ncand, 0);
        std::vector<std::string> zerr(ncand);
        auto zjob = [&](uint32_t k) {
            const uint64_t n = boff[k + 1] - boff[k];
            std::vector<uint8_t> z(ZSTD_compressBound(n));
            const size_t r = ZSTD_compress(z.data(), z.size(), hbytes + boff[k], n, 3);
            if (ZSTD_isError(r)) zerr[k] = "ZSTD_compress failed";
            else zlen[k] = r;
        };
        host_pool(c).run(ncand, [&](size_t k) {
            if (S[k].count) zjob((uint32_t)k);
        });
        for (auto& e : zerr)
            if (!e.empty()) fail(ASZ_ERR_DATA, e);
        lossless_s += secs_since(ph);

        // ---- 4. choose (generate_config.tpp:642-671)
        int best = -1;
        float best_dist = -1e8f;
        for (uint32_t k = 0; k < ncand; k++) {
            if (!S[k].count) continue;
            const float nf = (float)S[k].count;
            const float br = (float)zlen[k] * (8.0f / nf);
            const float mse = mse_sum[k] / nf;
            const float psnr = log10f(r2 / mse) * 10.0f;
            const float dist = psnr - slope * br;
            if (best < 0 || dist > best_dist) {
                best = (int)k;
                best_dist = dist;
            }
        }
        rep.best_dist[rep.searched] = best_dist;
        rep.best_axis[rep.searched] = best < 0 ? -1 : (int8_t)cands[best].axis;
        rep.best_order[rep.searched] = best < 0 ? -1 : (int8_t)cands[best].order;
        rep.searched++;
        if (best < 0) break;
        ps[cands[best].axis] *= 2;
        chosen.push_back(cands[best]);
    }
    std::memset(&out, 0, sizeof(out));
    out.layers = (int32_t)chosen.size();
    for (size_t r = 0; r < chosen.size(); r++) {
        out.axis[r] = (uint8_t)chosen[chosen.size() - 1 - r].axis;
        out.order[r] = (uint8_t)chosen[chosen.size() - 1 - r].order;
    }
    rep.layers = (uint32_t)chosen.size();
    rep.statistics_s = stat_s;
    rep.sample_prediction_s = pred_s;
    rep.lossless_estimation_s = lossless_s;
    rep.total_s = secs_since(t_start);
}

}  // namespace asz

extern "C" int asz_generate_plan(asz_ctx* c, const float* d_field, const uint64_t dims[3], const asz_params* p,
                                 float min_val, float max_val, uint32_t max_layer_num,
                                 uint32_t full_predictor_search_layer_num, uint32_t generate_config_sample_stride,
                                 asz_plan* plan, asz_plan_report* report) {
    return guarded(c, [&] {
        if (!c || !d_field || !dims || !p || !plan) asz::fail(ASZ_ERR_ARG, "null argument");
        HIPCHECK(hipSetDevice(c->device));
        asz_plan_report tmp;
        asz::generate_plan(c, d_field, dims, *p, min_val, max_val, max_layer_num, full_predictor_search_layer_num,
                           generate_config_sample_stride, *plan, report ? *report : tmp);
    });
}
