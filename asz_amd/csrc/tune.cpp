// Autotuner of the round kernels' tile shapes, and its cache file.
//
// For every large round that is not a block round, the tuner compresses a
// copy of the field under each candidate tile shape of that round (the
// others as chosen so far), times the round's kernel with an event pair
// (median of `reps` compressions) and keeps the fastest.  The tile shape
// changes how targets map to lanes and workgroups, never which value or QB
// lands where, so every choice gives the same stream.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "context.hpp"

namespace asz {
namespace {

TuneEntry& tune_slot(asz_ctx* c, const uint64_t dims[3], const asz_plan& plan, uint64_t bpr, uint64_t h1) {
    if (const TuneEntry* e = find_tune(c, dims, plan, bpr, h1)) return const_cast<TuneEntry&>(*e);
    TuneEntry e;
    std::memset(&e, 0, sizeof(e));
    for (int a = 0; a < 3; a++) e.dims[a] = dims[a];
    e.plan = canonical_plan(plan);
    e.bpr = bpr;
    e.h1 = h1;
    c->tunes.push_back(e);
    return c->tunes.back();
}

double median(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    return v[v.size() / 2];
}

}  // namespace
}  // namespace asz

using namespace asz;

extern "C" {

int asz_autotune(asz_ctx* c, const float* d_field, const uint64_t dims[3], const asz_plan* plan,
                 const asz_params* p, uint32_t reps, asz_tune_report* report) {
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        const uint64_t N = dims[0] * dims[1] * dims[2];
        if (N == 0) fail(ASZ_ERR_ARG, "empty field");
        reps = std::max<uint32_t>(1, std::min<uint32_t>(reps, 15));
        float mn = 0, mx = 0;
        if (asz_range_f32(c, d_field, N, &mn, &mx) != ASZ_OK) fail(ASZ_ERR_DATA, c->err);
        float* work = c->tune_work.get<float>(N);
        // the rounds and their kinds under the context's block-round setting
        const Geometry g = make_geometry(dims, *plan, p->block_pivot_ratio);
        const uint32_t coarse = c->coarse_max < 0 ? (uint32_t)ASZ_MAX_LAYERS : (uint32_t)c->coarse_max;
        const Plan P0 = make_plan(g, 1.0f, p->pivot_eb_ratio, p->eb_ratio, p->slope_ratio, 1.0f,
                                  p->h1_block_sampling_num, coarse);
        TuneEntry& te = tune_slot(c, dims, *plan, p->block_pivot_ratio, p->h1_block_sampling_num);
        std::memset(te.tune, 0, sizeof(te.tune));
        const uint64_t te_index = (uint64_t)(&te - c->tunes.data());
        asz_tune_report rep{};
        auto trial = [&](int r) {
            std::vector<double> ms;
            for (uint32_t k = 0; k < reps; k++) {
                HIPCHECK(hipMemcpyAsync(work, d_field, N * 4, hipMemcpyDeviceToDevice, c->stream));
                ms.push_back(compress_timed_round(c, work, dims, *plan, *p, mn, mx, r));
            }
            rep.trials++;
            return median(ms);
        };
        for (int r = 0; r < g.L; r++) {
            const RoundDesc& R = P0.rounds[r];
            if (R.kind == kRoundBlock || R.count * 64 < N) continue;
            std::vector<RoundTune> cands{RoundTune{0, 0, 0}};
            if (R.kind == kRoundSweep) {
                for (uint8_t seg : {24, 16, 12, 8})
                    for (uint8_t xsh : {6, 7, 8})
                        if (!(seg == kSweepSeg && xsh == R.xsh)) cands.push_back(RoundTune{seg, xsh, 0});
            } else {
                for (uint16_t rpw : {16, 32, 64, 128, 256, 512})
                    if (rpw != R.rpw) cands.push_back(RoundTune{0, 0, rpw});
            }
            double best = 0, base = 0;
            RoundTune pick{0, 0, 0};
            for (size_t i = 0; i < cands.size(); i++) {
                c->tunes[te_index].tune[r] = cands[i];
                const double t = trial(r);
                if (i == 0) base = best = t;
                else if (t < best) {
                    best = t;
                    pick = cands[i];
                }
            }
            c->tunes[te_index].tune[r] = pick;
            rep.rounds_tuned++;
            rep.default_ms += base;
            rep.tuned_ms += best;
        }
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (report) *report = rep;
    });
}

int asz_tune_lookup(asz_ctx* c, const uint64_t dims[3], const asz_plan* plan, uint64_t bpr, uint64_t h1,
                    asz_round_tune tune[ASZ_MAX_LAYERS]) {
    const TuneEntry* e = find_tune(c, dims, *plan, bpr, h1);
    if (!e) return 0;
    for (int r = 0; r < ASZ_MAX_LAYERS; r++) tune[r] = asz_round_tune{e->tune[r].seg, e->tune[r].xsh, e->tune[r].rpw};
    return 1;
}

// File: a header line "asz-tune 1", then one line per key:
//   d0 d1 d2 bpr h1 layers <axis order digits per round> <seg:xsh:rpw per round, comma separated>
int asz_tune_save(asz_ctx* c, const char* path) {
    return guarded(c, [&] {
        std::ofstream out(path, std::ios::trunc);
        if (!out) fail(ASZ_ERR_ARG, std::string("cannot write ") + path);
        out << "asz-tune 1\n";
        for (const TuneEntry& e : c->tunes) {
            out << e.dims[0] << ' ' << e.dims[1] << ' ' << e.dims[2] << ' ' << e.bpr << ' ' << e.h1 << ' '
                << e.plan.layers << ' ';
            for (int r = 0; r < e.plan.layers; r++) out << (int)e.plan.axis[r] << (int)e.plan.order[r];
            out << ' ';
            for (int r = 0; r < e.plan.layers; r++)
                out << (r ? "," : "") << (int)e.tune[r].seg << ':' << (int)e.tune[r].xsh << ':' << e.tune[r].rpw;
            out << '\n';
        }
        if (!out) fail(ASZ_ERR_ARG, std::string("cannot write ") + path);
    });
}

int asz_tune_load(asz_ctx* c, const char* path) {
    return guarded(c, [&] {
        std::ifstream in(path);
        if (!in) fail(ASZ_ERR_ARG, std::string("cannot read ") + path);
        std::string line;
        if (!std::getline(in, line) || line != "asz-tune 1") fail(ASZ_ERR_DATA, "not an asz-tune 1 file");
        std::vector<TuneEntry> got;
        while (std::getline(in, line)) {
            if (line.empty()) continue;
            std::istringstream ls(line);
            TuneEntry e;
            std::memset(&e, 0, sizeof(e));
            int layers = 0;
            std::string ao, tiles;
            if (!(ls >> e.dims[0] >> e.dims[1] >> e.dims[2] >> e.bpr >> e.h1 >> layers >> ao >> tiles) ||
                layers < 1 || layers > ASZ_MAX_LAYERS || ao.size() != 2u * (size_t)layers)
                fail(ASZ_ERR_DATA, "malformed tune line: " + line);
            e.plan.layers = layers;
            for (int r = 0; r < layers; r++) {
                const int ax = ao[2 * r] - '0', od = ao[2 * r + 1] - '0';
                if (ax < 0 || ax > 2 || od < 0 || od > 2) fail(ASZ_ERR_DATA, "malformed tune line: " + line);
                e.plan.axis[r] = (uint8_t)ax;
                e.plan.order[r] = (uint8_t)od;
            }
            std::istringstream ts(tiles);
            for (int r = 0; r < layers; r++) {
                unsigned seg = 0, xsh = 0, rpw = 0;
                char c1 = 0, c2 = 0, sep = 0;
                if (!(ts >> seg >> c1 >> xsh >> c2 >> rpw) || c1 != ':' || c2 != ':' || seg > kSweepSeg ||
                    (xsh && (xsh < 6 || xsh > 8)) || rpw > 1024)
                    fail(ASZ_ERR_DATA, "malformed tune line: " + line);
                if (r + 1 < layers && (!(ts >> sep) || sep != ','))
                    fail(ASZ_ERR_DATA, "malformed tune line: " + line);
                e.tune[r] = RoundTune{(uint8_t)seg, (uint8_t)xsh, (uint16_t)rpw};
            }
            got.push_back(e);
        }
        for (const TuneEntry& e : got) {
            TuneEntry& slot = tune_slot(c, e.dims, e.plan, e.bpr, e.h1);
            std::memcpy(slot.tune, e.tune, sizeof(slot.tune));
        }
    });
}

}  // extern "C"
