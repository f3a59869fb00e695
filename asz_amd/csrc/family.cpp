// Predictor-family selection on the GPU: param_c::select_predictor_family
// (predictor_family.tpp:217-251) and predictor_family_detail::evaluate
// (predictor_family.tpp:129-214).
//
// The probe plan is the last predictor_family_suffix_layers rounds of the
// plan at a probe block pivot ratio; up to predictor_family_sample_blocks
// blocks are drawn by stratified mt19937_64(0) sampling and kept while they
// fit the point budget.  Each block is compressed in its own context twice
// (cubic and quintic interior) by the probe.hip kernels, with per-block t
// chosen on the host from GPU-sampled keys in the reference's map order.
// The host then forms the reference's long double statistics in its order:
// squared prediction / reconstruction errors summed per target block and
// round, per-round QB entropies from unordered_map<int32_t, uint64_t> in
// insertion order (p = count * (1 / n), log2 by the x87 fyl2x the reference
// inlines), PSNR by fldlg2 / fyl2x, and score_gain = psnr_gain +
// default_slope (cubic_h1 - quintic_h1).  The choice and every printed value
// are the reference's.
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <unordered_map>

#include "context.hpp"
#include "launch_family.hpp"
#include "scoring.hpp"

namespace asz {

namespace {

using Clock = std::chrono::steady_clock;

uint64_t al256(uint64_t v) { return (v + 255) / 256 * 256; }

// x87 log2 and log10 as the reference's inlined __builtin_log2l /
// __builtin_log10l (fld1 or fldlg2, then fyl2x).
inline long double x87_log2(long double x) {
    long double r;
    asm("fld1\n\tfxch %%st(1)\n\tfyl2x" : "=t"(r) : "0"(x));
    return r;
}
inline long double x87_log10(long double x) {
    long double r;
    asm("fldlg2\n\tfxch %%st(1)\n\tfyl2x" : "=t"(r) : "0"(x));
    return r;
}

// predictor_family.tpp:112-126
long double entropy(const std::vector<int32_t>& qbs) {
    if (qbs.empty()) return 0;
    std::unordered_map<int32_t, uint64_t> freq;
    for (int32_t q : qbs) freq[q]++;
    const long double inv = 1.0L / (long double)qbs.size();
    long double h1 = 0;
    for (const auto& kv : freq) {
        const long double p = (long double)kv.second * inv;
        h1 -= p * x87_log2(p);
    }
    return h1;
}

struct Stats {
    uint64_t points = 0;
    long double pred_mse = 0, rec_mse = 0, psnr = 0, h1 = 0;
};

}  // namespace

void select_family(asz_ctx* c, const float* d_field, const uint64_t dims[3], const asz_plan& plan,
                   const asz_params& p, float mn, float mx, const asz_family_opts& o, asz_family_report& rep) {
    const auto t0 = Clock::now();
    std::memset(&rep, 0, sizeof(rep));
    if (plan.layers <= 0) fail(ASZ_ERR_ARG, "cannot select predictor family without a predictor plan");
    if (o.suffix_layers == 0 || o.sample_blocks == 0 || o.sample_point_limit == 0 || o.coarse_probe_bpr == 0 ||
        o.fine_probe_bpr == 0 || !(o.sample_fraction_limit > 0 && o.sample_fraction_limit <= 1))
        fail(ASZ_ERR_ARG, "invalid predictor family selection settings");
    hipStream_t st = c->stream;
    // probe (predictor_family.tpp:222-229)
    const uint32_t S = (uint32_t)std::min<uint64_t>(o.suffix_layers, (uint64_t)plan.layers);
    asz_plan pp{};
    pp.layers = (int32_t)S;
    for (uint32_t r = 0; r < S; r++) {
        pp.axis[r] = plan.axis[plan.layers - S + r];
        pp.order[r] = plan.order[plan.layers - S + r];
    }
    const uint64_t bpr = p.rel_eb <= o.fine_eb_threshold ? o.fine_probe_bpr : o.coarse_probe_bpr;
    const Geometry g = make_geometry(dims, pp, bpr);
    rep.suffix_layers = S;
    rep.probe_block_pivot_ratio = bpr;
    const uint64_t N = g.N();
    // stratified targets (predictor_family.tpp:61-80) and the point budget (:97-110, 230-234)
    std::vector<std::array<uint64_t, 3>> cand;
    {
        const uint64_t total = g.nblocks();
        const uint64_t count = std::min<uint64_t>(o.sample_blocks, total);
        std::mt19937_64 rng(0);
        for (uint64_t s = 0; s < count; s++) {
            const uint64_t b0 = (uint64_t)((unsigned __int128)s * total / count);
            const uint64_t b1 = (uint64_t)((unsigned __int128)(s + 1) * total / count);
            uint64_t idx = b0 + rng() % (b1 - b0);
            std::array<uint64_t, 3> t;
            t[2] = idx % g.block_num[2];
            idx /= g.block_num[2];
            t[1] = idx % g.block_num[1];
            t[0] = idx / g.block_num[1];
            cand.push_back(t);
        }
    }
    const uint64_t fraction_budget = (uint64_t)floorl((long double)N * (long double)o.sample_fraction_limit);
    const uint64_t budget = std::min<uint64_t>(o.sample_point_limit, fraction_budget);
    rep.sample_point_budget = budget;
    std::vector<std::array<uint64_t, 3>> targets;
    uint64_t sampled = 0;
    for (const auto& t : cand) {
        uint64_t pts = 1;
        for (int a = 0; a < 3; a++) {
            const uint64_t b = t[a] * g.block_shape[a];
            pts *= std::min((t[a] + 1) * g.block_shape[a] + 1, g.D[a]) - b;
        }
        if (sampled <= budget && pts <= budget - sampled) {
            targets.push_back(t);
            sampled += pts;
        }
    }
    rep.blocks = (uint32_t)targets.size();
    if (targets.empty()) {
        rep.sampled_block_points = 0;
        rep.quintic = 0;
        rep.time_s = std::chrono::duration<double>(Clock::now() - t0).count();
        return;
    }
    rep.sampled_block_points = sampled;

    // jobs: family-major (cubic then quintic), target order within
    const uint32_t T = (uint32_t)targets.size(), J = 2 * T;
    const float aeb = abs_eb_of(p.rel_eb, mn, mx);
    const float seed_eb = daz(aeb * p.pivot_eb_ratio);
    struct Job {
        uint64_t cbb[3], tb[3], halo[3], LD[3];
    };
    std::vector<Job> jobs(J);
    std::vector<ProbeBox> boxes(J);
    uint64_t total_box = 0, max_box = 0;
    for (uint32_t j = 0; j < J; j++) {
        const auto& t = targets[j % T];
        Job& jb = jobs[j];
        ProbeBox& B = boxes[j];
        std::memset(&B, 0, sizeof(B));
        uint64_t n = 1;
        for (int a = 0; a < 3; a++) {
            jb.tb[a] = t[a];
            jb.cbb[a] = t[a] ? t[a] - 1 : 0;
            const uint64_t cbeg = jb.cbb[a] * g.block_shape[a];
            const uint64_t cend = std::min((t[a] + 1) * g.block_shape[a] + 1, g.D[a]);
            const uint64_t left = 5 * (g.pivot_stride[a] / 2);
            jb.halo[a] = cbeg > left ? cbeg - left : 0;
            jb.LD[a] = cend - jb.halo[a];
            B.D[a] = (uint32_t)g.D[a];
            B.LD[a] = (uint32_t)jb.LD[a];
            B.halo[a] = (uint32_t)jb.halo[a];
            B.cb[a] = (uint32_t)cbeg;
            B.ps[a] = (uint32_t)g.pivot_stride[a];
            n *= jb.LD[a];
        }
        B.base = total_box;
        B.n = n;
        B.inv_half_pe = 0.5 / (double)seed_eb;
        B.two_pe = (double)seed_eb * 2.0;
        total_box += (n + 63) / 64 * 64;
        max_box = std::max(max_box, n);
    }
    float* orig = c->pg_err.get<float>(2 * total_box + 64);
    float* recon = orig + total_box;
    uint8_t* dmeta = c->pg_meta.get<uint8_t>(J * sizeof(ProbeBox) + 64);
    HIPCHECK(hipMemcpyAsync(dmeta, boxes.data(), J * sizeof(ProbeBox), hipMemcpyHostToDevice, st));
    launch_probe_init(reinterpret_cast<const ProbeBox*>(dmeta), J, max_box, d_field, orig, recon, st);
    HIPCHECK(hipGetLastError());

    // output slots of the target blocks: QBs and records per (job, round)
    uint64_t stride0[3] = {g.pivot_stride[0], g.pivot_stride[1], g.pivot_stride[2]};
    std::vector<std::vector<uint64_t>> qoff(J, std::vector<uint64_t>(S)), qcnt(J, std::vector<uint64_t>(S));
    uint64_t qtotal = 0;
    {
        uint64_t s[3] = {stride0[0], stride0[1], stride0[2]};
        for (uint32_t r = 0; r < S; r++) {
            const int d = pp.axis[r];
            for (uint32_t j = 0; j < J; j++) {
                const Job& jb = jobs[j];
                uint64_t n = 1;
                for (int a = 0; a < 3; a++) {
                    const uint64_t lo = jb.tb[a] * g.block_shape[a] - jb.halo[a];
                    const uint64_t hi = std::min((jb.tb[a] + 1) * g.block_shape[a] + 1, g.D[a]) - jb.halo[a];
                    const uint64_t start = a == d ? lo + s[a] / 2 : (lo == 0 ? 0 : lo + s[a]);
                    n *= start >= hi ? 0 : (hi - start + s[a] - 1) / s[a];
                }
                qoff[j][r] = qtotal;
                qcnt[j][r] = n;
                qtotal += n;
            }
            s[d] /= 2;
        }
    }
    int32_t* dqb = c->pg_qb.get<int32_t>(qtotal + 64);
    float* drec = c->scratch.get<float>(3 * qtotal + 64);

    // every round's block descriptors up front (they depend on geometry
    // only); per round the keys are sampled, the t chosen on the host workers
    // and written into that round's own descriptor slice, so a round's
    // upload never overwrites descriptors a launched kernel still reads
    struct RoundSet {
        std::vector<ProbeKeys> keys;
        std::vector<ProbeUnit> units;
        std::vector<uint64_t> nkeys;  // samples per unit (0: t = 1/4 without sampling)
        uint64_t key_total = 0;
        float layer_eb, slope, stat_eb;
    };
    std::vector<RoundSet> rs(S);
    uint64_t s[3] = {stride0[0], stride0[1], stride0[2]};
    for (uint32_t r = 0; r < S; r++) {
        RoundSet& R = rs[r];
        const int d = pp.axis[r];
        const int L1 = (int)S - (int)r - 1;
        R.layer_eb = (float)((double)aeb * std::pow((double)p.eb_ratio, (double)L1));
        R.slope = (float)((double)p.default_slope * std::pow((double)p.slope_ratio, (double)L1));
        R.stat_eb = daz(R.layer_eb * p.stat_eb_ratio);
        for (uint32_t j = 0; j < J; j++) {
            const Job& jb = jobs[j];
            for (uint64_t b0 = jb.cbb[0]; b0 <= jb.tb[0]; b0++)
                for (uint64_t b1 = jb.cbb[1]; b1 <= jb.tb[1]; b1++)
                    for (uint64_t b2 = jb.cbb[2]; b2 <= jb.tb[2]; b2++) {
                        const uint64_t bb[3] = {b0, b1, b2};
                        uint64_t lo[3], hi[3];
                        for (int a = 0; a < 3; a++) {
                            lo[a] = bb[a] * g.block_shape[a] - jb.halo[a];
                            hi[a] = std::min((bb[a] + 1) * g.block_shape[a] + 1, g.D[a]) - jb.halo[a];
                        }
                        uint64_t dist = s[d] / 2;
                        for (int a = d + 1; a < 3; a++) dist *= jb.LD[a];
                        // t-selection samples (compress.tpp:96-125): empty means t = 1/4
                        uint64_t ns = 0;
                        if (!(hi[d] <= lo[d] + 3 * s[d])) {
                            uint64_t slo[3], shi[3];
                            bool empty = false;
                            for (int a = 0; a < 3; a++) {
                                slo[a] = lo[a];
                                shi[a] = hi[a];
                                if (a != d && lo[a] != 0) slo[a] += s[a];
                            }
                            slo[d] += 3 * s[d] / 2;
                            shi[d] -= 3 * s[d] / 2;
                            for (int a = 0; a < 3; a++) empty |= slo[a] >= shi[a];
                            if (!empty) {
                                const uint64_t vol = (shi[0] - slo[0]) * (shi[1] - slo[1]) * (shi[2] - slo[2]);
                                const uint64_t den = s[2] * (s[0] * s[1] * 256);  // local h1_block_sampling_num
                                uint64_t ss = (uint64_t)std::cbrt((double)((float)vol / (float)den));
                                if (ss == 0) ss = 1;
                                ProbeKeys K{};
                                K.base = boxes[j].base;
                                ns = 1;
                                for (int a = 0; a < 3; a++) {
                                    K.LD[a] = (uint32_t)jb.LD[a];
                                    K.lo[a] = (uint32_t)slo[a];
                                    K.step[a] = (uint32_t)(s[a] * ss);
                                    K.cnt[a] = (uint32_t)((shi[a] - slo[a] + s[a] * ss - 1) / (s[a] * ss));
                                    ns *= K.cnt[a];
                                }
                                K.dist = dist;
                                K.key_off = R.key_total;
                                K.neg2 = 0.5f / (p.stat_eb_ratio * R.layer_eb);
                                R.key_total += ns;
                                R.keys.push_back(K);
                            }
                        }
                        R.nkeys.push_back(ns);
                        ProbeUnit U{};
                        U.base = boxes[j].base;
                        for (int a = 0; a < 3; a++) {
                            U.LD[a] = (uint32_t)jb.LD[a];
                            const uint64_t start = a == d ? lo[a] + s[a] / 2 : (lo[a] == 0 ? 0 : lo[a] + s[a]);
                            U.start[a] = (uint32_t)start;
                            U.stride[a] = (uint32_t)s[a];
                            U.n[a] = (uint32_t)(start >= hi[a] ? 0 : (hi[a] - start + s[a] - 1) / s[a]);
                        }
                        U.d = (uint32_t)d;
                        U.order = pp.order[r];
                        U.h = (uint32_t)(s[d] / 2);
                        U.right = (uint32_t)hi[d];
                        U.quintic = j >= T ? 1 : 0;
                        U.diag = (b0 == jb.tb[0] && b1 == jb.tb[1] && b2 == jb.tb[2]) ? 1 : 0;
                        U.dist = dist;
                        U.qb_off = U.rec_off = qoff[j][r];
                        U.eb_d = (double)R.layer_eb;
                        U.inv_half_eb = 0.5 / U.eb_d;
                        U.neg_two_eb = -(2.0 * U.eb_d);
                        U.t = 0.25f;
                        R.units.push_back(U);
                    }
        }
        s[d] /= 2;
    }
    // descriptor slices: keys descriptors then unit descriptors, per round
    std::vector<uint64_t> koff(S + 1, 0), uoff(S + 1, 0);
    uint64_t key_max = 0;
    for (uint32_t r = 0; r < S; r++) {
        koff[r + 1] = koff[r] + al256(rs[r].keys.size() * sizeof(ProbeKeys));
        key_max = std::max(key_max, rs[r].key_total);
    }
    uoff[0] = koff[S];
    for (uint32_t r = 0; r < S; r++) uoff[r + 1] = uoff[r] + al256(rs[r].units.size() * sizeof(ProbeUnit));
    uint8_t* hdesc = c->pin_huff.get<uint8_t>(uoff[S] + 256);
    uint8_t* ddesc = c->hmeta.get<uint8_t>(uoff[S] + 256);
    for (uint32_t r = 0; r < S; r++)
        std::memcpy(hdesc + koff[r], rs[r].keys.data(), rs[r].keys.size() * sizeof(ProbeKeys));
    HIPCHECK(hipMemcpyAsync(ddesc, hdesc, koff[S], hipMemcpyHostToDevice, st));
    int64_t* dkeys = c->keys.get<int64_t>(2 * key_max + 2);
    int64_t* hkeys = c->pin_xfer.get<int64_t>(2 * key_max + 2);
    for (uint32_t r = 0; r < S; r++) {
        RoundSet& R = rs[r];
        // keys -> host t (compress.tpp:117-143 with the reference's map order)
        if (R.key_total) {
            launch_probe_keys(reinterpret_cast<const ProbeKeys*>(ddesc + koff[r]), (uint32_t)R.keys.size(), recon,
                              dkeys, st);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(hkeys, dkeys, 2 * R.key_total * sizeof(int64_t), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            std::vector<uint64_t> off(R.units.size() + 1, 0);
            for (size_t u = 0; u < R.units.size(); u++) off[u + 1] = off[u] + R.nkeys[u];
            host_pool(c).run(R.units.size(), [&](size_t u) {
                if (R.nkeys[u])
                    R.units[u].t = select_t_exact(hkeys + 2 * off[u], R.nkeys[u], R.layer_eb, R.stat_eb, R.slope, mn,
                                                  mx);
            });
        }
        std::memcpy(hdesc + uoff[r], R.units.data(), R.units.size() * sizeof(ProbeUnit));
        HIPCHECK(hipMemcpyAsync(ddesc + uoff[r], hdesc + uoff[r], R.units.size() * sizeof(ProbeUnit),
                                hipMemcpyHostToDevice, st));
        launch_probe_round(reinterpret_cast<const ProbeUnit*>(ddesc + uoff[r]), (uint32_t)R.units.size(), orig, recon,
                           dqb, drec, st);
        HIPCHECK(hipGetLastError());
    }
    std::vector<int32_t> hq(qtotal);
    std::vector<float> hr(3 * qtotal);
    HIPCHECK(hipMemcpyAsync(hq.data(), dqb, qtotal * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(hr.data(), drec, 3 * qtotal * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));

    // evaluate's statistics (predictor_family.tpp:185-213)
    auto evaluate = [&](uint32_t fam) {
        std::vector<long double> pse(S, 0.0L), rse(S, 0.0L);
        std::vector<uint64_t> cnt(S, 0);
        std::vector<std::vector<int32_t>> lq(S);
        for (uint32_t k = 0; k < T; k++) {
            const uint32_t j = fam * T + k;
            for (uint32_t r = 0; r < S; r++) {
                long double ps_ = 0.0L, rs_ = 0.0L;
                const float* e = hr.data() + 3 * qoff[j][r];
                for (uint64_t i = 0; i < qcnt[j][r]; i++) {
                    const long double pe = (long double)e[3 * i] - (long double)e[3 * i + 2];
                    const long double re = (long double)e[3 * i + 1] - (long double)e[3 * i + 2];
                    ps_ += pe * pe;
                    rs_ += re * re;
                }
                pse[r] += ps_;
                rse[r] += rs_;
                cnt[r] += qcnt[j][r];
                lq[r].insert(lq[r].end(), hq.begin() + qoff[j][r], hq.begin() + qoff[j][r] + qcnt[j][r]);
            }
        }
        Stats R;
        long double wh1 = 0, pt = 0, rt = 0;
        for (uint32_t r = 0; r < S; r++) {
            R.points += cnt[r];
            pt += pse[r];
            rt += rse[r];
            wh1 += entropy(lq[r]) * (long double)cnt[r];
        }
        if (R.points == 0) fail(ASZ_ERR_DATA, "predictor-family sample contains no target points");
        const long double inv = 1.0L / (long double)R.points;
        R.pred_mse = pt * inv;
        R.rec_mse = rt * inv;
        const long double range = (long double)mx - (long double)mn;
        R.psnr = R.rec_mse == 0 ? (long double)INFINITY : x87_log10(range * range / R.rec_mse) * 10.0L;
        R.h1 = wh1 * inv;
        return R;
    };
    const Stats cu = evaluate(0), qu = evaluate(1);
    // predictor_family.tpp:244-246 as compiled: isinf() folds away under
    // -ffinite-math-only, and the slope term is added as slope * (h1_c - h1_q)
    const long double psnr_gain = qu.psnr - cu.psnr;
    const long double score = (cu.h1 - qu.h1) * (long double)p.default_slope + psnr_gain;
    rep.points = cu.points;
    rep.cubic_mse = cu.pred_mse;
    rep.quintic_mse = qu.pred_mse;
    rep.cubic_psnr = cu.psnr;
    rep.quintic_psnr = qu.psnr;
    rep.cubic_h1 = cu.h1;
    rep.quintic_h1 = qu.h1;
    rep.score_gain = score;
    rep.quintic = score > 0 ? 1 : 0;
    rep.time_s = std::chrono::duration<double>(Clock::now() - t0).count();
}

}  // namespace asz

extern "C" int asz_select_family(asz_ctx* c, const float* d_field, const uint64_t dims[3], const asz_plan* plan,
                                 const asz_params* p, float min_val, float max_val, const asz_family_opts* opts,
                                 asz_family_report* report) {
    return guarded(c, [&] {
        if (!c || !d_field || !dims || !plan || !p || !opts || !report) asz::fail(ASZ_ERR_ARG, "null argument");
        HIPCHECK(hipSetDevice(c->device));
        asz::select_family(c, d_field, dims, *plan, *p, min_val, max_val, *opts, *report);
    });
}
