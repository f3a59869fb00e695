// Plan geometry and per-round descriptors (host).
#include <algorithm>
#include <cmath>

#include "internal.hpp"
#include "layout.hpp"

namespace asz {

Geometry make_geometry(const uint64_t dims[3], const asz_plan& plan, uint64_t bpr) {
    Geometry g;
    if (plan.layers < 1 || plan.layers > ASZ_MAX_LAYERS)
        fail(ASZ_ERR_ARG, "predictor plan cannot contain more than 63 layers");
    if (bpr == 0) fail(ASZ_ERR_ARG, "block_pivot_ratio must be positive");
    g.L = plan.layers;
    for (int a = 0; a < 3; a++) {
        if (dims[a] < 7) fail(ASZ_ERR_ARG, "all three dimensions must be at least 7");
        if (dims[a] > 0x7FFFFFFFull) fail(ASZ_ERR_ARG, "dimension too large");
        g.D[a] = dims[a];
        g.pivot_stride[a] = 1;
    }
    for (int r = 0; r < g.L; r++) {
        if (plan.axis[r] > 2 || plan.order[r] > 2) fail(ASZ_ERR_ARG, "unsupported predictor direction");
        g.axis[r] = plan.axis[r];
        g.order[r] = plan.order[r];
        if (g.pivot_stride[plan.axis[r]] >= (1ull << 31)) fail(ASZ_ERR_ARG, "too many rounds on one axis");
        g.pivot_stride[plan.axis[r]] *= 2;
    }
    for (int a = 0; a < 3; a++) {
        g.block_shape[a] = g.pivot_stride[a] * bpr;
        g.block_num[a] = (g.D[a] + g.block_shape[a] - 1) / g.block_shape[a];
    }
    return g;
}

void set_plan_bounds(Plan& P, float abs_eb, float pivot_eb_ratio, float eb_ratio, float slope_ratio,
                     float default_slope) {
    P.pivot_eb = daz(abs_eb) * daz(pivot_eb_ratio);  // compress.tpp:166
    const int L = P.g.L;
    for (int r = 0; r < L; r++) {
        RoundDesc& R = P.rounds[r];
        // layer eb and slope (compress.tpp:194-195): libm pow, product in double, rounded to float
        const double f = std::pow((double)eb_ratio, (double)(L - r - 1));
        R.eb = (float)((double)abs_eb * f);
        R.layer_slope = (float)((double)default_slope * std::pow((double)slope_ratio, (double)(L - r - 1)));
        R.inv_half_eb = 0.5 / (double)R.eb;
        R.eb_d = (double)R.eb;
        R.neg_two_eb = -(2.0 * (double)R.eb);
    }
}

Plan make_plan(const Geometry& g, float abs_eb, float pivot_eb_ratio, float eb_ratio, float slope_ratio,
               float default_slope, uint64_t h1_num, uint32_t coarse_max, const RoundTune* tune,
               bool tsel_strides) {
    Plan P;
    P.g = g;
    P.pivot_count = 1;
    for (int a = 0; a < 3; a++) {
        P.pn[a] = (uint32_t)((g.D[a] + g.pivot_stride[a] - 1) / g.pivot_stride[a]);
        P.pivot_count *= P.pn[a];
    }
    const uint64_t nbl = g.nblocks();
    if (tsel_strides) {
        P.ss.assign((uint64_t)g.L * nbl, 0);
        P.ns.assign((uint64_t)g.L * nbl, 0);
    }
    uint64_t s[3] = {g.pivot_stride[0], g.pivot_stride[1], g.pivot_stride[2]};
    // every array starts on a 256-byte boundary of the symbol buffer (whole
    // cache lines for the kernels that write and read it)
    uint64_t qb_off = sym_align(P.pivot_count), qb_sum = P.pivot_count;
    for (int r = 0; r < g.L; r++) {
        const int d = g.axis[r];
        RoundDesc R{};
        R.layer = (uint32_t)r;
        R.d = (uint32_t)d;
        R.order = g.order[r];
        R.h = (uint32_t)(s[d] / 2);
        uint64_t dist = s[d] / 2;
        for (int a = d + 1; a < 3; a++) dist *= g.D[a];
        R.dist = dist;
        uint64_t count = 1;
        for (int a = 0; a < 3; a++) {
            R.D[a] = (uint32_t)g.D[a];
            R.s[a] = (uint32_t)s[a];
            // blocks wider than the field behave the same under any width >= D[a]
            R.B[a] = (uint32_t)std::min<uint64_t>(g.block_shape[a], 1ull << 31);
            R.nb[a] = (uint32_t)g.block_num[a];
            if (a == d) R.n[a] = g.D[a] > R.h ? (uint32_t)((g.D[a] - R.h + 2 * R.h - 1) / (2 * R.h)) : 0;
            else R.n[a] = (uint32_t)((g.D[a] + s[a] - 1) / s[a]);
            R.tab_off[a] = (uint32_t)P.tables.size();
            const uint64_t step = a == d ? s[a] : s[a];
            uint64_t acc = 0;
            std::vector<uint32_t> cnt(g.block_num[a]), pre(g.block_num[a]);
            for (uint64_t b = 0; b < g.block_num[a]; b++) {
                uint64_t lo = b * g.block_shape[a];
                uint64_t hi = std::min(lo + g.block_shape[a] + 1, g.D[a]);
                uint64_t start = a == d ? lo + s[a] / 2 : (lo == 0 ? 0 : lo + s[a]);
                uint64_t c = start >= hi ? 0 : (hi - start + step - 1) / step;  // asz_param.hpp:111-112
                cnt[b] = (uint32_t)c;
                pre[b] = (uint32_t)acc;
                acc += c;
            }
            R.C[a] = (uint32_t)acc;
            P.tables.insert(P.tables.end(), cnt.begin(), cnt.end());
            P.tables.insert(P.tables.end(), pre.begin(), pre.end());
            count *= acc;
        }
        R.count = count;
        R.qb_off = qb_off;
        qb_off = sym_align(qb_off + count);
        qb_sum += count;
        // kernel shape: register-window sweep for rounds along axis 0/1, row walk along DIM2 otherwise
        R.kind = d == 2 ? kRoundRow : kRoundSweep;
        // targets per block along d; any value above the target count gives block 0
        R.tpb = make_fastdiv((uint32_t)std::min<uint64_t>(g.block_shape[d] / s[d], 1ull << 30));
        const RoundTune T = tune ? tune[r] : RoundTune{0, 0, 0};
        if (R.kind == kRoundSweep) {
            const uint32_t seg_len = T.seg ? std::min<uint32_t>(T.seg, kSweepSeg) : kSweepSeg;
            R.seg_off = (uint32_t)P.segs.size();
            const uint32_t per_block = (uint32_t)(g.block_shape[d] / s[d]);  // targets per full block along d
            const uint32_t* cntd = P.tables.data() + R.tab_off[d];
            for (uint32_t b = 0; b < R.nb[d]; b++) {
                const uint32_t k0 = b * per_block;
                // equal segments of at most seg_len targets: a block run that fits one
                // segment is walked whole, so its QB run (orders along the round axis)
                // is written by one lane and a wave's runs form one contiguous span
                const uint32_t nseg = (cntd[b] + seg_len - 1) / seg_len;
                for (uint32_t q = 0; q < nseg; q++) {
                    const uint32_t c0 = (uint32_t)((uint64_t)cntd[b] * q / nseg);
                    const uint32_t c1 = (uint32_t)((uint64_t)cntd[b] * (q + 1) / nseg);
                    P.segs.push_back(SweepSeg{k0 + c0, c1 - c0, b, 0});
                }
            }
            R.nseg = (uint32_t)P.segs.size() - R.seg_off;
        }
        if (R.kind == kRoundSweep) {
            // a workgroup spans the next power of two of the DIM2 targets (64 to 256
            // lanes) and as many lattice rows of the other axis as fill its 256 lanes
            R.xsh = 6;
            while (R.xsh < 8 && (1u << R.xsh) < R.n[2]) R.xsh++;
            if (T.xsh) R.xsh = std::min<uint32_t>(std::max<uint32_t>(T.xsh, 6), 8);
            const uint32_t rows = 256u >> R.xsh;
            R.nwg = (uint64_t)((R.n[2] + (1u << R.xsh) - 1) >> R.xsh) * ((R.n[1 - d] + rows - 1) / rows) * R.nseg;
        } else {
            // about 4096 targets per workgroup, but at least ~2048 workgroups when the round has the rows
            const uint64_t rows = (uint64_t)R.n[0] * R.n[1];
            auto up4 = [](uint64_t v) { return (std::max<uint64_t>(v, 1) + 3) / 4 * 4; };
            const uint64_t want = up4((4096 + R.n[2] - 1) / std::max<uint32_t>(R.n[2], 1));
            const uint64_t cap = up4((rows + 2047) / 2048);
            const uint64_t g = 4 * kRowsPerWave;  // rows of one pass of the four waves
            uint64_t rpw = std::min<uint64_t>(std::max<uint64_t>(4, std::min(want, cap)), 1024);
            if (T.rpw) rpw = std::min<uint64_t>(T.rpw, 1024);
            R.rpw = (uint32_t)((rpw + g - 1) / g * g);
            R.nwg = (rows + R.rpw - 1) / R.rpw;
        }
        // a leading round with few targets per block: a block round, one
        // workgroup per block (a partial and a window each); its kernel
        // divides block-local target indices in float, exact below 2^24
        uint64_t most = 1;
        for (int a = 0; a < 3; a++) {
            const uint32_t* cnt = P.tables.data() + R.tab_off[a];
            most *= *std::max_element(cnt, cnt + R.nb[a]);
        }
        // and addresses a block's points by 32-bit offsets from its first target
        const uint64_t slab = (uint64_t)std::min<uint64_t>(g.block_shape[0] + 1, g.D[0]) * g.D[1] * g.D[2];
        if ((uint32_t)r == P.coarse && P.coarse < coarse_max && R.count <= kCoarseTargets * nbl &&
            most < (1ull << 24) && slab < (1ull << 31)) {
            R.kind = kRoundBlock;
            R.nwg = nbl;
            P.coarse++;
        }
        if (R.count == 0) R.nwg = 0;
        R.part_off = P.total_partials;
        P.total_partials += R.nwg;
        uint32_t win = kHistWinMin;
        const uint64_t budget = std::min(kHistWidthWords, std::max(kHistWidthMin, R.count / 4));
        while (win < kHistWinMax && (uint64_t)win * 2 * std::max<uint64_t>(R.nwg, 1) <= budget) win *= 2;
        R.lds_win = std::max(win, kLdsWinMin);
        R.hist_copies = (uint32_t)std::min<uint64_t>(kHistCopies, std::max<uint64_t>(1, R.nwg / kWgPerCopy));
        R.hist_off = P.total_hist_words;
        P.total_hist_words += R.nwg ? (uint64_t)R.hist_copies * R.lds_win : 0;
        // t-selection sample strides (compress.tpp:96-116, get_AB_map.tpp:9-12); 0 marks a fallback block
        for (uint64_t b0 = 0; tsel_strides && b0 < g.block_num[0]; b0++)
            for (uint64_t b1 = 0; b1 < g.block_num[1]; b1++)
                for (uint64_t b2 = 0; b2 < g.block_num[2]; b2++) {
                    uint64_t bb[3] = {b0, b1, b2};
                    uint64_t lo[3], hi[3];
                    bool empty = false;
                    for (int a = 0; a < 3; a++) {
                        lo[a] = bb[a] * g.block_shape[a];
                        hi[a] = std::min(lo[a] + g.block_shape[a] + 1, g.D[a]);
                    }
                    uint64_t bid = (b0 * g.block_num[1] + b1) * g.block_num[2] + b2;
                    if (hi[d] <= lo[d] + 3 * s[d]) { P.ss[(uint64_t)r * nbl + bid] = 0; continue; }
                    for (int a = 0; a < 3; a++)
                        if (a != d && lo[a] != 0) lo[a] += s[a];
                    lo[d] += 3 * s[d] / 2;
                    hi[d] -= 3 * s[d] / 2;
                    for (int a = 0; a < 3; a++) empty |= lo[a] >= hi[a];
                    if (empty) { P.ss[(uint64_t)r * nbl + bid] = 0; continue; }
                    uint64_t vol = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
                    uint64_t den = s[2] * (s[0] * s[1] * h1_num);
                    uint64_t ss = (uint64_t)std::cbrt((double)((float)vol / (float)den));
                    if (ss == 0) ss = 1;
                    if (ss > 0xFFFFFFFFull) ss = 0xFFFFFFFFull;
                    P.ss[(uint64_t)r * nbl + bid] = (uint32_t)ss;
                    uint64_t ns = 1;
                    for (int a = 0; a < 3; a++) {
                        uint64_t step = s[a] * ss;
                        ns *= (hi[a] - lo[a] + step - 1) / step;
                    }
                    P.ns[(uint64_t)r * nbl + bid] = (uint32_t)std::min<uint64_t>(ns, 0xFFFFFFFFull);
                }
        P.rounds.push_back(R);
        s[d] /= 2;
    }
    if (qb_sum != g.N()) fail(ASZ_ERR_DATA, "total_qb_num!=data.size()");
    P.qb_total = qb_off;
    set_plan_bounds(P, abs_eb, pivot_eb_ratio, eb_ratio, slope_ratio, default_slope);
    return P;
}

}  // namespace asz

extern "C" int asz_round_layout(const uint64_t dims[3], const asz_plan* plan, uint64_t bpr, int round,
                                uint64_t* count, uint64_t* pos, uint64_t* slot) {
    try {
        asz::Geometry g = asz::make_geometry(dims, *plan, bpr);
        if (round < 0 || round >= g.L) asz::fail(ASZ_ERR_ARG, "round out of range");
        asz::Plan P = asz::make_plan(g, 1.0f, 0.05f, 0.76f, 0.54f, 1.0f, 256);
        const asz::RoundDesc& R = P.rounds[round];
        const uint64_t n = (uint64_t)R.n[0] * R.n[1] * R.n[2];
        *count = n;
        if (!pos) return ASZ_OK;
        uint64_t t = 0;
        uint32_t k[3], i[3];
        for (k[0] = 0; k[0] < R.n[0]; k[0]++)
            for (k[1] = 0; k[1] < R.n[1]; k[1]++)
                for (k[2] = 0; k[2] < R.n[2]; k[2]++, t++) {
                    asz::target_coords(R, (int)R.d, k, i);
                    asz::TargetLoc L;
                    asz::locate(R, (int)R.d, k, i, L);
                    pos[t] = ((uint64_t)i[0] * R.D[1] + i[1]) * R.D[2] + i[2];
                    slot[t] = asz::qb_slot(R, P.tables.data(), L);
                }
        return ASZ_OK;
    } catch (const asz::Error& e) {
        return e.code;
    }
}
