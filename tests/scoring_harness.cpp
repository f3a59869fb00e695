// Checks the library's h1_psnr (flat bin counts, bin map rebuilt in
// first-occurrence order) against the straightforward form of
// get_h_psnr.tpp:5-31: a bin map filled entry by entry while the AB map is
// walked.  Both must give the same (h1, psnr) bit for bit.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <unordered_map>

#include "scoring.hpp"

using namespace asz;

static int64_t lround_f(float y) {
    float s = y + std::copysign(0.49999997f, y);
    if (!(s > -9223372036854775808.0f && s < 9223372036854775808.0f)) return INT64_MIN;
    return (int64_t)s;
}

static void direct_h1_psnr(const ABMap& m, float t, float abs_eb, float stat_eb, float min_val, float max_val,
                           float& h1, float& psnr) {
    const float neg_2_eb = 0.5f / abs_eb;
    const float two_t = t * 2.0f;
    const float two_eb = abs_eb * 2.0f;
    std::unordered_map<int64_t, uint64_t> qb_map;
    float mse = 0.0f;
    uint64_t n = 0;
    for (const auto& kv : m) {
        float af = (float)std::get<0>(kv.first), bf = (float)std::get<1>(kv.first);
        float err = std::fmaf(bf, 2.0f, af * two_t) * stat_eb;
        int64_t q = lround_f(err * neg_2_eb);
        qb_map[q] += kv.second;
        float de = std::fmaf(-(float)q, two_eb, err);
        double ded = (double)de;
        mse = (float)std::fma((double)kv.second, ded * ded, (double)mse);
        n += kv.second;
    }
    const float dn = (float)n;
    const float rec = 1.0f / dn;
    h1 = 0.0f;
    for (const auto& kv : qb_map) {
        float p = (float)kv.second * rec;
        h1 = std::fmaf(-p, log2f(p), h1);
    }
    const float range = max_val - min_val;
    psnr = log10f((range * range / mse) * dn) * 10.0f;
}

int main() {
    std::mt19937_64 rng(12345);
    const double spreads[] = {3.0, 40.0, 900.0, 3.0e4, 2.0e6, 5.0e18};
    const float ts[] = {1.0f / 16, 0.25f, 9.0f / 32, 5.0f / 16, 11.0f / 32, 0.5f, 1.0f};
    long cases = 0, bad = 0, far_cases = 0;
    for (int rep = 0; rep < 40; rep++) {
        for (double spread : spreads) {
            std::normal_distribution<double> nd(0.0, spread);
            const int n = 1 + (int)(rng() % 3000);
            std::vector<int64_t> keys(2 * n);
            for (int i = 0; i < n; i++) {
                // repeated keys give counts above one
                const int src = (i > 0 && rng() % 4 == 0) ? (int)(rng() % i) : i;
                if (src != i) {
                    keys[2 * i] = keys[2 * src];
                    keys[2 * i + 1] = keys[2 * src + 1];
                } else {
                    keys[2 * i] = (int64_t)std::llround(std::fmax(-9.0e18, std::fmin(9.0e18, nd(rng))));
                    keys[2 * i + 1] = (int64_t)std::llround(std::fmax(-9.0e18, std::fmin(9.0e18, nd(rng))));
                }
            }
            ABMap m;
            build_ab_map(keys.data(), n, m);
            FlatAB f;
            flatten_ab(m, f);
            const float abs_eb = (float)(0.5 + (rng() % 1000) / 250.0);
            const float stat_eb = abs_eb * 0.2f * 0.5f;
            for (float t : ts) {
                float h1a, pa, h1b, pb;
                direct_h1_psnr(m, t, abs_eb, stat_eb, -1.0f, 3.0f, h1a, pa);
                h1_psnr(f, t, abs_eb, stat_eb, -1.0f, 3.0f, h1b, pb);
                cases++;
                if (spread >= 2.0e6) far_cases++;
                if (std::memcmp(&h1a, &h1b, 4) != 0 || std::memcmp(&pa, &pb, 4) != 0) {
                    if (bad < 5)
                        std::printf("mismatch spread %g n %d t %g: h1 %a vs %a psnr %a vs %a\n", spread, n, t, h1a,
                                    h1b, pa, pb);
                    bad++;
                }
            }
        }
    }
    std::printf("cases %ld far %ld mismatches %ld\n", cases, far_cases, bad);
    return bad ? 1 : 0;
}
