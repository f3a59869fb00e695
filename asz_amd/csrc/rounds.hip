// Range scan, pivots and the level-synchronous interpolation rounds.
//
// One launch per round covers every block.  The reference runs block -> round
// (compress.tpp:179-208); every round-r read touches points written in rounds
// < r or the pivots, and every target is owned by exactly one block, so the
// round-major order here reads exactly the same values.  The field is updated
// in place: targets hold the original value until their own thread rewrites
// them, neighbours (coarse points, even multiples of h along the round axis)
// are never written during the round.
//
// Two kernel shapes:
//  * k_round_sweep (round axis 0 or 1): a wave owns 64 consecutive lattice
//    columns along DIM2 and walks a segment (<= kSeg targets) of one block
//    along the round axis, keeping the coarse neighbours in a register window
//    (each coarse value is loaded once per segment).  QB values go straight
//    to their slots when the serialization order runs along DIM2 ("012"), are
//    staged in LDS and written as contiguous runs when it runs along the round
//    axis ("120" on axis 0, "021" on axis 1), and are scattered otherwise.
//  * k_round_row (round axis 2, every order): waves walk whole lattice rows
//    along DIM2, one target per lane per step; k_round_quad for h = 1, two
//    targets per lane in one 16-byte load and store.
// Error statistics and QB ranges are reduced per workgroup into partials and
// folded right after each round (k_fold_round); each compress workgroup also
// leaves an LDS histogram window of its QBs for the Huffman coder.
#include "device.hpp"
#include "launch.hpp"
#include "layout.hpp"

namespace asz {
namespace dev {

// ------------------------------------------------------------------ range scan

__device__ __forceinline__ uint32_t order_key(float x) {
    x = daz(x);
    if (x == 0.0f) x = 0.0f;  // -0 == +0
    uint32_t b = __float_as_uint(x);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Range scan.  Per field, kRangeWords words at out: [0] min key, [1] max key,
// [2] non-finite flag, [4:5] / [6:7] the first / last index of an element
// whose DAZ value is zero (uint64; first = ~0 when there is none), which the
// zero ties of min_element / max_element need (their own bits).  With xs,
// field blockIdx.y of several (its array xs[y], its words at out + 8 y).
__global__ __launch_bounds__(256) void k_range(const float* __restrict__ x, uint64_t n, uint32_t* out,
                                               const float* const* __restrict__ xs) {
    __shared__ uint32_t s_mn[4], s_mx[4], s_bad[4];
    __shared__ unsigned long long s_zf[4], s_zl[4];
    if (xs) {
        x = xs[blockIdx.y];
        out += kRangeWords * blockIdx.y;
    }
    constexpr uint32_t kZeroKey = 0x80000000u;  // order_key of every DAZ-zero value
    uint32_t mn = 0xFFFFFFFFu, mx = 0u, bad = 0u;
    unsigned long long zf = ~0ull, zl = 0ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n4 = n / 4;
    const float4* x4 = reinterpret_cast<const float4*>(x);
    auto scan4 = [&](const float4 v, uint64_t i) {
        const float e[4] = {v.x, v.y, v.z, v.w};
        uint32_t key[4];
        bool z = false;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            bad |= !isfinite(e[k]);
            key[k] = order_key(e[k]);
            mn = min(mn, key[k]);
            mx = max(mx, key[k]);
            z |= key[k] == kZeroKey;
        }
        // zero ties are rare: the index bookkeeping runs only in waves that meet one
        if (__ballot(z)) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (key[k] == kZeroKey) {
                    zf = min(zf, (unsigned long long)(4 * i + k));
                    zl = max(zl, (unsigned long long)(4 * i + k));
                }
            }
        }
    };
    // a workgroup takes 4 x 256 consecutive float4 (16 KB) at a time: four
    // coalesced 16-byte loads in flight per thread, then their keys
    const uint64_t span = 4ull * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * span + threadIdx.x;
    for (; i + 3 * blockDim.x < n4; i += span * gridDim.x) {
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = x4[i + u * blockDim.x];
#pragma unroll
        for (int u = 0; u < 4; u++) scan4(v[u], i + u * blockDim.x);
    }
    for (; i < n4; i += blockDim.x) scan4(x4[i], i);
    for (uint64_t i = n4 * 4 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        bad |= !isfinite(x[i]);
        const uint32_t key = order_key(x[i]);
        mn = min(mn, key);
        mx = max(mx, key);
        if (key == kZeroKey) {
            zf = min(zf, (unsigned long long)i);
            zl = max(zl, (unsigned long long)i);
        }
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    bad = wave_max(bad);
    zf = wave_min(zf);
    zl = wave_max(zl);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        s_mn[w] = mn;
        s_mx[w] = mx;
        s_bad[w] = bad;
        s_zf[w] = zf;
        s_zl[w] = zl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) {
            mn = min(mn, s_mn[k]);
            mx = max(mx, s_mx[k]);
            bad |= s_bad[k];
            zf = min(zf, s_zf[k]);
            zl = max(zl, s_zl[k]);
        }
        // only a workgroup that improves on the value already there issues the
        // atomic: thousands of atomics on one address serialise in L2 (a stale
        // read only costs a redundant atomic)
        const volatile uint32_t* vo = out;
        if (mn < vo[0]) atomicMin(&out[0], mn);
        if (mx > vo[1]) atomicMax(&out[1], mx);
        if (bad) atomicOr(&out[2], 1u);
        unsigned long long* z = reinterpret_cast<unsigned long long*>(out + 4);
        if (zf != ~0ull) {
            atomicMin(&z[0], zf);
            atomicMax(&z[1], zl);
        }
    }
}

// The field's error-bound scalars from its range scan, on the device, so that
// compression can follow the scan without a host round trip: min / max as
// load_input takes them (a DAZ-zero extreme keeps the bits of the first
// smallest / last largest element), the absolute bound, the pivot bound
// (compress.tpp:166) and every layer's bound (compress.tpp:194: the host
// gives (double) eb_ratio^(L-r-1) and the layer slopes, which do not depend on
// the range).  res[0..1] = min / max bits, res[2] = status (0 valid, 1
// non-finite input, 2 max == min).  One workgroup of 64 threads.
__global__ __launch_bounds__(64) void k_field_desc(const uint32_t* __restrict__ rng, const float* __restrict__ x,
                                                   RangeSeed seed, FieldDesc* fd, uint32_t* res) {
    auto key_to_float = [](uint32_t k) {
        const uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
        return __uint_as_float(b);
    };
    float vmin = key_to_float(rng[0]), vmax = key_to_float(rng[1]);
    const unsigned long long* z = reinterpret_cast<const unsigned long long*>(rng + 4);
    if (vmin == 0.0f && z[0] != ~0ull) vmin = x[z[0]];
    if (vmax == 0.0f && z[0] != ~0ull) vmax = x[z[1]];
    const float range = daz(daz(vmax) - daz(vmin));
    const float aeb = daz(daz(seed.rel_eb) * range);
    const float pe = daz(aeb) * daz(seed.pivot_eb_ratio);
    const uint32_t r = threadIdx.x;
    if (r < seed.L) {
        FieldRound q;
        q.eb = (float)((double)aeb * seed.eb_factor[r]);
        q.layer_slope = seed.layer_slope[r];
        q.inv_half_eb = 0.5 / (double)q.eb;
        q.eb_d = (double)q.eb;
        q.neg_two_eb = -(2.0 * (double)q.eb);
        fd->r[r] = q;
    }
    if (r == 0) {
        fd->work = seed.work;
        fd->inv_half_pe = 0.5 / (double)pe;
        fd->two_pe = (double)pe * 2.0;
        fd->range = daz(vmax) - daz(vmin);
        fd->pad = 0;
        res[0] = __float_as_uint(vmin);
        res[1] = __float_as_uint(vmax);
        res[2] = rng[2] ? 1u : (daz(vmin) == daz(vmax) ? 2u : 0u);
        res[3] = 0;
        __threadfence_system();  // res is pinned host memory, read after later kernels' results
    }
}

// ------------------------------------------------------------------ partial reductions


__device__ __forceinline__ void init_acc(DevAcc* acc) {
    const uint32_t i = threadIdx.x;
    if (i == 0) {
        acc->err = ErrAcc{0.0, 0.0};
        acc->fallbacks = 0;
    }
    if (i < ASZ_MAX_LAYERS + 2) acc->arr[i] = ArrayAcc{INT32_MAX, INT32_MIN, {0, 0}};
    if (i < ASZ_MAX_LAYERS) {
        acc->rerr[i] = ErrAcc{0.0, 0.0};
        acc->fold_count[i] = 0;
    }
}

__global__ __launch_bounds__(128) void k_init_acc(DevAcc* acc) {
    init_acc(acc + blockIdx.x);  // one workgroup per field
}

// A compression call's start in one launch: workgroups [0, nfields) set the
// fields' accumulators (as k_init_acc), the others clear n4 16-byte words.
__global__ __launch_bounds__(256) void k_init_call(DevAcc* acc, uint32_t nfields, uint4* v, uint64_t n4) {
    if (blockIdx.x < nfields) {
        init_acc(acc + blockIdx.x);
        return;
    }
    const uint64_t stride = (uint64_t)(gridDim.x - nfields) * blockDim.x;
    for (uint64_t i = (uint64_t)(blockIdx.x - nfields) * blockDim.x + threadIdx.x; i < n4; i += stride)
        v[i] = make_uint4(0, 0, 0, 0);
}

// ------------------------------------------------------------------ pivots (compress.tpp:161-175)

// One thread per pivot (lattice i_a = 0 mod pivot_stride_a, C order); every
// workgroup leaves one Partial (error, QB range) at part[blockIdx.x].
__device__ __forceinline__ void pivot_coords(const PivotDesc& P, uint64_t i, uint64_t& pos) {
    const uint32_t k2 = (uint32_t)(i % P.pn[2]);
    const uint64_t k01 = i / P.pn[2];
    const uint32_t k1 = (uint32_t)(k01 % P.pn[1]), k0 = (uint32_t)(k01 / P.pn[1]);
    pos = ((uint64_t)k0 * P.ps[0] * P.D[1] + (uint64_t)k1 * P.ps[1]) * P.D[2] + (uint64_t)k2 * P.ps[2];
}

__global__ __launch_bounds__(256) void k_pivots_compress(float* __restrict__ work, int32_t* __restrict__ qb,
                                                         PivotDesc P, Partial* __restrict__ part, FieldSet fs) {
    __shared__ double s_err[4], s_sq[4];
    __shared__ int32_t s_min[4], s_max[4];
    if (fs.fd) {  // field blockIdx.y of a multi-field call
        const uint32_t f = blockIdx.y;
        const FieldDesc& F = fs.fd[f];
        work = F.work;
        qb += f * fs.sym;
        part += f * fs.fpart;
        P.inv_half_pe = F.inv_half_pe;
        P.two_pe = F.two_pe;
    }
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double err = 0, sq = 0;
    int32_t qmin = INT32_MAX, qmax = INT32_MIN;
    if (i < P.count) {
        uint64_t pos;
        pivot_coords(P, i, pos);
        const float x = work[pos];
        const int32_t q = lround_q((double)daz(x) * P.inv_half_pe);
        const float r = (float)((double)q * P.two_pe);
        work[pos] = r;
        qb[i] = q;
        const double e = (double)daz(x) - (double)r;
        err = fabs(e);
        sq = e * e;
        qmin = qmax = q;
    }
    err = wave_max(err);
    sq = wave_sum(sq);
    qmin = wave_min(qmin);
    qmax = wave_max(qmax);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) { s_err[w] = err; s_sq[w] = sq; s_min[w] = qmin; s_max[w] = qmax; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) {
            err = fmax(err, s_err[k]);
            sq += s_sq[k];
            qmin = min(qmin, s_min[k]);
            qmax = max(qmax, s_max[k]);
        }
        part[blockIdx.x] = Partial{err, sq, qmin, qmax};
    }
}

__global__ __launch_bounds__(256) void k_pivots_decompress(float* __restrict__ out, const int32_t* __restrict__ qb,
                                                           PivotDesc P) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.count) return;
    uint64_t pos;
    pivot_coords(P, i, pos);
    out[pos] = (float)((double)qb[i] * P.two_pe);
}

// ------------------------------------------------------------------ row rounds (axis 2)

// Round along DIM2.  Workgroup x owns lattice rows [x rpw, (x+1) rpw) of the
// (k0, k1) lattice; wave w takes kRowsPerWave consecutive rows at a time and
// lane l the targets k2 = l, l+64, ... of each, so loads and the "012" QB
// stores are coalesced along DIM2, and the loads of all the rows are issued
// before any of them is used.  Per-row quantities (blocks and QB bases along
// axes 0/1) are wave-uniform.  With h == 1 and an even row length a row is
// rewritten whole: each lane stores the pair {coarse left neighbour, target},
// so no cache line is written partially (the neighbour is rewritten with the
// value it holds).
//
// The serialization order is a template parameter: with a run-time switch on
// R.order inside this divergent loop, the gfx950 code generated by this
// toolchain added j1 instead of j2 on the "012" path (seen as QBs stored one
// row-stride off); compile-time orders leave no switch to lower.
template <bool QUIN, bool DEC, int ORDER>
__global__ __launch_bounds__(256) void k_round_row(RoundDesc R, float* __restrict__ work, int32_t* __restrict__ qb,
                                                   const uint32_t* __restrict__ tabs, const float* __restrict__ tblk,
                                                   Partial* part, uint32_t* __restrict__ phist,
                                                   uint32_t* __restrict__ gwin, FieldSet fs) {
    extern __shared__ uint32_t lh[];  // compress: LDS histogram window of R.lds_win bins
    if (!DEC && fs.fd) {  // field blockIdx.y of a multi-field call
        const uint32_t f = blockIdx.y;
        const FieldDesc& F = fs.fd[f];
        work = F.work;
        qb += f * fs.sym;
        tblk += f * fs.tblk;
        part += f * fs.part;
        phist += f * fs.phist;
        gwin += f * fs.ghist;
        field_round(F, R);
    }
    constexpr int RU = (int)kRowsPerWave;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t win = R.lds_win;
    if (!DEC) {
        for (uint32_t i = threadIdx.x; i < win; i += blockDim.x) lh[i] = 0;
        __syncthreads();
    }
    const uint32_t h = R.h, n1 = R.n[1], n2 = R.n[2];
    const int64_t D2 = R.D[2], hh = h;
    const uint64_t rows = (uint64_t)R.n[0] * n1;
    const bool pair = h == 1 && (R.D[2] & 1u) == 0 && (reinterpret_cast<uintptr_t>(work) & 7u) == 0;
    double err = 0, sq = 0;
    int32_t qmin = INT32_MAX, qmax = INT32_MIN;
    for (uint32_t rr0 = wave * RU; rr0 < R.rpw; rr0 += 4 * RU) {
        const uint64_t row0 = (uint64_t)blockIdx.x * R.rpw + rr0;
        if (row0 >= rows) break;
        float* rowp[RU];
        const float* trow[RU];
        uint32_t j0[RU], j1[RU], c0[RU], c1[RU], p0[RU], p1[RU];
        bool ok[RU];
#pragma unroll
        for (int u = 0; u < RU; u++) {
            ok[u] = row0 + u < rows;
            const uint64_t row = ok[u] ? row0 + u : row0;
            const uint32_t k0 = (uint32_t)(row / n1), k1 = (uint32_t)(row - (uint64_t)k0 * n1);
            uint32_t b0, b1;
            side_loc(R.s[0] * k0, k0, R.B[0], R.s[0], b0, j0[u]);
            side_loc(R.s[1] * k1, k1, R.B[1], R.s[1], b1, j1[u]);
            c0[u] = tabs[R.tab_off[0] + b0];
            p0[u] = tabs[R.tab_off[0] + R.nb[0] + b0];
            c1[u] = tabs[R.tab_off[1] + b1];
            p1[u] = tabs[R.tab_off[1] + R.nb[1] + b1];
            rowp[u] = work + ((uint64_t)(R.s[0] * k0) * R.D[1] + R.s[1] * k1) * R.D[2];
            trow[u] = tblk + ((uint64_t)b0 * R.nb[1] + b1) * R.nb[2];
        }
        for (uint32_t k2 = lane; k2 < n2; k2 += 64) {
            const uint32_t i2 = h + 2 * h * k2;
            const uint32_t b2 = fdiv(k2, R.tpb);
            const uint32_t j2 = k2 - b2 * R.tpb.d;
            const uint32_t right = (uint32_t)min<uint64_t>((uint64_t)(b2 + 1) * R.B[2] + 1, R.D[2]);
            const uint32_t c2 = tabs[R.tab_off[2] + b2], p2 = tabs[R.tab_off[2] + R.nb[2] + b2];
            // every load of the RU rows first, unconditionally (positions clamped into
            // the row; predict_w reads only the in-block ones): window w[u][m] at
            // i2 + (2m - 5) h, the target, and the block's t
            float w[RU][6], xv[RU], tv[RU];
#pragma unroll
            for (int u = 0; u < RU; u++) {
#pragma unroll
                for (int m = 0; m < 6; m++) {
                    if (!QUIN && (m == 0 || m == 5)) {
                        w[u][m] = 0.0f;
                        continue;
                    }
                    const int64_t pos = min(max((int64_t)i2 + (2 * m - 5) * hh, (int64_t)0), D2 - 1);
                    w[u][m] = rowp[u][pos];
                }
                xv[u] = DEC ? 0.0f : rowp[u][i2];
                tv[u] = trow[u][b2];
            }
#pragma unroll
            for (int u = 0; u < RU; u++) {
                if (!ok[u]) continue;
                const uint32_t cnt[3] = {c0[u], c1[u], c2}, pre[3] = {p0[u], p1[u], p2}, jj[3] = {j0[u], j1[u], j2};
                const uint64_t slot = R.qb_off + block_base(R, cnt, pre) + local_slot(ORDER, jj, cnt);
                const float p = predict_w<QUIN>(i2, h, right, w[u], tv[u]);
                float xr;
                if (DEC) {
                    xr = reconstruct(p, qb[slot], R.neg_two_eb);
                } else {
                    const float x = xv[u];
                    const int32_t q = quantize<2>(p, x, R);
                    xr = reconstruct(p, q, R.neg_two_eb);
                    qb[slot] = q;
                    const double e = (double)daz(x) - (double)xr;
                    err = fmax(err, fabs(e));
                    sq += e * e;
                    qmin = min(qmin, q);
                    qmax = max(qmax, q);
                    hist_add(lh, win, gwin, q);
                }
                if (pair) *reinterpret_cast<float2*>(rowp[u] + i2 - 1) = make_float2(w[u][2], xr);
                else rowp[u][i2] = xr;
            }
        }
    }
    if (!DEC) finish_wg(err, sq, qmin, qmax, R, blockIdx.x, part, lh, phist);
}

// Axis-2 rounds with h = 1 on rows of a multiple of four elements (16-byte
// aligned field): targets are the odd elements of every lattice row, the
// coarse values the even ones.  The workgroup layout is k_round_row's (waves
// walk R.rpw rows per workgroup, kRowsPerWave at a time), but a lane reads
// and writes four consecutive elements (two targets) with one 16-byte access
// each, instead of five loads and one store per target.
template <bool QUIN, bool DEC, int ORDER>
__global__ __launch_bounds__(256) void k_round_quad(RoundDesc R, float* __restrict__ work, int32_t* __restrict__ qb,
                                                    const uint32_t* __restrict__ tabs, const float* __restrict__ tblk,
                                                    Partial* part, uint32_t* __restrict__ phist,
                                                    uint32_t* __restrict__ gwin, FieldSet fs) {
    extern __shared__ uint32_t lh[];  // compress: LDS histogram window of R.lds_win bins
    if (!DEC && fs.fd) {  // field blockIdx.y of a multi-field call
        const uint32_t f = blockIdx.y;
        const FieldDesc& F = fs.fd[f];
        work = F.work;
        qb += f * fs.sym;
        tblk += f * fs.tblk;
        part += f * fs.part;
        phist += f * fs.phist;
        gwin += f * fs.ghist;
        field_round(F, R);
    }
    constexpr int RU = (int)kRowsPerWave;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t win = R.lds_win;
    if (!DEC) {
        for (uint32_t i = threadIdx.x; i < win; i += blockDim.x) lh[i] = 0;
        __syncthreads();
    }
    const uint32_t n1 = R.n[1];
    const int64_t D2 = R.D[2];
    const uint64_t rows = (uint64_t)R.n[0] * n1;
    double err = 0, sq = 0;
    int32_t qmin = INT32_MAX, qmax = INT32_MIN;
    for (uint32_t rr0 = wave * RU; rr0 < R.rpw; rr0 += 4 * RU) {
        // Lane l of a wave holds elements 4l .. 4l+3 of a 256-element segment
        // (coarse 4l, 4l+2; targets 4l+1, 4l+3) in one float4; the other
        // coarse neighbours come from lanes l-1, l+1 (l+2 quintic) by
        // shuffles, and across the segment's ends from the halo float4s
        // x[seg-4 .. seg-1] (lane 0) and x[seg+256 .. seg+263] (lane 63).
        // Each lane stores its four elements back whole (the coarse values
        // unchanged), so rows are rewritten in full lines.
        const uint64_t row0 = (uint64_t)blockIdx.x * R.rpw + rr0;
        if (row0 >= rows) break;
        float* rowp[RU];
        const float* trow[RU];
        uint32_t j0[RU], j1[RU], c0[RU], c1[RU], p0[RU], p1[RU];
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const uint64_t row = row0 + u < rows ? row0 + u : row0;
            const uint32_t k0 = (uint32_t)(row / n1), k1 = (uint32_t)(row - (uint64_t)k0 * n1);
            uint32_t b0, b1;
            side_loc(R.s[0] * k0, k0, R.B[0], R.s[0], b0, j0[u]);
            side_loc(R.s[1] * k1, k1, R.B[1], R.s[1], b1, j1[u]);
            c0[u] = tabs[R.tab_off[0] + b0];
            p0[u] = tabs[R.tab_off[0] + R.nb[0] + b0];
            c1[u] = tabs[R.tab_off[1] + b1];
            p1[u] = tabs[R.tab_off[1] + R.nb[1] + b1];
            rowp[u] = work + ((uint64_t)(R.s[0] * k0) * R.D[1] + R.s[1] * k1) * R.D[2];
            trow[u] = tblk + ((uint64_t)b0 * R.nb[1] + b1) * R.nb[2];
        }
        for (uint32_t seg = 0; seg < (uint32_t)D2; seg += 256) {
            const uint32_t e0 = seg + 4 * lane;  // this lane's first element
            const bool live = e0 < (uint32_t)D2;
            // the lane's two targets i2 = e0 + 1 + 2e: block along DIM2, local
            // index, block end and the block's table entries
            uint32_t b2[2], j2[2], right[2], c2[2], p2[2];
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const uint32_t k2 = live ? (e0 + 2 * e) / 2 : 0;
                b2[e] = fdiv(k2, R.tpb);
                j2[e] = k2 - b2[e] * R.tpb.d;
                right[e] = (uint32_t)min<uint64_t>((uint64_t)(b2[e] + 1) * R.B[2] + 1, R.D[2]);
                c2[e] = tabs[R.tab_off[2] + b2[e]];
                p2[e] = tabs[R.tab_off[2] + R.nb[2] + b2[e]];
            }
            // halo: lane 0 the coarse x[seg-4], x[seg-2]; lane 63 x[seg+256],
            // x[seg+258] and (quintic) x[seg+260]
            const bool hl = lane == 0 && seg > 0, hr = lane == 63 && seg + 256 < (uint32_t)D2;
            const bool hr2 = QUIN && lane == 63 && seg + 260 < (uint32_t)D2;
            const uint32_t ha = hl ? seg - 4 : (hr ? seg + 256 : e0), hb = ha + 2, hc = hr2 ? seg + 260 : e0;
            float4 v[RU];
            float h0[RU], h1[RU], h2[RU], tv[RU][2];
#pragma unroll
            for (int u = 0; u < RU; u++) {
                const float4* r4 = reinterpret_cast<const float4*>(rowp[u]);
                v[u] = live ? r4[e0 / 4] : make_float4(0.f, 0.f, 0.f, 0.f);
                h0[u] = hl || hr ? rowp[u][ha] : 0.0f;
                h1[u] = hl || hr ? rowp[u][hb] : 0.0f;
                h2[u] = hr2 ? rowp[u][hc] : 0.0f;
                tv[u][0] = trow[u][b2[0]];
                tv[u][1] = trow[u][b2[1]];
            }
#pragma unroll
            for (int u = 0; u < RU; u++) {
                // coarse neighbours: the previous lane's .x / .z, the next lanes' .x / .z
                float px = __shfl_up(v[u].x, 1, 64), pz = __shfl_up(v[u].z, 1, 64);
                float nx = __shfl_down(v[u].x, 1, 64), nz = __shfl_down(v[u].z, 1, 64);
                float n2x = __shfl_down(v[u].x, 2, 64);
                const float hx1 = __shfl(h0[u], 63, 64);
                if (lane == 0) {
                    px = h0[u];
                    pz = h1[u];
                }
                if (lane == 62) n2x = hx1;
                if (lane == 63) {
                    nx = h0[u];
                    nz = h1[u];
                    n2x = h2[u];
                }
                if (row0 + u >= rows || !live) continue;
                float4 out = v[u];
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    const uint32_t i2 = e0 + 1 + 2 * e;  // target 4l+1, then 4l+3
                    if (i2 >= (uint32_t)D2) break;
                    // window at i2 - 5 .. i2 + 5 over the coarse values
                    float w[6];
                    if (e == 0) {
                        w[0] = px; w[1] = pz; w[2] = v[u].x; w[3] = v[u].z; w[4] = nx; w[5] = nz;
                    } else {
                        w[0] = pz; w[1] = v[u].x; w[2] = v[u].z; w[3] = nx; w[4] = nz; w[5] = n2x;
                    }
                    const uint32_t cnt[3] = {c0[u], c1[u], c2[e]}, pre[3] = {p0[u], p1[u], p2[e]};
                    const uint32_t jj[3] = {j0[u], j1[u], j2[e]};
                    const uint64_t slot = R.qb_off + block_base(R, cnt, pre) + local_slot(ORDER, jj, cnt);
                    const float p = predict_w<QUIN>(i2, 1, right[e], w, tv[u][e]);
                    float xr;
                    if (DEC) {
                        xr = reconstruct(p, qb[slot], R.neg_two_eb);
                    } else {
                        const float x = e == 0 ? v[u].y : v[u].w;
                        const int32_t q = quantize<2>(p, x, R);
                        xr = reconstruct(p, q, R.neg_two_eb);
                        qb[slot] = q;
                        const double er = (double)daz(x) - (double)xr;
                        err = fmax(err, fabs(er));
                        sq += er * er;
                        qmin = min(qmin, q);
                        qmax = max(qmax, q);
                        hist_add(lh, win, gwin, q);
                    }
                    if (e == 0) out.y = xr;
                    else out.w = xr;
                }
                reinterpret_cast<float4*>(rowp[u])[e0 / 4] = out;
            }
        }
    }
    if (!DEC) finish_wg(err, sq, qmin, qmax, R, blockIdx.x, part, lh, phist);
}

// ------------------------------------------------------------------ sweep rounds (axis 0 / 1)

// WMODE 0: order "012" (QB runs along DIM2)      -> direct per-lane writes
// WMODE 1: order runs along the round axis       -> LDS tile, contiguous runs
// WMODE 2: order runs along the other axis       -> scattered per-lane writes
template <int AX, bool QUIN, bool DEC, int WMODE>
__global__ __launch_bounds__(256) void k_round_sweep(RoundDesc R, float* __restrict__ work, int32_t* __restrict__ qb,
                                                     const uint32_t* __restrict__ tabs,
                                                     const float* __restrict__ tblk, const SweepSeg* __restrict__ segs,
                                                     Partial* part, uint32_t* __restrict__ phist,
                                                     uint32_t* __restrict__ gwin, FieldSet fs) {
    constexpr int OA = 1 - AX;  // the non-round axis other than DIM2
    // grid z: field f's segments at f * nseg .. (f + 1) * nseg - 1
    uint32_t sz = blockIdx.z;
    if (!DEC && fs.fd) {
        const uint32_t f = blockIdx.z / R.nseg;
        sz -= f * R.nseg;
        const FieldDesc& F = fs.fd[f];
        work = F.work;
        qb += f * fs.sym;
        tblk += f * fs.tblk;
        part += f * fs.part;
        phist += f * fs.phist;
        gwin += f * fs.ghist;
        field_round(F, R);
    }
    __shared__ int32_t tile[WMODE == 1 ? 4 : 1][WMODE == 1 ? kSweepSeg : 1][65];
    __shared__ uint64_t lbase[WMODE == 1 ? 4 : 1][64];
    extern __shared__ uint32_t lh[];  // compress: LDS histogram window of R.lds_win bins
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (!DEC) {
        for (uint32_t i = threadIdx.x; i < R.lds_win; i += blockDim.x) lh[i] = 0;
        if (WMODE != 1) __syncthreads();
    }
    const SweepSeg S = segs[R.seg_off + sz];
    // lanes: 1 << xsh consecutive DIM2 columns (whole waves) by 256 >> xsh rows of axis OA
    const uint32_t k2 = (blockIdx.x << R.xsh) + (threadIdx.x & ((1u << R.xsh) - 1));
    const uint32_t ko = blockIdx.y * (256u >> R.xsh) + (threadIdx.x >> R.xsh);
    const bool valid = k2 < R.n[2] && ko < R.n[OA];
    const uint32_t kk2 = valid ? k2 : 0, kko = valid ? ko : 0;
    const uint32_t h = R.h;

    // block coordinates and local indices of this lane's column
    uint32_t b[3], j[3];
    b[AX] = S.b;
    j[AX] = S.k0 - S.b * (R.B[AX] / (2 * h));
    side_loc(R.s[2] * kk2, kk2, R.B[2], R.s[2], b[2], j[2]);
    side_loc(R.s[OA] * kko, kko, R.B[OA], R.s[OA], b[OA], j[OA]);
    uint32_t cnt[3], pre[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        cnt[a] = tabs[R.tab_off[a] + b[a]];
        pre[a] = tabs[R.tab_off[a] + R.nb[a] + b[a]];
    }
    const uint64_t base = R.qb_off + block_base(R, cnt, pre);
    const uint32_t right = min((S.b + 1) * R.B[AX] + 1, R.D[AX]);
    const float t = tblk[((uint64_t)b[0] * R.nb[1] + b[1]) * R.nb[2] + b[2]];
    const uint64_t sd = AX == 0 ? (uint64_t)R.D[1] * R.D[2] : (uint64_t)R.D[2];
    const uint64_t so = OA == 0 ? (uint64_t)R.D[1] * R.D[2] : (uint64_t)R.D[2];
    const uint64_t col = (uint64_t)(R.s[OA] * kko) * so + (uint64_t)R.s[2] * kk2;
    const int64_t Dd = R.D[AX];

    // QB slot of target c of the segment
    auto slot_of = [&](uint32_t c) -> uint64_t {
        uint32_t jj[3] = {j[0], j[1], j[2]};
        jj[AX] += c;
        return base + local_slot(R.order, jj, cnt);
    };
    constexpr uint64_t kNoLane = ~0ull;  // lbase of a lane past the round's targets
    if (WMODE == 1) {
        lbase[wave][lane] = valid ? slot_of(0) : kNoLane;
        if (DEC) {
            __syncthreads();
            for (uint32_t e = lane; e < 64 * S.cnt; e += 64) {
                const uint32_t jl = e / S.cnt, c = e - jl * S.cnt;
                if (lbase[wave][jl] != kNoLane) tile[wave][c][jl] = qb[lbase[wave][jl] + c];
            }
        }
        __syncthreads();
    }

    double err = 0, sq = 0;
    int32_t qmin = INT32_MAX, qmax = INT32_MIN;
    if (valid) {
        // Every load of the segment is issued before any use: cw[m] holds the
        // coarse value at i0 + (2m - 5) h, xs[c] the original value of target c.
        constexpr int SEG = (int)kSweepSeg;
        const uint32_t cnt_s = S.cnt;
        const uint32_t i0 = h + 2 * h * S.k0;
        const int64_t hh = (int64_t)h;
        float cw[SEG + 5];
#pragma unroll
        for (int m = 0; m < SEG + 5; m++) {
            const int64_t c = (int64_t)i0 + (2 * m - 5) * hh;
            const bool need = QUIN ? m <= (int)cnt_s + 4 : (m >= 1 && m <= (int)cnt_s + 3);
            cw[m] = (need && c >= 0 && c < Dd) ? work[col + (uint64_t)c * sd] : 0.0f;
        }
        float xs[SEG];
        if (!DEC) {
#pragma unroll
            for (int c = 0; c < SEG; c++)
                xs[c] = c < (int)cnt_s ? work[col + (uint64_t)(i0 + 2 * h * c) * sd] : 0.0f;
        }
#pragma unroll
        for (int c = 0; c < SEG; c++) {
            if (c >= (int)cnt_s) continue;  // no early exit: every index stays a compile-time constant
            const uint32_t i = i0 + 2 * h * c;
            float* rp = work + col + (uint64_t)i * sd;
            const float w[6] = {cw[c], cw[c + 1], cw[c + 2], cw[c + 3], cw[c + 4], cw[c + 5]};
            const float p = predict_w<QUIN>(i, h, right, w, t);
            if (DEC) {
                const int32_t q = WMODE == 1 ? tile[wave][c][lane] : qb[slot_of(c)];
                *rp = reconstruct(p, q, R.neg_two_eb);
            } else {
                const float x = xs[c];
                const int32_t q = quantize<AX>(p, x, R);
                const float xr = reconstruct(p, q, R.neg_two_eb);
                *rp = xr;
                if (WMODE == 1) tile[wave][c][lane] = q;
                else qb[slot_of(c)] = q;
                const double e = (double)daz(x) - (double)xr;
                err = fmax(err, fabs(e));
                sq += e * e;
                qmin = min(qmin, q);
                qmax = max(qmax, q);
                hist_add(lh, R.lds_win, gwin, q);
            }
        }
    }
    if (!DEC) {
        if (WMODE == 1) {
            __syncthreads();
            for (uint32_t e = lane; e < 64 * S.cnt; e += 64) {
                const uint32_t jl = e / S.cnt, c = e - jl * S.cnt;
                if (lbase[wave][jl] != kNoLane) qb[lbase[wave][jl] + c] = tile[wave][c][jl];
            }
        }
        finish_wg(err, sq, qmin, qmax, R, blockIdx.x + (uint64_t)gridDim.x * (blockIdx.y + (uint64_t)gridDim.y * sz),
                  part, lh, phist);
    }
}

// ------------------------------------------------------------------ QB order moves

// Rounds whose serialization order runs along neither DIM2 nor the round axis
// ("120" / "021" on the row kernel, the other lattice axis on the sweep
// kernel) would store their QBs 4 bytes at a time, one cache line per lane.
// Their round kernels write the QBs in "012" order instead (coalesced, into a
// staging array), and k_qb_order moves every block's array into its order;
// before a decompression round it moves them back.  Within a block of counts
// (c0, c1, c2) the "012" index of (j0, j1, j2) is (j0 c1 + j1) c2 + j2, so
// order "120" is the transpose [c0][c1 c2] -> [c1 c2][c0] and order "021" the
// transpose [c1][c2] -> [c2][c1] of each j0 slab: a = p Q + q  <->  b = q P + p
// per slab.  A workgroup moves one tile of kXT^2 elements of one slab through
// LDS: TP rows along p (the next power of two of the round's longest p, at
// most kXT) by kXT^2 / TP along q, so a block with short runs is still moved
// by few workgroups.  Each side is accessed in its own row order, so whole
// rows (or, when the tile spans the full row, the whole tile) are contiguous
// on both sides.
constexpr uint32_t kXT = 64;

template <bool TO_ORDER>
__global__ __launch_bounds__(256) void k_qb_order(RoundDesc R, const uint32_t* __restrict__ tabs, int32_t* staged,
                                                  int32_t* ordered, uint64_t staged_stride, uint64_t ordered_stride,
                                                  uint32_t tiles, uint32_t tp_shift) {
    // a tile of TP x TQ = kXT^2 elements, TP = 2^tp_shift <= kXT (rows: p)
    __shared__ int32_t tile[kXT * (kXT + 1)];
    const uint32_t TP = 1u << tp_shift, TQ = (kXT * kXT) >> tp_shift, ld = TQ + 1;
    const uint32_t f = blockIdx.y;
    const uint64_t bid = blockIdx.x / tiles;
    const uint32_t t = (uint32_t)(blockIdx.x - bid * tiles);
    const uint32_t b2 = (uint32_t)(bid % R.nb[2]), b1 = (uint32_t)((bid / R.nb[2]) % R.nb[1]),
                   b0 = (uint32_t)(bid / ((uint64_t)R.nb[2] * R.nb[1]));
    const uint32_t bb[3] = {b0, b1, b2};
    uint32_t cnt[3], pre[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        cnt[a] = tabs[R.tab_off[a] + bb[a]];
        pre[a] = tabs[R.tab_off[a] + R.nb[a] + bb[a]];
    }
    const uint64_t base = block_base(R, cnt, pre);
    const bool o120 = R.order == 0;
    const uint32_t slabs = o120 ? 1 : cnt[0], P = o120 ? cnt[0] : cnt[1];
    const uint64_t Q = o120 ? (uint64_t)cnt[1] * cnt[2] : cnt[2];
    if (slabs == 0 || P == 0 || Q == 0) return;  // a block without targets
    const uint32_t tq_n = (uint32_t)((Q + TQ - 1) / TQ), tp_n = (P + TP - 1) >> tp_shift;
    const uint32_t tq = t % tq_n, rest = t / tq_n, tp = rest % tp_n, slab = rest / tp_n;
    if (slab >= slabs) return;
    const uint32_t p0 = tp * TP;
    const uint64_t q0 = (uint64_t)tq * TQ;
    const uint32_t np = min(TP, P - p0), nq = (uint32_t)min<uint64_t>(TQ, Q - q0);
    const uint64_t slab_off = base + (uint64_t)slab * P * Q;
    int32_t* A = staged + f * staged_stride + slab_off;   // "012": p Q + q
    int32_t* B = ordered + f * ordered_stride + slab_off;  // order:  q P + p
    const uint32_t n = np * nq;
    // e / d for e < kXT^2 and d <= kXT^2 through a float reciprocal: (e + 1/2) / d
    // lies at least 1/(2d) from an integer and the float error (below
    // 2^-22 (e / d)) stays far below that, so the floor is exact
    const float inv_nq = 1.0f / (float)nq, inv_np = 1.0f / (float)np;
    auto qdiv = [](uint32_t e, float inv) { return (uint32_t)(((float)e + 0.5f) * inv); };
    constexpr uint32_t kPer = kXT * kXT / 256;  // elements per thread when the tile is full
    // every load of a thread is issued before the first LDS store waits for one
    int32_t v[kPer];
    if (TO_ORDER) {
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            const uint32_t e = threadIdx.x + 256 * k;
            if (e < n) {
                const uint32_t pl = qdiv(e, inv_nq), ql = e - pl * nq;
                v[k] = A[(uint64_t)(p0 + pl) * Q + q0 + ql];
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            const uint32_t e = threadIdx.x + 256 * k;
            if (e < n) {
                const uint32_t pl = qdiv(e, inv_nq), ql = e - pl * nq;
                tile[pl * ld + ql] = v[k];
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            const uint32_t e = threadIdx.x + 256 * k;
            if (e >= n) break;
            const uint32_t ql = qdiv(e, inv_np), pl = e - ql * np;
            B[(q0 + ql) * P + p0 + pl] = tile[pl * ld + ql];
        }
    } else {
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            const uint32_t e = threadIdx.x + 256 * k;
            if (e < n) {
                const uint32_t ql = qdiv(e, inv_np), pl = e - ql * np;
                v[k] = B[(q0 + ql) * P + p0 + pl];
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            const uint32_t e = threadIdx.x + 256 * k;
            if (e < n) {
                const uint32_t ql = qdiv(e, inv_np), pl = e - ql * np;
                tile[pl * ld + ql] = v[k];
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            const uint32_t e = threadIdx.x + 256 * k;
            if (e >= n) break;
            const uint32_t pl = qdiv(e, inv_nq), ql = e - pl * nq;
            A[(uint64_t)(p0 + pl) * Q + q0 + ql] = tile[pl * ld + ql];
        }
    }
}

}  // namespace dev

// ------------------------------------------------------------------ launchers

void launch_range(const float* x, uint64_t n, uint32_t* d_out3, hipStream_t st, const float* const* d_xs,
                  uint32_t nfields) {
    // about 2048 workgroups in all (the result words are preset by the caller)
    const uint64_t per = std::max<uint64_t>(64, 2048 / std::max<uint32_t>(nfields, 1));
    const uint64_t blocks = std::min<uint64_t>((n / 4 + 255) / 256 + 1, per);
    hipLaunchKernelGGL(dev::k_range, dim3((uint32_t)blocks, d_xs ? nfields : 1), dim3(256), 0, st, x, n, d_out3,
                       d_xs);
}
void launch_field_desc(const uint32_t* rng, const float* x, const RangeSeed& seed, FieldDesc* fd, uint32_t* res,
                       hipStream_t st) {
    hipLaunchKernelGGL(dev::k_field_desc, dim3(1), dim3(64), 0, st, rng, x, seed, fd, res);
}

void launch_pivots(bool compress, float* work, int32_t* qb, const PivotDesc& P, Partial* part, hipStream_t st,
                   const FieldSet& fs, uint32_t nfields) {
    const dim3 grid((uint32_t)((P.count + 255) / 256), compress ? nfields : 1);
    if (compress)
        hipLaunchKernelGGL(dev::k_pivots_compress, grid, dim3(256), 0, st, work, qb, P, part, fs);
    else
        hipLaunchKernelGGL(dev::k_pivots_decompress, grid, dim3(256), 0, st, work, (const int32_t*)qb, P);
}

void launch_init_acc(DevAcc* acc, hipStream_t st, uint32_t nfields) {
    hipLaunchKernelGGL(dev::k_init_acc, dim3(nfields), dim3(128), 0, st, acc);
}

void launch_init_call(DevAcc* acc, uint32_t nfields, uint32_t* words, uint64_t n, hipStream_t st) {
    if (n % 4 || reinterpret_cast<uintptr_t>(words) % 16) throw Error(ASZ_ERR_ARG, "unaligned call-start words");
    const uint64_t n4 = n / 4, blocks = std::min<uint64_t>((n4 + 255) / 256, 4096);
    hipLaunchKernelGGL(dev::k_init_call, dim3((uint32_t)(nfields + blocks)), dim3(256), 0, st, acc, nfields,
                       reinterpret_cast<uint4*>(words), n4);
}

namespace {

struct RoundArgs {
    float* w;
    int32_t* qb;
    const uint32_t* tabs;
    const float* tblk;
    const SweepSeg* segs;
    Partial* part;
    uint32_t* phist;
    uint32_t* gwin;
    FieldSet fs;
    uint32_t nf;      // fields (grid dimension)
    bool aligned16;   // every field's array starts on a 16-byte boundary
};

template <bool Q, bool DEC, int ORDER>
void row(const RoundDesc& R, const RoundArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((dev::k_round_row<Q, DEC, ORDER>), dim3((uint32_t)R.nwg, a.nf), dim3(256),
                       DEC ? 0 : R.lds_win * 4, st, R, a.w, a.qb, a.tabs, a.tblk, a.part, a.phist, a.gwin, a.fs);
}

template <bool Q, bool DEC, int ORDER>
void quad(const RoundDesc& R, const RoundArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((dev::k_round_quad<Q, DEC, ORDER>), dim3((uint32_t)R.nwg, a.nf), dim3(256),
                       DEC ? 0 : R.lds_win * 4, st, R, a.w, a.qb, a.tabs, a.tblk, a.part, a.phist, a.gwin, a.fs);
}

template <bool Q, bool DEC>
void row_order(const RoundDesc& R, const RoundArgs& a, hipStream_t st) {
    // h = 1 on rows of a multiple of four elements, every field 16-byte aligned: the quad kernel
    if (R.h == 1 && (R.D[2] & 3u) == 0 && a.aligned16) {
        if (R.order == 0) quad<Q, DEC, 0>(R, a, st);
        else if (R.order == 1) quad<Q, DEC, 1>(R, a, st);
        else quad<Q, DEC, 2>(R, a, st);
        return;
    }
    if (R.order == 0) row<Q, DEC, 0>(R, a, st);
    else if (R.order == 1) row<Q, DEC, 1>(R, a, st);
    else row<Q, DEC, 2>(R, a, st);
}

template <int AX, bool Q, bool DEC, int WM>
void sweep(const RoundDesc& R, const RoundArgs& a, hipStream_t st) {
    const uint32_t rows = 256u >> R.xsh;
    dim3 grid((R.n[2] + (1u << R.xsh) - 1) >> R.xsh, (R.n[1 - AX] + rows - 1) / rows, R.nseg * a.nf);
    hipLaunchKernelGGL((dev::k_round_sweep<AX, Q, DEC, WM>), grid, dim3(256), DEC ? 0 : R.lds_win * 4, st, R, a.w,
                       a.qb, a.tabs, a.tblk, a.segs, a.part, a.phist, a.gwin, a.fs);
}

template <bool DEC>
void dispatch(const RoundDesc& R, bool q, const RoundArgs& a, hipStream_t st) {
    if (R.kind == kRoundRow) {
        if (q) row_order<true, DEC>(R, a, st);
        else row_order<false, DEC>(R, a, st);
        return;
    }
    // sweep: write mode from the serialization order
    const int fastest = R.order == 2 ? 2 : (R.order == 0 ? 0 : 1);
    const int wm = fastest == 2 ? 0 : (fastest == (int)R.d ? 1 : 2);
    const int key = ((int)R.d * 2 + (q ? 1 : 0)) * 3 + wm;
    switch (key) {
        case 0: sweep<0, false, DEC, 0>(R, a, st); break;
        case 1: sweep<0, false, DEC, 1>(R, a, st); break;
        case 2: sweep<0, false, DEC, 2>(R, a, st); break;
        case 3: sweep<0, true, DEC, 0>(R, a, st); break;
        case 4: sweep<0, true, DEC, 1>(R, a, st); break;
        case 5: sweep<0, true, DEC, 2>(R, a, st); break;
        case 6: sweep<1, false, DEC, 0>(R, a, st); break;
        case 7: sweep<1, false, DEC, 1>(R, a, st); break;
        case 8: sweep<1, false, DEC, 2>(R, a, st); break;
        case 9: sweep<1, true, DEC, 0>(R, a, st); break;
        case 10: sweep<1, true, DEC, 1>(R, a, st); break;
        default: sweep<1, true, DEC, 2>(R, a, st); break;
    }
}

}  // namespace

void launch_round_compress(const RoundDesc& R, bool quintic, float* work, int32_t* qb, const uint32_t* tabs,
                           const float* tblk, const SweepSeg* segs, Partial* part, uint32_t* phist, uint32_t* gwin,
                           hipStream_t st, const FieldSet& fs, uint32_t nfields) {
    if (R.count == 0) return;
    const bool a16 = fs.fd ? fs.aligned16 != 0 : (reinterpret_cast<uintptr_t>(work) & 15u) == 0;
    dispatch<false>(R, quintic, RoundArgs{work, qb, tabs, tblk, segs, part, phist, gwin, fs, nfields, a16}, st);
}

bool round_needs_staging(const RoundDesc& R) {
    if (R.count < kStageMinCount || R.order == 2) return false;
    if (R.kind == kRoundRow) return true;  // "120" / "021" along DIM2
    const uint32_t fastest = R.order == 0 ? 0 : 1;
    return fastest != R.d;  // the sweep's scattered write mode
}

// Tile rows (along p, the order's run axis) for a round: the next power of two
// of its largest run count, at most kXT; the tile holds kXT^2 elements.
static uint32_t qb_order_tp_shift(const RoundDesc& R, const uint32_t cmax[3]) {
    const uint64_t P = R.order == 0 ? cmax[0] : cmax[1];
    uint32_t sh = 0;
    while ((1ull << sh) < P && (1u << sh) < dev::kXT) sh++;
    return sh;
}

uint32_t qb_order_tiles(const RoundDesc& R, const uint32_t cmax[3]) {
    const bool o120 = R.order == 0;
    const uint64_t slabs = o120 ? 1 : cmax[0], P = o120 ? cmax[0] : cmax[1];
    const uint64_t Q = o120 ? (uint64_t)cmax[1] * cmax[2] : cmax[2];
    const uint32_t sh = qb_order_tp_shift(R, cmax);
    const uint64_t TP = 1ull << sh, TQ = ((uint64_t)dev::kXT * dev::kXT) >> sh;
    return (uint32_t)(slabs * ((P + TP - 1) / TP) * ((Q + TQ - 1) / TQ));
}

void launch_qb_order(const RoundDesc& R, const uint32_t* tabs, const uint32_t cmax[3], int32_t* staged,
                     int32_t* ordered, uint64_t staged_stride, uint64_t ordered_stride, uint32_t nfields,
                     bool to_order, hipStream_t st) {
    if (R.count == 0) return;
    const uint32_t tiles = qb_order_tiles(R, cmax);
    const uint32_t sh = qb_order_tp_shift(R, cmax);
    const uint64_t nbl = (uint64_t)R.nb[0] * R.nb[1] * R.nb[2];
    if (nbl * tiles > 0x7FFFFFFFull) fail(ASZ_ERR_ARG, "field dimensions exceed the launch grid");
    const dim3 grid((uint32_t)(nbl * tiles), nfields);
    if (to_order)
        hipLaunchKernelGGL(dev::k_qb_order<true>, grid, dim3(256), 0, st, R, tabs, staged, ordered, staged_stride,
                           ordered_stride, tiles, sh);
    else
        hipLaunchKernelGGL(dev::k_qb_order<false>, grid, dim3(256), 0, st, R, tabs, staged, ordered, staged_stride,
                           ordered_stride, tiles, sh);
}

void launch_round_decompress(const RoundDesc& R, bool quintic, float* out, const int32_t* qb, const uint32_t* tabs,
                             const float* tblk, const SweepSeg* segs, hipStream_t st) {
    if (R.count == 0) return;
    dispatch<true>(R, quintic,
                   RoundArgs{out, const_cast<int32_t*>(qb), tabs, tblk, segs, nullptr, nullptr, nullptr, FieldSet{}, 1,
                             (reinterpret_cast<uintptr_t>(out) & 15u) == 0},
                   st);
}

}  // namespace asz
