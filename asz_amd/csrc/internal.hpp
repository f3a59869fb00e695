// Internal declarations shared by the host orchestration (codec.cpp) and the
// HIP kernels.  Field layout in HBM: one float32 array, C order (DIM2 fastest),
// the same as the reference's data / compressed_data vectors.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/asz_gpu.h"

namespace asz {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

// ---------------------------------------------------------------- geometry

// Plan-derived geometry (generate_config.tpp:51-68).
struct Geometry {
    uint64_t D[3];
    int L;
    uint8_t axis[ASZ_MAX_LAYERS], order[ASZ_MAX_LAYERS];
    uint64_t pivot_stride[3], block_shape[3], block_num[3];
    uint64_t nblocks() const { return block_num[0] * block_num[1] * block_num[2]; }
    uint64_t N() const { return D[0] * D[1] * D[2]; }
};
Geometry make_geometry(const uint64_t dims[3], const asz_plan& plan, uint64_t bpr);

// Exact unsigned division by a constant 1 <= d < 2^31: n / d == (n * m) >> sh
// for every n < 2^31, with sh = 32 + ceil(log2 d) and m = ceil(2^sh / d)
// (m < 2^33, so n * m < 2^64; the rounding error n(m - 2^sh/d)/2^sh stays
// below 1/d).
struct FastDiv {
    uint64_t m;
    uint32_t sh, d;
};
inline FastDiv make_fastdiv(uint32_t d) {
    uint32_t l = 0;
    while ((1ull << l) < d) l++;
    FastDiv f;
    f.sh = 32 + l;
    f.m = ((1ull << f.sh) + d - 1) / d;
    f.d = d;
    return f;
}

// Per-axis block tables of one round.  cnt[b] = targets of block b along the
// axis, pre[b] = exclusive prefix of cnt (asz_param.hpp:105-115 per axis).
// The QB base of block (b0,b1,b2) in the round's array is
// pre0[b0]*C1*C2 + cnt0[b0]*(pre1[b1]*C2 + cnt1[b1]*pre2[b2]), which is the
// block-major concatenation of compress.tpp:61-62.
struct RoundDesc {
    uint32_t D[3];
    uint32_t s[3];      // strides of this round, before stride[d] is halved
    uint32_t h;         // s[d] / 2
    uint32_t n[3];      // global target counts per axis
    uint32_t B[3];      // block shape
    uint32_t nb[3];     // block counts
    uint32_t C[3];      // per-axis target totals (sum of cnt)
    uint32_t tab_off[3];// offsets of the cnt/pre tables of each axis in the table buffer
    uint32_t d, order;
    uint64_t dist;      // h * prod_{a > d} D[a]
    uint64_t qb_off;    // offset of this round's array in the QB buffer (a multiple of kSymAlign)
    uint64_t count;     // targets of this round (= C0*C1*C2)
    float eb;           // layer error bound
    float layer_slope;  // layer default slope (compress.tpp:195)
    double inv_half_eb; // 0.5 / eb           (axis-0 quantizer form)
    double eb_d;        // (double) eb         (axes-1/2 quantizer form)
    double neg_two_eb;  // -(2.0 * (double) eb) (reconstruction)
    uint32_t kind;      // kRoundRow or kRoundSweep
    uint32_t seg_off;   // first sweep segment of this round
    uint32_t nseg;      // sweep segments of this round
    uint32_t rpw;       // row kernel: lattice rows per workgroup (multiple of kRowsPerWave * 4)
    uint64_t part_off;  // first per-workgroup partial of this round
    uint64_t nwg;       // workgroups of this round's kernel
    uint64_t hist_off;    // first word of this round's histogram copies (hist_copies x lds_win words)
    uint32_t hist_copies; // copies of the LDS window the workgroups add into (workgroup w: copy w % hist_copies)
    uint32_t lds_win;     // bins of each workgroup's LDS histogram (power of 2, >= kLdsWinMin)
    uint32_t layer;     // index of this round in the plan
    uint32_t xsh;       // sweep: log2 of a workgroup's lanes along DIM2 (6..8); 256 >> xsh lattice rows each
    FastDiv tpb;        // targets per full block along the round axis, B[d] / (2h)
};

// Multi-field calls: F fields of one shape compressed with one plan, each
// per-round kernel covering all of them with the field index as a grid
// dimension.  The plan tables are shared; what depends on a field's value
// range (pivot and layer error bounds, layer slopes, the range itself) comes
// from its FieldDesc, and every per-call buffer holds one slice per field at
// the FieldSet's stride (in elements of that buffer).  fd == nullptr: a
// single field, described by the kernel's own arguments.
struct FieldRound {
    double inv_half_eb, eb_d, neg_two_eb;  // as RoundDesc
    float eb, layer_slope;
};
struct FieldDesc {
    float* work;               // the field, rewritten with its reconstruction
    double inv_half_pe, two_pe;  // pivot quantizer (as PivotDesc)
    float range;               // max_val - min_val
    uint32_t pad;
    FieldRound r[ASZ_MAX_LAYERS];
};
struct FieldSet {
    const FieldDesc* fd;
    uint64_t sym;    // t array + QB arrays (int32)
    uint64_t tblk;   // per-block float t
    uint64_t part;   // round partials
    uint64_t phist;  // per-workgroup histogram windows (words)
    uint64_t fpart;  // pivot partials then fold partials
    uint64_t ghist;  // global histogram windows (words)
    uint64_t hres;   // host result slots (words)
    uint32_t aligned16;  // with fd set: every field's array starts on a 16-byte boundary
    uint32_t pad;
};

constexpr uint32_t kRoundRow = 0;     // round axis 2: waves walk lattice rows along DIM2
constexpr uint32_t kRoundSweep = 1;   // register-window sweep along the round axis (axis 0 / 1)
constexpr uint32_t kRoundBlock = 2;   // block round: a workgroup per block selects its t and predicts its targets
// A leading round runs as a block round when it has at most kCoarseTargets
// targets per block (one 256-thread workgroup per block does its selection
// and its targets in one launch).
constexpr uint64_t kCoarseTargets = 16384;
constexpr uint32_t kSweepSeg = 24;    // most targets per sweep segment (register window length)
constexpr uint32_t kRowsPerWave = 4;  // row kernel: lattice rows a wave processes together

// QB histograms.  Every round kernel bins its QBs into a per-workgroup LDS
// histogram of lds_win symbols around 0.  At the end of the workgroup its
// nonzero bins are added with atomics into one of the round's hist_copies
// copies of that window (workgroup w into copy w % hist_copies, so that at
// most kWgPerCopy-odd workgroups share an address); the rare QB outside the
// LDS histogram goes to the array's global window of kHistGlobal bins
// (centre kHistCentre) directly.  The fold then sums the copies, a few KB
// per round whatever its workgroup count.  No global atomic sits in a round
// kernel's main loop: gfx9 counts them in vmcnt with the loads, so a slow one
// would stall every later load.  Rounds with few workgroups (the coarse
// ones, whose QBs spread widest) get the widest LDS windows: the window
// doubles while 2 x window x workgroups stays within a quarter word per
// target, clamped to [kHistWidthMin, kHistWidthWords].  Arrays whose QB
// range leaves the global window are histogrammed again by a dedicated pass.
constexpr int kHistWin = 128;                   // window of the t / pivot arrays
constexpr uint32_t kHistWinMin = 128, kHistWinMax = 4096;
constexpr uint32_t kLdsWinMin = 1024;  // 4 KB: the axis-0/1 sweep then fits five workgroups per CU
constexpr uint64_t kHistWidthWords = 1 << 20, kHistWidthMin = 1 << 15;
constexpr uint32_t kHistCopies = 16, kWgPerCopy = 32;
constexpr uint32_t kHistGlobal = 65536;
constexpr int32_t kHistCentre = 32768;

// Each array's global window is followed by a tail that carries what the host
// needs to build that array's codebook and to report errors, so that one
// copy per array brings everything back (while later rounds still run).
struct HistTail {
    int32_t qmin, qmax;   // QB range of the array (empty: qmin > qmax)
    uint32_t fallbacks;   // t array only: blocks that fell back to t = 4/16
    uint32_t pad;
    double max_err, sq;   // pivots and rounds: max |x - x'| and sum (x - x')^2
};
constexpr uint64_t kHistTailWords = sizeof(HistTail) / 4;
constexpr uint64_t kHistStride = kHistGlobal + kHistTailWords;  // words per array

// Host-visible result slot of one array (pinned, device-written): the
// kernel that finishes the array writes the tail and the bins of [qmin, qmax]
// (when inside the global window), then, after a system-scope fence, the
// call's epoch into the flag.  The host builds the codebook once it sees it.
constexpr uint64_t kHostSlotHead = 4;                                  // flag + pad (the tail follows)
constexpr uint64_t kHostSlotBins = kHostSlotHead + kHistTailWords;     // first bin word
constexpr uint64_t kHostSlotWords = kHostSlotBins + kHistGlobal;
constexpr uint64_t kCodebookThreads = 4;  // host threads building codebooks

// A sweep segment: targets k0 .. k0+cnt-1 (global indices along the round
// axis) inside block b along that axis.
struct SweepSeg {
    uint32_t k0, cnt, b, pad;
};

// Per-workgroup reduction result of a round kernel.
struct Partial {
    double max_err, sq;
    int32_t qmin, qmax;
};

// Arrays of the compression symbol buffer (t array, pivots, each round)
// start at multiples of kSymAlign elements.
constexpr uint64_t kSymAlign = 64;
inline uint64_t sym_align(uint64_t n) { return (n + kSymAlign - 1) / kSymAlign * kSymAlign; }

struct Plan {
    Geometry g;
    std::vector<RoundDesc> rounds;
    std::vector<uint32_t> tables;  // per round, per axis: cnt[nb] then pre[nb]
    std::vector<uint32_t> ss;      // t-selection sampling stride per (round, block): r * nblocks + b
    std::vector<uint32_t> ns;      // t-selection sample count per (round, block), 0 for fallback blocks
    std::vector<SweepSeg> segs;    // sweep segments of every sweep round
    uint64_t total_partials = 0;   // per-workgroup partials over all rounds
    uint64_t total_hist_words = 0; // the rounds' histogram copies (words)
    uint64_t pivot_count;
    uint64_t qb_total;             // QB buffer elements: pivots and rounds, each array aligned (kSymAlign)
    uint32_t coarse = 0;           // leading rounds of kind kRoundBlock
    uint32_t pn[3];                // pivot lattice counts per axis
    float pivot_eb;
};
// Tile shape of one round's kernel, chosen by the autotuner (0: the default
// rule): sweep rounds' segment length (targets a lane walks along the round
// axis, at most kSweepSeg) and lanes along DIM2 (1 << xsh, 6 <= xsh <= 8);
// row rounds' lattice rows per workgroup (rounded up to a multiple of
// 4 kRowsPerWave).  The QB layout and every value are the same for any tile
// shape.
struct RoundTune {
    uint8_t seg, xsh;
    uint16_t rpw;
};
// coarse_max: the most leading rounds that may run as block rounds (those
// with at most kCoarseTargets targets per block).  tune: one entry per round
// or nullptr.  tsel_strides: fill ss / ns, the t selection's per-block sample
// strides (compression only; decompression leaves them empty).
Plan make_plan(const Geometry& g, float abs_eb, float pivot_eb_ratio, float eb_ratio, float slope_ratio,
               float default_slope, uint64_t h1_num, uint32_t coarse_max = 0, const RoundTune* tune = nullptr,
               bool tsel_strides = true);
// The per-call scalars of a plan (pivot and layer error bounds, layer slopes);
// the tables do not depend on them.
void set_plan_bounds(Plan& P, float abs_eb, float pivot_eb_ratio, float eb_ratio, float slope_ratio,
                     float default_slope);

// Accumulators written by the pivot / round kernels (device memory).
struct ArrayAcc {
    int32_t qmin, qmax;  // QB range of one array, feeds the Huffman histogram
    uint32_t pad[2];
};
struct ErrAcc {
    double max_err;      // max |x - x'|
    double sq_sum;       // sum (x - x')^2
};
struct DevAcc {
    ErrAcc err;                        // pivots
    uint32_t fallbacks;                // t selection: blocks that fell back to t = 4/16
    uint32_t pad;
    uint32_t fold_count[ASZ_MAX_LAYERS];  // k_fold_rounds: workgroups done per round (reset by the last one)
    ArrayAcc arr[ASZ_MAX_LAYERS + 2];  // [0] t array, [1] pivots, [2 + r] round r
    ErrAcc rerr[ASZ_MAX_LAYERS];       // round r
};

// ---------------------------------------------------------------- fp helpers (host)

// Reference process mode is FTZ/DAZ (crtfastmath.o from the -Ofast link).
inline float daz(float x) {
    return (x != 0.0f && x > -1.17549435e-38f && x < 1.17549435e-38f) ? (x < 0 ? -0.0f : 0.0f) : x;
}

// ---------------------------------------------------------------- Huffman (host)

struct HuffArray {
    // inputs
    uint64_t count = 0;          // symbols in the array
    int64_t offset = 0;          // min symbol
    uint64_t range = 0;          // max - min + 1
    bool is_int64 = false;
    // outputs of build_codebook
    std::vector<uint8_t> header; // type byte, offset, n, maxval, tree bits (or the single-symbol form)
    std::vector<uint64_t> code;  // per symbol index: code (LSB = first bit)
    std::vector<uint8_t> len;    // per symbol index: length, 0 = absent
    uint64_t payload_bits = 0;
    bool single = false;         // maxval == 1 path (huffman.hpp:193-210, 342-347)
    // encoder tables (pack_book): code in bits 0..55, length in bits 56..63;
    // and the code's low 32 bits (the whole code when max_len <= 32)
    std::vector<uint64_t> book;
    std::vector<uint32_t> code32;
    int max_len = 0;
};
// Build the reference tree codebook from a dense histogram (huffman.hpp:263-320, 650-691).
void build_codebook(HuffArray& a, const uint32_t* hist);
// The encoder's table of a built codebook (fails on codes over 56 bits).
void pack_book(HuffArray& a);

// Decode tables built from a serialized tree (huffman.hpp:693-783).
struct DecodeInfo {
    int64_t offset = 0;
    uint64_t nleaf = 0;
    bool single = false;
    uint64_t single_count = 0;
    uint64_t payload_bits = 0;
    uint64_t payload_off = 0;    // byte offset of the payload in the stream
    uint64_t end_off = 0;        // byte offset after this array
    int max_len = 0;
    std::vector<uint32_t> table; // multi-level lookup, see huffman_host.cpp
};
// Parse one HUF<T> block at stream offset off. Throws on malformed input.
DecodeInfo parse_huffman(const uint8_t* s, uint64_t n, uint64_t off, uint64_t count);

// Lookup-table geometry for the GPU decoder.
constexpr int kRootBits = 12;     // root table (LDS)
constexpr int kSubTableBits = 8;  // largest subtable below it (global memory)

}  // namespace asz
