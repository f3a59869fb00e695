// Host-callable launchers of the HIP kernels (defined in the .hip files).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "internal.hpp"

namespace asz {

struct PivotDesc {
    uint32_t D[3];
    uint32_t ps[3];   // pivot stride
    uint32_t pn[3];   // pivot lattice counts
    uint64_t count;   // pn[0] * pn[1] * pn[2]
    double inv_half_pe;  // 0.5 / (double) pivot_eb
    double two_pe;       // (double) pivot_eb * 2
};

// t-selection inputs for one round (compress.tpp:90-144).
struct TselDesc {
    RoundDesc R;
    uint32_t L, r;
    float stat_eb_ratio;
    float range;            // max_val - min_val (float)
    float t_precision;
    float t_min, t_max;     // gradient-tuned form: the projection interval
};

// AB-key sampling over a whole axis for the default slope (get_default_slope.tpp:15-25).
struct SlopeSampleDesc {
    uint32_t D[3];
    uint32_t lo[3], hi[3], step[3];  // step = stride * sampling stride
    uint32_t cnt[3];                 // samples per axis
    uint64_t dist;
    float neg_2_stat_eb;
};

// The default slope's samples of several fields of one shape in one launch.
constexpr uint32_t kSlopeMaxFields = 16;
struct SlopeFieldsDesc {
    SlopeSampleDesc S[3];           // one per axis (neg_2_stat_eb unused)
    const float* f[kSlopeMaxFields];
    float scale[kSlopeMaxFields];   // 1 / (abs_eb stat_eb_ratio) per field
    uint64_t off[3];                // key offsets of the axes in one field's slice
    uint64_t per;                   // keys per field
    uint32_t nf;
};

// Results: kRangeWords words at d_out (preset: min key ~0, max key 0, flag 0,
// first zero index ~0, last 0).  d_xs (device array of nfields pointers,
// every field n elements): one launch for all of them, field f's words at
// d_out + kRangeWords f.
void launch_range(const float* x, uint64_t n, uint32_t* d_out, hipStream_t st, const float* const* d_xs = nullptr,
                  uint32_t nfields = 1);
// Range scan result words per field (k_range): min / max keys, non-finite
// flag, first / last DAZ-zero index.
constexpr uint32_t kRangeWords = 8;
// Host-side inputs of k_field_desc: what does not depend on the range.
struct RangeSeed {
    float* work;
    float rel_eb, pivot_eb_ratio;
    uint32_t L;
    double eb_factor[ASZ_MAX_LAYERS];   // (double) eb_ratio^(L-1-r), libm pow
    float layer_slope[ASZ_MAX_LAYERS];  // compress.tpp:195
};
// The error-bound scalars of one field from its range words (rng, preset and
// filled by launch_range), into fd; res: min bits, max bits, status.
void launch_field_desc(const uint32_t* rng, const float* x, const RangeSeed& seed, FieldDesc* fd, uint32_t* res,
                       hipStream_t st);
// compress: one Partial per 256 pivots into part[0 ..]; decompress: part unused.
// With fs.fd set, nfields fields (compress only; the pointers are field 0's slices).
void launch_pivots(bool compress, float* work, int32_t* qb, const PivotDesc& P, Partial* part, hipStream_t st,
                   const FieldSet& fs = FieldSet{}, uint32_t nfields = 1);
// part / phist: this round's per-workgroup partials and histogram windows (at
// R.part_off); gwin: the round array's global histogram window.  With fs.fd
// set, one launch covers nfields fields.
void launch_round_compress(const RoundDesc& R, bool quintic, float* work, int32_t* qb, const uint32_t* tabs,
                           const float* tblk, const SweepSeg* segs, Partial* part, uint32_t* phist, uint32_t* gwin,
                           hipStream_t st, const FieldSet& fs = FieldSet{}, uint32_t nfields = 1);
void launch_round_decompress(const RoundDesc& R, bool quintic, float* out, const int32_t* qb, const uint32_t* tabs,
                             const float* tblk, const SweepSeg* segs, hipStream_t st);
// Rounds of at least kStageMinCount targets whose serialization order runs
// along neither DIM2 nor the round axis are compressed with their QBs in
// "012" order into a staging array (R.order = 2, R.qb_off = 0) and moved into
// their order by launch_qb_order (to_order); decompression moves them back
// first.  staged / ordered: the round array's first QB in each layout (field
// f at f * stride); cmax: the largest per-axis block counts of the round.
constexpr uint64_t kStageMinCount = 1ull << 18;
bool round_needs_staging(const RoundDesc& R);
void launch_qb_order(const RoundDesc& R, const uint32_t* tabs, const uint32_t cmax[3], int32_t* staged,
                     int32_t* ordered, uint64_t staged_stride, uint64_t ordered_stride, uint32_t nfields,
                     bool to_order, hipStream_t st);
// Resets the accumulators: errors and fallbacks to 0, every QB range to empty.
void launch_init_acc(DevAcc* acc, hipStream_t st, uint32_t nfields = 1);
// launch_init_acc and the clearing of n words (a multiple of 4, 16-byte
// aligned) in one launch.
void launch_init_call(DevAcc* acc, uint32_t nfields, uint32_t* words, uint64_t n, hipStream_t st);

// grad: gradient-tuned candidates (projected Newton on the sampled residual
// energy, snapped to the t grid) instead of the five block candidates.
// Fallback blocks are counted in acc->fallbacks.  With fs.fd set, one launch
// covers nfields fields (acc: one DevAcc per field).
void launch_tsel(const TselDesc& T, const float* work, const uint32_t* ss, float* tblk, int32_t* t_int, DevAcc* acc,
                 uint64_t nblocks, bool grad, hipStream_t st, const FieldSet& fs = FieldSet{},
                 uint32_t nfields = 1);
void launch_slope_sample(const SlopeSampleDesc& S, const float* field, int64_t* keys, hipStream_t st);
void launch_slope_sample_fields(const SlopeFieldsDesc& B, int64_t* keys, hipStream_t st);
// launch_tsel for round T.r together with launch_fold_rounds for rounds
// r0 .. r0 + nr - 1 (finished before T.r), in one launch.
struct FoldDesc;
void launch_fold_tsel(const TselDesc& T, const float* work, const uint32_t* ss, float* tblk, int32_t* t_int,
                      DevAcc* acc, uint64_t nblocks, bool grad, const FoldDesc* descs, uint32_t r0, uint32_t nr,
                      uint32_t max_groups, const Partial* part, const uint32_t* phist, Partial* fpart,
                      uint32_t* ghist, uint32_t* hres, uint32_t epoch, hipStream_t st, const FieldSet& fs = FieldSet{},
                      uint32_t nfields = 1);

// Block round (coarse.hip): round T.R of kind kRoundBlock with its t
// selection, one workgroup per (block, field).  tsel: 0 takes each block's t
// from tblk (the round's slice), 1 selects it with the five-candidate scan,
// 2 with the gradient-tuned form (the t array sym and tblk receive it).
// gwin: the round array's global window; the pointers are field 0's (with
// fs.fd set, field f's slices at the FieldSet strides).
void launch_round_block(const TselDesc& T, uint32_t tsel, bool quintic, float* work, int32_t* sym, int32_t* qb,
                        const uint32_t* tabs, const uint32_t* ss, float* tblk, Partial* part, uint32_t* phist,
                        uint32_t* gwin, DevAcc* acc, hipStream_t st, const FieldSet& fs, uint32_t nfields);

// Residual energy of round R per block (forward: E[b], FP64) or its adjoint
// (g[b * np + j] = gE[b] * dE_b/dtheta_j, FP32), np = 1 (cubic t) or 2 (quintic a, b).
void launch_energy(const RoundDesc& R, uint64_t nblocks, bool quintic, bool adjoint, const float* work,
                   const float* theta, const float* gE, double* E, float* g, hipStream_t st);
void launch_tsel_keys(const TselDesc& T, const float* work, const uint32_t* ss, const uint64_t* key_off,
                      int64_t* keys, uint64_t nblocks, hipStream_t st);

// Huffman kernels (huffman.hip)
struct HuffArrayDev {
    uint64_t src_off;     // first symbol in the concatenated symbol buffer
    uint64_t count;       // symbols
    int64_t offset;       // min symbol
    uint64_t range;       // max - min + 1
    uint64_t book_off;    // offset of the codebook in the book buffer
    uint64_t hist_off;    // offset of the exact-range histogram (k_hist)
    uint64_t hchunk0;     // first k_hist chunk of this array
    uint64_t chunk0;      // first encode chunk of this array
    uint64_t nchunk;      // encode chunks (0 for an empty or single-symbol array)
    uint64_t payload_bit; // absolute bit position of the payload in the stream
    uint32_t lds_book;    // codebook copied to LDS by the encoder (range <= kLdsBook)
    uint32_t pad;
};
constexpr uint32_t kHistChunk = 16384;   // symbols per histogram workgroup
constexpr uint32_t kEncChunk = 4096;     // symbols per encode workgroup (256 threads x 16)
constexpr uint32_t kLdsBook = 1024;      // LDS codebook entries of the encoder

void launch_hist(const int32_t* sym, const HuffArrayDev* arrays, const uint32_t* chunk_arr, uint64_t nchunks,
                 uint32_t* hist, hipStream_t st);
// nfields > 1: field f's arrays s_stride symbols and its windows g_stride words
// after field 0's.
void launch_hist_small(const int32_t* s0, uint64_t n0, const int32_t* s1, uint64_t n1, uint32_t* g0, uint32_t* g1,
                       hipStream_t st, uint32_t nfields = 1, uint64_t s_stride = 0, uint64_t g_stride = 0);
// Fold of a round's per-workgroup partials (error, QB range) and histogram
// windows into the round array's global window (kHistStride words per array
// in ghist: bins, then the HistTail) and acc->arr[2 + r].  One descriptor per
// round, built once per plan (make_fold_descs) and kept in device memory.
struct FoldDesc {
    uint64_t part_off, nwg, hist_off, fpart_off;
    uint32_t win, copies, slices, pgroups, groups, pad;
};
// A round's fold workgroups: win / 256 bin slices, each summing its 256 bins
// over the round's histogram copies into the global window, then
// ceil(nwg / kFoldPartials) groups folding the workgroups' Partials.
constexpr uint32_t kFoldPartials = 1024;
inline void make_fold_descs(const std::vector<RoundDesc>& rounds, std::vector<FoldDesc>& out, uint64_t& fpart_total) {
    out.clear();
    fpart_total = 0;
    for (const RoundDesc& R : rounds) {
        FoldDesc F{R.part_off, R.nwg, R.hist_off, fpart_total, R.lds_win, R.hist_copies, 0, 0, 0, 0};
        if (R.nwg) {
            F.slices = R.lds_win / 256;
            F.pgroups = (uint32_t)((R.nwg + kFoldPartials - 1) / kFoldPartials);
            F.groups = F.slices + F.pgroups;
        }
        fpart_total += F.pgroups;
        out.push_back(F);
    }
}
// Rounds r0 .. r0 + nr - 1 in one launch; max_groups = their largest groups.
// hres: device address of the pinned host result slots (kHostSlotWords per array).
// With fs.fd set, one launch covers nfields fields (acc: one DevAcc per field).
void launch_fold_rounds(const FoldDesc* descs, uint32_t r0, uint32_t nr, uint32_t max_groups, const Partial* part,
                        const uint32_t* phist, Partial* fpart, DevAcc* acc, uint32_t* ghist, uint32_t* hres,
                        uint32_t epoch, hipStream_t st, const FieldSet& fs = FieldSet{}, uint32_t nfields = 1);
// Finishes the t array (k = 0, range and fallbacks from acc) or the pivot
// array (k = 1, folding the pivot kernel's nparts Partials) into its host slot.
void launch_hist_tail(const DevAcc* acc, uint32_t k, const Partial* part, uint64_t nparts, const uint32_t* gw,
                      uint32_t* hres, uint32_t epoch, hipStream_t st, uint32_t nfields = 1, uint64_t part_stride = 0,
                      uint64_t g_stride = 0, uint64_t h_stride = 0);
// Huffman encoder of the arrays' chunks into the stream buffer (header bytes
// are scattered in after it has run, with the same stream).  Every code at
// most 32 bits (max_len <= 32): a pack pass codes each chunk from bit 0 into
// its slot of the staging buffer (kEncChunk words per chunk: stage holds
// nchunks * kEncChunk words) and counts its bits into chunk_bits, a scan
// turns the counts into chunk_off (and zeroes the payload words that chunks
// share), and a place pass copies each slot to its bit position.  Longer
// codes: a bit-count pass, the scan, and an encode pass that ORs the codes
// straight into the stream (stage unused).  The codebook of array k starts at
// its book_off in three parallel tables: book (code in bits 0..55, length in
// bits 56..63), code32 (the code, when max_len <= 32) and len8 (the length).
// chunk_bits / chunk_off: nchunks entries.
constexpr uint64_t kEncGroups = 2048;    // most workgroups of a persistent encoder pass
void launch_encode(const int32_t* sym, const HuffArrayDev* arrays, uint32_t narr, uint64_t nchunks,
                   const uint64_t* book, const uint32_t* code32, const uint8_t* len8, uint32_t* chunk_bits,
                   uint64_t* chunk_off, uint32_t* stream_words, int max_len, uint32_t* stage, hipStream_t st);
// The encoder's two halves for codes of at most 32 bits, so that arrays can be
// packed as their codebooks arrive: the pack pass over chunks [c_begin,
// c_end) of the listed arrays (any subset whose chunk numbers cover that
// range), then, once every chunk is packed, the scan and place passes over all
// nchunks chunks of the full array list.
void launch_encode_pack(const int32_t* sym, const HuffArrayDev* arrays, uint32_t narr, uint64_t c_begin,
                        uint64_t c_end, const uint32_t* code32, const uint8_t* len8, uint32_t* chunk_bits,
                        uint32_t* stage, hipStream_t st);
void launch_encode_place(const HuffArrayDev* arrays, uint32_t narr, uint64_t nchunks, uint32_t* chunk_bits,
                         uint64_t* chunk_off, const uint32_t* stage, uint32_t* stream_words, hipStream_t st);

// Decoder (self-synchronising parallel decode of one bit stream per array).
struct DecArrayDev {
    uint64_t payload_bit;  // absolute bit position of the payload in the stream buffer
    uint64_t bits;         // payload bits
    uint64_t count;        // symbols to produce
    uint64_t out_off;      // first output slot
    uint64_t table_off;    // decode table offset (words)
    uint64_t sub0;         // first subsequence index of this array (= wg0 * 256)
    uint64_t nsub;         // subsequences
    uint64_t wg0, nwg;     // decoder workgroups of this array
    int64_t offset;        // symbol offset
    uint32_t root_bits;
    uint32_t max_len;
    uint32_t lead;         // bits the sync pass starts a subsequence's parse early (<= kMaxSyncLead)
    uint32_t pair_off;     // the sync pass's root table (decode table words), kNoPairTable: none
};
// The sync pass's root table: the root table with every entry whose root-bit
// pattern holds two whole codes replaced by kPairFlag | len1 << 8 | len1 +
// len2 (the sync pass counts codewords and needs no symbols).
constexpr uint32_t kPairFlag = 0x40000000u, kNoPairTable = 0xFFFFFFFFu;
constexpr uint32_t kSubBits = 512;  // bits per subsequence
// The sync lead: about 48 codewords of the array's mean length, by which a
// parse from an arbitrary bit has almost always fallen into step with the
// true codeword boundaries (on the 512^3 stream: 0.07 % of the subsequences
// of its largest array start out of step at 96 bits, under 0.001 % at 159;
// of its 6.2-bit array 13.5 % at 96 bits, 0.3 % at 298), so that few
// workgroups need a fix pass.
constexpr uint32_t kMinSyncLead = 64, kMaxSyncLead = kSubBits;
inline uint32_t sync_lead(uint64_t payload_bits, uint64_t count) {
    const uint64_t l = count ? (48 * payload_bits + count - 1) / count : kMaxSyncLead;
    return (uint32_t)std::min<uint64_t>(kMaxSyncLead, std::max<uint64_t>(kMinSyncLead, l));
}

// wg_arr[w] = array of decoder workgroup w (each workgroup holds 256 subsequences of one array)
void launch_dec_sync(const uint32_t* words, const DecArrayDev* arrays, const uint32_t* wg_arr, uint64_t nwg,
                     const uint32_t* tables, uint64_t* start, uint64_t* exit_pos, uint32_t* count, hipStream_t st);
void launch_dec_fix(const uint32_t* words, const DecArrayDev* arrays, const uint32_t* wg_arr, uint64_t nwg,
                    const uint32_t* tables, uint64_t* start, const uint64_t* exit_in, uint64_t* exit_out,
                    uint32_t* count, unsigned int* changed, hipStream_t st);
void launch_dec_offsets(const DecArrayDev* arrays, uint32_t narrays, const uint32_t* wg_arr, uint64_t nwg,
                        const uint32_t* count, uint64_t* wg_sum, uint64_t* wg_off, unsigned int* err, hipStream_t st);
void launch_dec_write(const uint32_t* words, const DecArrayDev* arrays, const uint32_t* wg_arr, uint64_t nwg,
                      const uint32_t* tables, const uint64_t* start, const uint32_t* count, const uint64_t* wg_off,
                      int32_t* out, unsigned int* err, hipStream_t st);

}  // namespace asz
