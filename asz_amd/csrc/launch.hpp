// Host-callable launchers of the HIP kernels (defined in the .hip files).
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace asz {

struct PivotDesc {
    uint32_t D[3];
    uint32_t ps[3];   // pivot stride
    uint32_t pn[3];   // pivot lattice counts
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
};

// AB-key sampling over a whole axis for the default slope (get_default_slope.tpp:15-25).
struct SlopeSampleDesc {
    uint32_t D[3];
    uint32_t lo[3], hi[3], step[3];  // step = stride * sampling stride
    uint32_t cnt[3];                 // samples per axis
    uint64_t dist;
    float neg_2_stat_eb;
};

void launch_range(const float* x, uint64_t n, uint32_t* d_out3, hipStream_t st);
void launch_zero_index(const float* x, uint64_t n, unsigned long long* d_out2, hipStream_t st);
void launch_pivots(bool compress, float* work, int32_t* qb, const PivotDesc& P, void* eacc, void* aacc,
                   hipStream_t st);
// part / phist: this round's per-workgroup partials and histogram windows (at
// R.part_off); gwin: the round array's global histogram window.
void launch_round_compress(const RoundDesc& R, bool quintic, float* work, int32_t* qb, const uint32_t* tabs,
                           const float* tblk, const SweepSeg* segs, Partial* part, uint32_t* phist, uint32_t* gwin,
                           hipStream_t st);
void launch_round_decompress(const RoundDesc& R, bool quintic, float* out, const int32_t* qb, const uint32_t* tabs,
                             const float* tblk, const SweepSeg* segs, hipStream_t st);
// Resets the accumulators: errors and fallbacks to 0, every QB range to empty.
void launch_init_acc(DevAcc* acc, hipStream_t st);
// ranges: 4 words per round (see launch_hist_reduce); writes DevAcc::arr[2 + r] and rerr[r]
void launch_reduce_partials(const Partial* part, const uint64_t* ranges, uint32_t rounds, DevAcc* acc,
                            hipStream_t st);

void launch_tsel(const TselDesc& T, const float* work, const uint32_t* ss, float* tblk, int32_t* t_int,
                 unsigned int* fallbacks, uint64_t nblocks, hipStream_t st);
void launch_slope_sample(const SlopeSampleDesc& S, const float* field, int64_t* keys, hipStream_t st);
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
    uint32_t lds_book;    // codebook copied to LDS by the encoder (range <= kLdsBook, codes <= 24 bits)
    uint32_t pad;
};
constexpr uint32_t kHistChunk = 16384;   // symbols per histogram workgroup
constexpr uint32_t kEncChunk = 4096;     // symbols per encode workgroup (256 threads x 16)
constexpr uint32_t kLdsBook = 4096;      // LDS codebook entries of the encoder

void launch_hist(const int32_t* sym, const HuffArrayDev* arrays, const uint32_t* chunk_arr, uint64_t nchunks,
                 uint32_t* hist, hipStream_t st);
void launch_hist_small(const int32_t* s0, uint64_t n0, const int32_t* s1, uint64_t n1, uint32_t* g0, uint32_t* g1,
                       hipStream_t st);
// ranges: (first partial, workgroups, first histogram word, window bins) per round
void launch_hist_reduce(const uint32_t* phist, const uint64_t* ranges, uint32_t rounds, uint32_t* ghist,
                        hipStream_t st);
void launch_hist_compact(const DevAcc* acc, uint32_t narr, const uint32_t* ghist, uint32_t* out, hipStream_t st);
// slot_words: scratch words per chunk (>= kEncChunk * max_len / 32 + 2); with
// max_len > 32 the slots must be zeroed first.
void launch_encode_chunks(const int32_t* sym, const HuffArrayDev* arrays, uint32_t narr, uint64_t nchunks,
                          const uint64_t* book, uint32_t* scratch, uint64_t slot_words, uint32_t* chunk_bits,
                          int max_len, hipStream_t st);
void launch_chunk_scan(const HuffArrayDev* arrays, uint32_t narrays, const uint32_t* chunk_bits, uint64_t* chunk_off,
                       hipStream_t st);
void launch_compact_bits(const uint32_t* scratch, uint64_t slot_words, const uint32_t* chunk_bits,
                         const uint64_t* chunk_off, uint64_t nchunks, uint32_t* stream_words, hipStream_t st);

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
};
constexpr uint32_t kSubBits = 512;  // bits per subsequence

// wg_arr[w] = array of decoder workgroup w (each workgroup holds 256 subsequences of one array)
void launch_dec_sync(const uint32_t* words, const DecArrayDev* arrays, const uint32_t* wg_arr, uint64_t nwg,
                     const uint32_t* tables, uint64_t* start, uint64_t* exit_pos, uint32_t* count, hipStream_t st);
void launch_dec_fix(const uint32_t* words, const DecArrayDev* arrays, const uint32_t* wg_arr, uint64_t nwg,
                    const uint32_t* tables, uint64_t* start, const uint64_t* exit_in, uint64_t* exit_out,
                    uint32_t* count, unsigned int* changed, hipStream_t st);
void launch_dec_offsets(const DecArrayDev* arrays, uint32_t narrays, uint64_t nwg, const uint32_t* count,
                        uint64_t* wg_sum, uint64_t* wg_off, unsigned int* err, hipStream_t st);
void launch_dec_write(const uint32_t* words, const DecArrayDev* arrays, const uint32_t* wg_arr, uint64_t nwg,
                      const uint32_t* tables, const uint64_t* start, const uint32_t* count, const uint64_t* wg_off,
                      int32_t* out, unsigned int* err, hipStream_t st);

}  // namespace asz
