// GPU side of the Huffman coder for the .asz stream (huffman.hpp).
//
// Histogram.  Every round workgroup bins its QBs in an LDS window (symbols
// [-win/2, win/2), win chosen per round) and adds the window into one of a
// few interleaved copies per round with atomics; the rare QB outside the
// window goes to the array's global window (kHistGlobal bins around 0).
// k_fold_rounds (or the fold workgroups beside the next round's t selection)
// sums the copies into the global window and folds the workgroups' errors
// and QB ranges; k_hist_small bins the two short arrays (t values and
// pivots) into their global windows.  The workgroup that finishes an
// array writes its QB range, errors and used bins straight into pinned host
// memory and raises a flag, so the host builds that codebook while later
// rounds run, without a copy or an API call in between.  An array whose range
// leaves the global window is binned again by k_hist over its exact range.
//
// Encoding.  The host builds the reference-tree codebooks.  The arrays are
// cut into chunks of kEncChunk symbols.  With every code at most 32 bits,
// k_encode_pack reads the symbols once and packs each chunk from bit 0 into a
// staging slot, counting its bits; k_chunk_scan turns the counts into bit
// positions per array, and k_encode_place shifts each slot into its position
// in the stream buffer (which holds the header bytes): interior words are
// stored, the two boundary words shared with the neighbours are ORed in.
// Longer codes: k_chunk_bits counts, k_chunk_scan places, k_encode_long ORs
// the codes into the stream.
#include "device.hpp"
#include "fold.hpp"
#include "launch.hpp"

namespace asz {
namespace dev {

// ------------------------------------------------------------------ histogram

// Exact-range histogram of the arrays listed in chunk_arr (a workgroup per
// kHistChunk symbols), for arrays whose QB range leaves the global window.
constexpr int kExactWin = 4096;

__global__ __launch_bounds__(256) void k_hist(const int32_t* __restrict__ sym, const HuffArrayDev* __restrict__ arrays,
                                              const uint32_t* __restrict__ chunk_arr, uint32_t* __restrict__ hist) {
    __shared__ uint32_t bins[kExactWin];
    const uint32_t k = chunk_arr[blockIdx.x];
    const HuffArrayDev A = arrays[k];
    const uint64_t c = blockIdx.x - A.hchunk0;
    const uint64_t first = c * kHistChunk;
    const uint64_t last = min(first + kHistChunk, A.count);
    // LDS window centred on the symbol 0 (QBs concentrate there)
    int64_t wlo = -A.offset - kExactWin / 2;
    if (wlo + kExactWin > (int64_t)A.range) wlo = (int64_t)A.range - kExactWin;
    if (wlo < 0) wlo = 0;
    for (int i = threadIdx.x; i < kExactWin; i += blockDim.x) bins[i] = 0;
    __syncthreads();
    const int32_t* s = sym + A.src_off;
    uint32_t* h = hist + A.hist_off;
    for (uint64_t i = first + threadIdx.x; i < last; i += blockDim.x) {
        const uint64_t idx = (uint64_t)((int64_t)s[i] - A.offset);
        const uint64_t w = idx - (uint64_t)wlo;
        if (w < (uint64_t)kExactWin) atomicAdd(&bins[w], 1u);
        else atomicAdd(&h[idx], 1u);
    }
    __syncthreads();
    const uint64_t wn = min<uint64_t>(kExactWin, A.range - (uint64_t)wlo);
    for (uint64_t i = threadIdx.x; i < wn; i += blockDim.x)
        if (bins[i]) atomicAdd(&h[(uint64_t)wlo + i], bins[i]);
}

// t array (y = 0) and pivot array (y = 1): a workgroup per kHistChunk
// symbols, LDS window flushed with atomics (these arrays are short).
__global__ __launch_bounds__(256) void k_hist_small(const int32_t* __restrict__ s0, uint64_t n0,
                                                    const int32_t* __restrict__ s1, uint64_t n1, uint32_t* g0,
                                                    uint32_t* g1, uint64_t s_stride, uint64_t g_stride) {
    __shared__ uint32_t lh[kHistWin];
    const bool second = blockIdx.y == 1;
    // field blockIdx.z: its arrays s_stride symbols and g_stride words further
    const int32_t* s = (second ? s1 : s0) + blockIdx.z * s_stride;
    const uint64_t n = second ? n1 : n0;
    uint32_t* gw = (second ? g1 : g0) + blockIdx.z * g_stride;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kHistWin; i += blockDim.x) lh[i] = 0;
    __syncthreads();
    const uint64_t first = (uint64_t)blockIdx.x * kHistChunk, last = min<uint64_t>(n, first + kHistChunk);
    for (uint64_t i = first + threadIdx.x; i < last; i += blockDim.x) hist_add(lh, kHistWin, gw, s[i]);
    __syncthreads();
    if (threadIdx.x < (uint32_t)kHistWin && lh[threadIdx.x])
        atomicAdd(&gw[kHistCentre - kHistWin / 2 + threadIdx.x], lh[threadIdx.x]);
}

// Fold of rounds r0 .. r0 + gridDim.y - 1 (grid x: the rounds' fold
// workgroups, z: fields).
__global__ __launch_bounds__(256) void k_fold_rounds(const FoldDesc* __restrict__ descs, uint32_t r0,
                                                     const Partial* __restrict__ part,
                                                     const uint32_t* __restrict__ phist, Partial* fpart, DevAcc* acc,
                                                     uint32_t* ghist, uint32_t* hres, uint32_t epoch, FieldSet fs) {
    if (fs.fd) {  // field blockIdx.z of a multi-field call
        const uint32_t f = blockIdx.z;
        part += f * fs.part;
        phist += f * fs.phist;
        fpart += f * fs.fpart;
        acc += f;
        ghist += f * fs.ghist;
        hres += f * fs.hres;
    }
    fold_group(descs, r0 + blockIdx.y, blockIdx.x, part, phist, fpart, acc, ghist, hres, epoch);
}

// Finishes the t array (k = 0: range and fallback count from acc) or the
// pivot array (k = 1: fold of the pivot kernel's Partials) for the host.
__global__ __launch_bounds__(256) void k_hist_tail(const DevAcc* __restrict__ acc, uint32_t k,
                                                   const Partial* __restrict__ part, uint64_t nparts,
                                                   const uint32_t* gw, uint32_t* hres, uint32_t epoch,
                                                   uint64_t part_stride, uint64_t g_stride, uint64_t h_stride) {
    // field blockIdx.x: its accumulators, partials, window and host slots
    acc += blockIdx.x;
    if (part) part += blockIdx.x * part_stride;
    gw += blockIdx.x * g_stride;
    hres += blockIdx.x * h_stride;
    HistTail t{};
    if (k == 0) {
        t = HistTail{acc->arr[0].qmin, acc->arr[0].qmax, acc->fallbacks, 0, 0.0, 0.0};
    } else {
        const Partial all = fold_partials(part, 0, nparts);
        t = HistTail{all.qmin, all.qmax, 0, 0, all.max_err, all.sq};
    }
    publish_host(gw, t, hres + k * kHostSlotWords, epoch);
}

// ------------------------------------------------------------------ encode

// book entry: code in bits 0..55, length in bits 56..63
__device__ __forceinline__ uint32_t book_len(uint64_t e) { return (uint32_t)(e >> 56); }
__device__ __forceinline__ uint64_t book_code(uint64_t e) { return e & ((1ull << 56) - 1); }

constexpr int kEncPerThread = kEncChunk / 256;  // 16
constexpr uint32_t kLdsLens = 8192;             // LDS code-length table of the bit-count pass

// The array of chunk c, from array k on (a workgroup's chunks increase, so
// its array index only moves forward; the table stays in global memory: one
// call may code many fields' arrays).
__device__ __forceinline__ uint32_t chunk_array(const HuffArrayDev* __restrict__ arrays, uint32_t narr, uint64_t c,
                                                uint32_t k) {
    while (k + 1 < narr && c >= arrays[k].chunk0 + arrays[k].nchunk) k++;
    return k;
}

// Symbols of chunk c of array A, coalesced: lane l of step j reads symbol
// 256 j + l (0 past the array's end).
__device__ __forceinline__ void load_chunk(const int32_t* __restrict__ sym, const HuffArrayDev& A, uint64_t c,
                                           int32_t v[kEncPerThread]) {
    const uint64_t first = (c - A.chunk0) * kEncChunk;
    const uint32_t n = (uint32_t)min<uint64_t>(kEncChunk, A.count - first);
    const int32_t* s = sym + A.src_off + first;
#pragma unroll
    for (int j = 0; j < kEncPerThread; j++) {
        const uint32_t i = j * 256 + threadIdx.x;
        v[j] = i < n ? s[i] : 0;
    }
}

// Bit-count pass: chunk_bits[c] = the code bits of chunk c.  Persistent:
// workgroup g takes chunks g, g + G, ... (G = gridDim.x), the next chunk's
// symbols in flight while the current one is summed, with the array's code
// lengths (len8: one byte per symbol of every array, at its book offset) in
// LDS when its alphabet fits.
__global__ __launch_bounds__(256) void k_chunk_bits(const int32_t* __restrict__ sym,
                                                    const HuffArrayDev* __restrict__ arrays, uint32_t narr,
                                                    uint64_t nchunks, const uint8_t* __restrict__ len8,
                                                    uint32_t* __restrict__ chunk_bits) {
    __shared__ uint8_t llen[kLdsLens];
    __shared__ uint32_t s_sum[4];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t G = gridDim.x;
    uint64_t c = blockIdx.x;
    if (c >= nchunks) return;
    uint32_t k = chunk_array(arrays, narr, c, 0);
    int32_t v[kEncPerThread];
    load_chunk(sym, arrays[k], c, v);
    int tab_k = -1;
    while (true) {
        const HuffArrayDev A = arrays[k];
        const uint8_t* L = len8 + A.book_off;
        const bool lds = A.range <= kLdsLens;
        if (lds && tab_k != (int)k) {
            for (uint32_t i = threadIdx.x; i < A.range; i += blockDim.x) llen[i] = L[i];
            tab_k = (int)k;
        }
        lds_barrier();  // table complete; the previous chunk's sums read
        const uint64_t first = (c - A.chunk0) * kEncChunk;
        const uint32_t n = (uint32_t)min<uint64_t>(kEncChunk, A.count - first);
        const uint64_t cn = c + G;
        const uint32_t kn = cn < nchunks ? chunk_array(arrays, narr, cn, k) : k;
        uint32_t bits = 0;
#pragma unroll
        for (int j = 0; j < kEncPerThread; j++) {
            const uint32_t i = j * 256 + threadIdx.x;
            if (i < n) {
                const int64_t idx = (int64_t)v[j] - A.offset;
                if (lds) bits += llen[idx];
                else bits += L[idx];
            }
        }
        if (cn < nchunks) load_chunk(sym, arrays[kn], cn, v);
        bits = wave_sum(bits);
        if (lane == 0) s_sum[wave] = bits;
        lds_barrier();
        if (threadIdx.x == 0) chunk_bits[c] = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
        if (cn >= nchunks) break;
        c = cn;
        k = kn;
    }
}

// chunk_off[c] = the bit position of chunk c in the stream: its array's
// payload start plus the sum of the bit counts of the array's earlier chunks.
// One workgroup (1024 threads) per 1024 consecutive chunks, thread t owning
// chunk 1024 g + t: a workgroup scan of the group's counts (wave shuffles,
// then the 16 wave totals), rebased at every array that starts inside the
// group; the array that started before the group adds the sum of its earlier
// chunks, which the workgroup reads coalesced.  It also zeroes the stream
// words that hold a chunk's first or last bit: the encoder ORs those (they
// are shared with a neighbouring chunk or with header bytes, which are
// scattered in after the encoder) and stores every other payload word whole.
__global__ __launch_bounds__(1024) void k_chunk_scan(const HuffArrayDev* __restrict__ arrays, uint32_t narr,
                                                     uint64_t nchunks, const uint32_t* __restrict__ chunk_bits,
                                                     uint64_t* chunk_off, uint32_t* words) {
    __shared__ uint64_t ex[1024];
    __shared__ uint64_t wsum[16];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t g0 = (uint64_t)blockIdx.x * 1024, c = g0 + threadIdx.x;
    const bool live = c < nchunks;
    const uint32_t bits = live ? chunk_bits[c] : 0u;
    // the array of a chunk: the last one whose first chunk is <= it (an array
    // without chunks has its successor's first chunk)
    auto array_of = [&](uint64_t x) {
        uint32_t lo = 0, hi = narr - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) / 2;
            if (arrays[mid].chunk0 <= x) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    };
    // the sum of the first array's chunks before the group
    const uint64_t a0 = arrays[array_of(g0)].chunk0;
    uint64_t pre = 0;
    for (uint64_t x = a0 + threadIdx.x; x < g0; x += 1024) pre += chunk_bits[x];
    pre = wave_sum(pre);
    if (lane == 0) wsum[wave] = pre;
    __syncthreads();
    pre = 0;
    for (int w = 0; w < 16; w++) pre += wsum[w];
    __syncthreads();
    // exclusive scan of the group's counts
    uint64_t incl = bits;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t t = __shfl_up(incl, o, 64);
        if (lane >= (uint32_t)o) incl += t;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint64_t excl = incl - bits;
    for (uint32_t w = 0; w < wave; w++) excl += wsum[w];
    ex[threadIdx.x] = excl;
    __syncthreads();
    if (!live) return;
    const HuffArrayDev A = arrays[array_of(c)];
    const uint64_t start = A.chunk0 > g0 ? A.chunk0 - g0 : 0;  // the array's first thread in the group
    const uint64_t off = A.payload_bit + excl - ex[start] + (A.chunk0 < g0 ? pre : 0);
    chunk_off[c] = off;
    words[off >> 5] = 0;
    if (bits) words[(off + bits - 1) >> 5] = 0;
}

// Pack pass (every code at most 32 bits) over chunks [c_begin, c_end):
// workgroup g packs chunks c_begin + g, c_begin + g + G, ... (persistent, the
// symbols of the next two chunks in flight while one is packed), each into
// its slot of the staging buffer (slot c: kEncChunk words at stage + c *
// kEncChunk, the chunk's first code at bit 0), and leaves the chunk's bit
// count in chunk_bits[c].  A chunk is read as 1024 int4 groups of four
// consecutive symbols, thread t owning groups t, t + 256, t + 512 and t + 768
// (one in each quarter of the chunk), so every load is a coalesced 16-byte
// load.  A thread joins its group's four codes into one 64-bit value (first
// symbol in the low bits); four workgroup scans of the group bit counts (two
// per 32-bit word: a quarter holds at most 32768 bits) place each group in
// the chunk, and each group is ORed into an LDS bit buffer with one to three
// LDS atomics.  A group longer than 64 bits (four codes averaging more than
// 16 bits) is ORed code by code from its symbols, read again.  The codes come
// from code32 / len8 (a 32-bit code and a byte length per symbol of every
// array, at its book offset), in LDS as (code, length) pairs when the array's
// alphabet is at most kLdsBook symbols.
__device__ __forceinline__ void load_groups(const int32_t* __restrict__ sym, const HuffArrayDev& A, uint64_t c,
                                            int4 v[4]) {
    const uint64_t first = (c - A.chunk0) * kEncChunk;
    const uint32_t n = (uint32_t)min<uint64_t>(kEncChunk, A.count - first);
    // 16-byte aligned: arrays start on 64-symbol boundaries, chunks on kEncChunk
    // symbols; a group that starts before n ends inside the array's padding
    const int4* s = reinterpret_cast<const int4*>(sym + A.src_off + first);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t g = 256 * j + threadIdx.x;
        v[j] = 4 * g < n ? s[g] : make_int4(0, 0, 0, 0);
    }
}

// LDS of an encoder workgroup: the chunk's bit buffer (32 bits per symbol at
// most, and the words the last group may spill into), the array's codebook as
// (code, length) pairs, and the four waves' group totals.
struct alignas(16) EncLds {
    uint32_t lbuf[kEncChunk + 4];
    uint2 lbook[kLdsBook];
    uint32_t s_tot[4][2];
};

// Array k's codebook into LDS unless it is there already (book_k: the array
// whose codebook LDS holds).  A fixed number of loads (no loop of unknown trip
// count), so that the waits for symbols loaded before it stay exact.  The
// caller's barrier completes it.
__device__ __forceinline__ void enc_book(EncLds& E, const HuffArrayDev& A, const uint32_t* C, const uint8_t* L,
                                         uint32_t k, int& book_k) {
    if (!A.lds_book || book_k == (int)k) return;
    static_assert(kLdsBook == 4 * 256, "four codebook entries per thread");
    const uint32_t t = threadIdx.x;
    uint32_t cb[4], lb[4];
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
        const uint32_t e = min(t + 256 * i, (uint32_t)A.range - 1);
        cb[i] = C[e];
        lb[i] = L[e];
    }
#pragma unroll
    for (uint32_t i = 0; i < 4; i++)
        if (t + 256 * i < A.range) E.lbook[t + 256 * i] = make_uint2(cb[i], lb[i]);
    book_k = (int)k;
}

// The thread's four groups (group j: symbols 1024 j + 4 t .. + 3 of the chunk,
// the first n live) joined: acc[j] their codes from bit 0, gb[j] their bit
// count.  Symbol - min in 32 bits (the index is below the alphabet size); the
// padding past n (read as 0) is clamped into the table and given no bits.
// LDS or global tables chosen at compile time (a run-time choice would merge
// the two into flat loads).  No branch per symbol: the codes are joined with
// shifts below 64 (a group of more than 64 bits is redone code by code).
template <bool LDS>
__device__ __forceinline__ void enc_join(const EncLds& E, const HuffArrayDev& A, const uint32_t* C, const uint8_t* L,
                                         const int4 v[4], uint32_t n, uint64_t acc[4], uint32_t gb[4]) {
    const uint32_t t = threadIdx.x;
    const uint32_t off32 = (uint32_t)A.offset, top = (uint32_t)A.range - 1;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t i0 = 1024 * j + 4 * t;
        const int32_t s[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
        uint32_t cd[4], ln[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const uint32_t idx = min((uint32_t)s[e] - off32, top);
            if constexpr (LDS) {
                const uint2 b = E.lbook[idx];
                cd[e] = b.x;
                ln[e] = b.y;
            } else {
                cd[e] = C[idx];
                ln[e] = L[idx];
            }
            const bool live = i0 + e < n;
            cd[e] = live ? cd[e] : 0u;
            ln[e] = live ? ln[e] : 0u;
        }
        const uint32_t p1 = ln[0], p2 = p1 + ln[1], p3 = p2 + ln[2];
        acc[j] = (uint64_t)cd[0] | ((uint64_t)cd[1] << (p1 & 63)) | ((uint64_t)cd[2] << (p2 & 63)) |
                 ((uint64_t)cd[3] << (p3 & 63));
        gb[j] = p3 + ln[3];
    }
}

// The joined groups into the (zero) bit buffer from bit 0: positions from
// inclusive wave scans of the group counts (DPP; two counts per 32-bit word,
// a quarter holds at most 32768 bits) and the four wave totals; groups of at
// most 64 bits ORed in with one to three LDS atomics and no wait between,
// longer ones code by code from their symbols (read again).  Returns the
// chunk's bit count; the buffer is complete on return.
__device__ __forceinline__ uint32_t enc_place_lds(EncLds& E, const HuffArrayDev& A, const uint32_t* C,
                                                  const uint8_t* L, const int32_t* sym, uint64_t first, uint32_t n,
                                                  const uint64_t acc[4], const uint32_t gb[4]) {
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const uint32_t p01 = gb[0] | (gb[1] << 16), p23 = gb[2] | (gb[3] << 16);
    const uint32_t i01 = wave_incl_scan(p01), i23 = wave_incl_scan(p23);
    if (lane == 63) {
        E.s_tot[wave][0] = i01;
        E.s_tot[wave][1] = i23;
    }
    lds_barrier();
    uint32_t a01 = 0, a23 = 0, w01 = 0, w23 = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4; w++) {
        const uint32_t x01 = E.s_tot[w][0], x23 = E.s_tot[w][1];
        a01 += x01;
        a23 += x23;
        if (w < wave) {
            w01 += x01;
            w23 += x23;
        }
    }
    const uint32_t T0 = a01 & 0xFFFFu, T1 = a01 >> 16, T2 = a23 & 0xFFFFu, T3 = a23 >> 16;
    const uint32_t e01 = w01 + i01 - p01, e23 = w23 + i23 - p23;
    const uint32_t pos[4] = {e01 & 0xFFFFu, T0 + (e01 >> 16), T0 + T1 + (e23 & 0xFFFFu), T0 + T1 + T2 + (e23 >> 16)};
    uint32_t* lbuf = E.lbuf;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t g = gb[j], p = pos[j];
        if (g && g <= 64) {
            const uint32_t w = p >> 5, s2 = p & 31;
            const uint64_t lo = acc[j] << s2;
            atomicOr(&lbuf[w], (uint32_t)lo);
            if (s2 + g > 32) atomicOr(&lbuf[w + 1], (uint32_t)(lo >> 32));
            if (s2 + g > 64) atomicOr(&lbuf[w + 2], (uint32_t)(acc[j] >> (64 - s2)));
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        uint32_t p = pos[j];
        if (gb[j] > 64) {
            const int32_t* s = sym + A.src_off + first;
            const uint32_t i0 = 1024 * j + 4 * t;
            for (int e = 0; e < 4; e++) {
                if (i0 + e >= n) break;
                const uint32_t idx = (uint32_t)((int64_t)s[i0 + e] - A.offset);
                uint32_t cd, ln;
                if (A.lds_book) {
                    cd = E.lbook[idx].x;
                    ln = E.lbook[idx].y;
                } else {
                    cd = C[idx];
                    ln = L[idx];
                }
                const uint32_t w = p >> 5, s2 = p & 31;
                if (ln) atomicOr(&lbuf[w], cd << s2);
                if (s2 + ln > 32) atomicOr(&lbuf[w + 1], cd >> (32 - s2));
                p += ln;
            }
        }
    }
    lds_barrier();
    return T0 + T1 + T2 + T3;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_encode_pack(
    const int32_t* __restrict__ sym, const HuffArrayDev* __restrict__ arrays, uint32_t narr, uint64_t c_begin,
    uint64_t c_end, const uint32_t* __restrict__ code32, const uint8_t* __restrict__ len8,
    uint32_t* __restrict__ chunk_bits, uint32_t* __restrict__ stage) {
    __shared__ EncLds E;
    const uint32_t t = threadIdx.x;
    const uint64_t G = gridDim.x;
    int book_k = -1;

    // Packs chunk c of array k from the symbols in v, then starts the load of
    // chunk c + 2G into v (two chunks in flight while this one is packed).
    // Workgroup g takes chunks g, g + G, ...: the few chunks of the arrays
    // whose codebook is too large for LDS (the slow ones) spread over many
    // workgroups.
    auto pack = [&](int4 (&v)[4], uint64_t c, uint32_t k) {
        const HuffArrayDev A = arrays[k];
        const uint64_t first = (c - A.chunk0) * kEncChunk;
        const uint32_t n = (uint32_t)min<uint64_t>(kEncChunk, A.count - first);
        const uint32_t* C = code32 + A.book_off;
        const uint8_t* L = len8 + A.book_off;
        enc_book(E, A, C, L, k, book_k);
        lds_barrier();  // LDS codebook complete; the previous chunk's buffer written out
        uint64_t acc[4];
        uint32_t gb[4];
        if (A.lds_book) enc_join<true>(E, A, C, L, v, n, acc, gb);
        else enc_join<false>(E, A, C, L, v, n, acc, gb);
        // the load of chunk c + 2G stays in flight while this one is packed
        // (the LDS-only barriers below do not wait for it)
        const uint64_t c2 = c + 2 * G;
        if (c2 < c_end) load_groups(sym, arrays[chunk_array(arrays, narr, c2, k)], c2, v);
        const uint32_t total = enc_place_lds(E, A, C, L, sym, first, n, acc, gb);
        const uint32_t nw4 = (total + 127) / 128;  // 16-byte units of the buffer (at most kEncChunk / 4)
        static_assert(kEncChunk / 4 == 4 * 256, "four buffer quads per thread at most");
        // a fixed number of (predicated) stores: the wait for the next
        // chunk's symbols then still knows how many memory operations follow
        // them; each quad is cleared by the thread that copied it out
        uint4* dst = reinterpret_cast<uint4*>(stage + c * kEncChunk);
        uint4* src = reinterpret_cast<uint4*>(E.lbuf);
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) {
            if (t + 256 * i < nw4) {
                dst[t + 256 * i] = src[t + 256 * i];
                src[t + 256 * i] = make_uint4(0, 0, 0, 0);
            }
        }
        if (t == 0) chunk_bits[c] = total;
    };

    uint64_t c = c_begin + blockIdx.x;
    if (c >= c_end) return;
    // the bit buffer starts zero (the first chunk's barrier orders this)
    static_assert((kEncChunk + 4) % 4 == 0, "buffer of whole quads");
    for (uint32_t i = t; i < (kEncChunk + 4) / 4; i += blockDim.x)
        reinterpret_cast<uint4*>(E.lbuf)[i] = make_uint4(0, 0, 0, 0);
    uint32_t k = chunk_array(arrays, narr, c, 0);
    int4 va[4], vb[4];
    load_groups(sym, arrays[k], c, va);
    uint32_t kb = k;
    if (c + G < c_end) {
        kb = chunk_array(arrays, narr, c + G, k);
        load_groups(sym, arrays[kb], c + G, vb);
    }
    // chunks alternate between the two register sets: va holds c, vb c + G
    while (true) {
        pack(va, c, k);
        if (c + G >= c_end) break;
        pack(vb, c + G, kb);
        c += 2 * G;
        if (c >= c_end) break;
        k = chunk_array(arrays, narr, c, kb);
        if (c + G < c_end) kb = chunk_array(arrays, narr, c + G, k);
    }
}

// Place pass: the staged bits of chunk c (one wave; a workgroup's four waves
// take four consecutive chunks) at bit chunk_off[c] of the stream.  The wave
// writes the chunk's aligned 16-byte quads of stream words, lane l taking
// quads q0 + l, q0 + l + 64, ... four at a time (their loads all issued
// before the stores).  Stream word i of the
// chunk (i = 0 at word chunk_off[c] / 32) is (s[i] << sh) | (s[i-1] >> (32 -
// sh)) over the staged words s (zero outside the chunk's nsw words), so a
// quad needs the five staged words from i0 - 1 on, which lie in two aligned
// staged quads: two 16-byte loads.  A quad wholly inside the chunk is stored
// with one 16-byte store; the chunk's first and last words (zeroed by
// k_chunk_scan, shared with the neighbouring chunks or header bytes) are ORed
// in; words of a quad outside the chunk are left alone.
__global__ __launch_bounds__(256) void k_encode_place(const uint32_t* __restrict__ chunk_bits,
                                                      const uint64_t* __restrict__ chunk_off, uint64_t nchunks,
                                                      const uint32_t* __restrict__ stage, uint32_t* words) {
    const uint64_t c = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (c >= nchunks) return;  // no barrier below: each wave works alone
    const uint32_t bits = chunk_bits[c];
    const uint64_t off = chunk_off[c];
    asm volatile("" ::"v"(bits), "v"(off));  // both loads in flight before the branch needs either
    if (!bits) return;
    const uint32_t sh = (uint32_t)(off & 31);
    const int64_t nsw = (bits + 31) >> 5, ndw = (sh + bits + 31) >> 5;  // ndw <= nsw + 1
    const uint4* src4 = reinterpret_cast<const uint4*>(stage + c * kEncChunk);  // 16 KB slots
    const uint64_t g0 = off >> 5;  // the chunk's first stream word
    const uint64_t q0 = g0 >> 2, q1 = (g0 + (uint64_t)ndw - 1) >> 2;
    // staged word i0 - 1 of every quad sits at position rem of staged quad m
    // (i0 - 1 = 4 m + rem; rem is the same for every quad of the chunk)
    const int64_t a0 = (int64_t)(4 * q0) - (int64_t)g0 - 1;  // >= -4
    const uint32_t rem = (uint32_t)(a0 & 3);
    // staged quad m: loaded unconditionally from a clamped index (so that
    // every load of a group is in flight at once), then words past nsw (or
    // before 0) cleared
    const int64_t mmax = (nsw - 1) >> 2;
    auto load_quad = [&](int64_t m) { return src4[min(max(m, (int64_t)0), mmax)]; };
    auto clip_quad = [&](uint4 v, int64_t m, bool use) -> uint4 {
        const int64_t j = 4 * m;
        const bool in = use && m >= 0 && j < nsw;
        v.x = in ? v.x : 0u;
        v.y = in && j + 1 < nsw ? v.y : 0u;
        v.z = in && j + 2 < nsw ? v.z : 0u;
        v.w = in && j + 3 < nsw ? v.w : 0u;
        return v;
    };
    for (uint64_t qb = q0; qb <= q1; qb += 4 * 64) {
        // output quad q needs staged quads m and m + 1, and quad q + 1 (the
        // next lane) needs m + 1 and m + 2: each lane loads its quad m (lanes
        // up to q1 + 1) and takes m + 1 from the next lane; lane 63 takes it
        // from lane 0 of the next group, or loads it after the last group
        const int64_t m0 = (a0 + 4 * (int64_t)(qb + lane - q0) - (int64_t)rem) / 4;  // exact: the numerator is 4 m
        uint4 A[4], B[4];
#pragma unroll
        for (int u = 0; u < 4; u++) A[u] = load_quad(m0 + 64 * u);
        uint4 tail = load_quad(m0 + 193);
#pragma unroll
        for (int u = 0; u < 4; u++) A[u] = clip_quad(A[u], m0 + 64 * u, qb + lane + 64 * u <= q1 + 1);
        tail = clip_quad(tail, m0 + 193, lane == 63 && qb + 255 <= q1);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            B[u] = make_uint4(__shfl_down(A[u].x, 1, 64), __shfl_down(A[u].y, 1, 64), __shfl_down(A[u].z, 1, 64),
                              __shfl_down(A[u].w, 1, 64));
            const uint4 n0 = u < 3 ? make_uint4(__shfl(A[u + 1].x, 0, 64), __shfl(A[u + 1].y, 0, 64),
                                                __shfl(A[u + 1].z, 0, 64), __shfl(A[u + 1].w, 0, 64))
                                   : tail;
            if (lane == 63) B[u] = n0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint64_t q = qb + lane + 64 * u;
            if (q > q1) break;
            // w[k]: staged word i0 - 1 + k
            const uint32_t x[8] = {A[u].x, A[u].y, A[u].z, A[u].w, B[u].x, B[u].y, B[u].z, B[u].w};
            uint32_t w[5];
            if (rem == 0) { w[0] = x[0]; w[1] = x[1]; w[2] = x[2]; w[3] = x[3]; w[4] = x[4]; }
            else if (rem == 1) { w[0] = x[1]; w[1] = x[2]; w[2] = x[3]; w[3] = x[4]; w[4] = x[5]; }
            else if (rem == 2) { w[0] = x[2]; w[1] = x[3]; w[2] = x[4]; w[3] = x[5]; w[4] = x[6]; }
            else { w[0] = x[3]; w[1] = x[4]; w[2] = x[5]; w[3] = x[6]; w[4] = x[7]; }
            uint32_t val[4];
#pragma unroll
            for (int e = 0; e < 4; e++) val[e] = sh ? (w[e + 1] << sh) | (w[e] >> (32 - sh)) : w[e + 1];
            const int64_t i0 = (int64_t)(4 * q) - (int64_t)g0;
            if (i0 >= 1 && i0 + 3 <= ndw - 2) {
                *reinterpret_cast<uint4*>(words + 4 * q) = make_uint4(val[0], val[1], val[2], val[3]);
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int64_t i = i0 + e;
                    if (i < 0 || i >= ndw) continue;
                    if (i == 0 || i == ndw - 1) atomicOr(&words[4 * q + e], val[e]);
                    else words[4 * q + e] = val[e];
                }
            }
        }
    }
}

// Long-code encode pass (some code longer than 32 bits): every chunk's codes
// ORed straight into the stream at its bit position chunk_off[c], from the
// 64-bit book (code in bits 0..55, length in bits 56..63), after the
// workgroup has zeroed the words that lie wholly inside the chunk.
// Persistent: workgroup g codes chunks g, g + G, ...; thread t owns the 16
// consecutive symbols 16 t .. 16 t + 15 of the chunk, and a workgroup scan of
// the per-thread bit counts gives each thread's first bit.
__global__ __launch_bounds__(256) void k_encode_long(const int32_t* __restrict__ sym,
                                                     const HuffArrayDev* __restrict__ arrays, uint32_t narr,
                                                     uint64_t nchunks, const uint64_t* __restrict__ book,
                                                     const uint64_t* __restrict__ chunk_off, uint32_t* words) {
    __shared__ uint32_t s_scan[4];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t G = gridDim.x;
    uint64_t c = blockIdx.x;
    if (c >= nchunks) return;
    uint32_t k = chunk_array(arrays, narr, c, 0);
    while (true) {
        const HuffArrayDev A = arrays[k];
        const uint64_t first = (c - A.chunk0) * kEncChunk;
        const uint32_t n = (uint32_t)min<uint64_t>(kEncChunk, A.count - first);
        const uint64_t* b = book + A.book_off;
        const uint32_t my0 = threadIdx.x * kEncPerThread;
        const int32_t* s = sym + A.src_off + first;
        uint64_t e64[kEncPerThread];
        uint32_t mybits = 0;
#pragma unroll
        for (int j = 0; j < kEncPerThread; j++) {
            const uint32_t i = my0 + j;
            uint64_t ent = 0;
            if (i < n) ent = b[(int64_t)s[i] - A.offset];
            e64[j] = ent;
            mybits += book_len(ent);
        }
        uint32_t incl = mybits;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(incl, o, 64);
            if (lane >= (uint32_t)o) incl += u;
        }
        if (lane == 63) s_scan[wave] = incl;
        __syncthreads();
        uint32_t wave_off = 0;
        for (uint32_t w = 0; w < wave; w++) wave_off += s_scan[w];
        // the words strictly inside the chunk are its own: zeroed here before
        // the codes are ORed in (k_chunk_scan zeroed the first and last)
        const uint32_t total = s_scan[0] + s_scan[1] + s_scan[2] + s_scan[3];
        const uint64_t off = chunk_off[c];
        if (total)
            for (uint64_t w = (off >> 5) + 1 + threadIdx.x; w < ((off + total - 1) >> 5); w += blockDim.x) words[w] = 0;
        __syncthreads();
        uint64_t p = off + wave_off + incl - mybits;
        for (int j = 0; j < kEncPerThread; j++) {
            const uint64_t ent = e64[j];
            const uint32_t l = book_len(ent);
            if (!l) continue;
            const uint64_t cd = book_code(ent);
            const uint32_t s2 = (uint32_t)(p & 31);
            uint32_t* w = words + (p >> 5);
            atomicOr(&w[0], (uint32_t)(cd << s2));
            if (s2 + l > 32) atomicOr(&w[1], (uint32_t)(cd >> (32 - s2)));
            if (s2 + l > 64) atomicOr(&w[2], (uint32_t)(cd >> (64 - s2)));
            p += l;
        }
        __syncthreads();  // s_scan free for the next chunk
        const uint64_t cn = c + G;
        if (cn >= nchunks) break;
        k = chunk_array(arrays, narr, cn, k);
        c = cn;
    }
}
}  // namespace dev

void launch_hist(const int32_t* sym, const HuffArrayDev* arrays, const uint32_t* chunk_arr, uint64_t nchunks,
                 uint32_t* hist, hipStream_t st) {
    if (!nchunks) return;
    hipLaunchKernelGGL(dev::k_hist, dim3((uint32_t)nchunks), dim3(256), 0, st, sym, arrays, chunk_arr, hist);
}
void launch_hist_small(const int32_t* s0, uint64_t n0, const int32_t* s1, uint64_t n1, uint32_t* g0, uint32_t* g1,
                       hipStream_t st, uint32_t nfields, uint64_t s_stride, uint64_t g_stride) {
    const uint64_t nc = (std::max(n0, n1) + kHistChunk - 1) / kHistChunk;
    if (!nc) return;
    hipLaunchKernelGGL(dev::k_hist_small, dim3((uint32_t)nc, 2, nfields), dim3(256), 0, st, s0, n0, s1, n1, g0, g1,
                       s_stride, g_stride);
}
void launch_fold_rounds(const FoldDesc* descs, uint32_t r0, uint32_t nr, uint32_t max_groups, const Partial* part,
                        const uint32_t* phist, Partial* fpart, DevAcc* acc, uint32_t* ghist, uint32_t* hres,
                        uint32_t epoch, hipStream_t st, const FieldSet& fs, uint32_t nfields) {
    if (!nr || !max_groups) return;
    hipLaunchKernelGGL(dev::k_fold_rounds, dim3(max_groups, nr, nfields), dim3(256), 0, st, descs, r0, part, phist,
                       fpart, acc, ghist, hres, epoch, fs);
}
void launch_hist_tail(const DevAcc* acc, uint32_t k, const Partial* part, uint64_t nparts, const uint32_t* gw,
                      uint32_t* hres, uint32_t epoch, hipStream_t st, uint32_t nfields, uint64_t part_stride,
                      uint64_t g_stride, uint64_t h_stride) {
    hipLaunchKernelGGL(dev::k_hist_tail, dim3(nfields), dim3(256), 0, st, acc, k, part, nparts, gw, hres, epoch,
                       part_stride, g_stride, h_stride);
}
namespace {
// Workgroups of a persistent pass: one resident wave of them (CUs x the
// workgroups that fit on a CU), at most kEncGroups and at most the chunks.
uint32_t enc_groups(const void* kernel, uint64_t nchunks) {
    int dev_id = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev_id);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0);
    const uint64_t resident = (uint64_t)std::max(cus, 1) * (uint64_t)std::max(per_cu, 1);
    return (uint32_t)std::min<uint64_t>(nchunks, std::min<uint64_t>(resident, kEncGroups));
}
}  // namespace

void launch_encode_pack(const int32_t* sym, const HuffArrayDev* arrays, uint32_t narr, uint64_t c_begin,
                        uint64_t c_end, const uint32_t* code32, const uint8_t* len8, uint32_t* chunk_bits,
                        uint32_t* stage, hipStream_t st) {
    if (c_end <= c_begin || !narr) return;
    hipLaunchKernelGGL(dev::k_encode_pack, dim3(enc_groups((const void*)dev::k_encode_pack, c_end - c_begin)),
                       dim3(256), 0, st, sym, arrays, narr, c_begin, c_end, code32, len8, chunk_bits, stage);
}

void launch_encode_place(const HuffArrayDev* arrays, uint32_t narr, uint64_t nchunks, uint32_t* chunk_bits,
                         uint64_t* chunk_off, const uint32_t* stage, uint32_t* stream_words, hipStream_t st) {
    if (!nchunks || !narr) return;
    if (nchunks > 0x7FFFFFFFull) fail(ASZ_ERR_ARG, "symbol arrays exceed the encoder's launch grid");
    hipLaunchKernelGGL(dev::k_chunk_scan, dim3((uint32_t)((nchunks + 1023) / 1024)), dim3(1024), 0, st, arrays, narr,
                       nchunks, chunk_bits, chunk_off, stream_words);
    hipLaunchKernelGGL(dev::k_encode_place, dim3((uint32_t)((nchunks + 3) / 4)), dim3(256), 0, st, chunk_bits,
                       chunk_off, nchunks, stage, stream_words);
}

void launch_encode(const int32_t* sym, const HuffArrayDev* arrays, uint32_t narr, uint64_t nchunks,
                   const uint64_t* book, const uint32_t* code32, const uint8_t* len8, uint32_t* chunk_bits,
                   uint64_t* chunk_off, uint32_t* stream_words, int max_len, uint32_t* stage, hipStream_t st) {
    if (!nchunks || !narr) return;
    if (nchunks > 0x7FFFFFFFull) fail(ASZ_ERR_ARG, "symbol arrays exceed the encoder's launch grid");
    if (max_len <= 32) {
        if (!stage) fail(ASZ_ERR_ARG, "encoder staging buffer missing");
        launch_encode_pack(sym, arrays, narr, 0, nchunks, code32, len8, chunk_bits, stage, st);
        launch_encode_place(arrays, narr, nchunks, chunk_bits, chunk_off, stage, stream_words, st);
        return;
    }
    hipLaunchKernelGGL(dev::k_chunk_bits, dim3(enc_groups((const void*)dev::k_chunk_bits, nchunks)), dim3(256), 0, st,
                       sym, arrays, narr, nchunks, len8, chunk_bits);
    hipLaunchKernelGGL(dev::k_chunk_scan, dim3((uint32_t)((nchunks + 1023) / 1024)), dim3(1024), 0, st, arrays, narr,
                       nchunks, chunk_bits, chunk_off, stream_words);
    hipLaunchKernelGGL(dev::k_encode_long, dim3(enc_groups((const void*)dev::k_encode_long, nchunks)), dim3(256), 0,
                       st, sym, arrays, narr, nchunks, book, chunk_off, stream_words);
}

}  // namespace asz

