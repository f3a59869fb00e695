// GPU side of the Huffman coder for the .asz stream (huffman.hpp).
//
// Histogram.  The round kernels leave one LDS window histogram per workgroup
// (symbols [-win/2, win/2), win chosen per round) and add the rest of their
// QBs to the array's global window (kHistGlobal bins around 0) with atomics.
// k_hist_reduce folds the per-workgroup windows into the global windows
// without contended atomics (each workgroup of the reduction owns a slice of
// the partials), k_hist_small bins the two short arrays (t values and
// pivots) the same way, and k_hist_compact packs each array's [min, max]
// range of bins next to the accumulators so that the host reads everything
// with one copy.  An array whose range leaves the global window is binned
// again by k_hist over its exact range.
//
// Encoding.  The host builds the reference-tree codebooks.  k_encode_chunks
// codes each chunk of kEncChunk symbols (one workgroup) from an LDS copy of
// the codebook when it is small, packs the codes LSB-first in an LDS bit
// buffer and stores it at the start of the chunk's own scratch slot, with the
// chunk's bit count.  k_chunk_scan turns the counts into absolute bit
// positions per array, and k_compact_bits shifts every chunk's bits into
// place in the stream buffer (which already holds the header bytes): interior
// words are stored, the two boundary words shared with the neighbours are
// ORed in.
#include "device.hpp"
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
                                                    uint32_t* g1) {
    __shared__ uint32_t lh[kHistWin];
    const bool second = blockIdx.y == 1;
    const int32_t* s = second ? s1 : s0;
    const uint64_t n = second ? n1 : n0;
    uint32_t* gw = second ? g1 : g0;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kHistWin; i += blockDim.x) lh[i] = 0;
    __syncthreads();
    const uint64_t first = (uint64_t)blockIdx.x * kHistChunk, last = min<uint64_t>(n, first + kHistChunk);
    for (uint64_t i = first + threadIdx.x; i < last; i += blockDim.x) hist_add(lh, kHistWin, gw, s[i]);
    __syncthreads();
    if (threadIdx.x < (uint32_t)kHistWin && lh[threadIdx.x])
        atomicAdd(&gw[kHistCentre - kHistWin / 2 + threadIdx.x], lh[threadIdx.x]);
}

// grid (kHistReduceGroups, rounds): workgroup (g, r) sums slice g of round r's
// per-workgroup windows and adds the slice total to the global window.
// ranges[4r ..]: first partial, workgroups, first histogram word, window bins.
constexpr uint32_t kHistReduceGroups = 64;

__global__ __launch_bounds__(256) void k_hist_reduce(const uint32_t* __restrict__ phist,
                                                     const uint64_t* __restrict__ ranges, uint32_t* ghist) {
    const uint32_t r = blockIdx.y;
    const uint64_t n = ranges[4 * r + 1], off = ranges[4 * r + 2], win = ranges[4 * r + 3];
    const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t p0 = (uint64_t)blockIdx.x * per, p1 = min(n, p0 + per);
    if (p0 >= p1) return;
    uint32_t* gw = ghist + (uint64_t)(2 + r) * kHistGlobal + kHistCentre - win / 2;
    for (uint64_t bin = threadIdx.x; bin < win; bin += blockDim.x) {
        uint32_t s = 0;
        for (uint64_t p = p0; p < p1; p++) s += phist[off + p * win + bin];
        if (s) atomicAdd(&gw[bin], s);
    }
}

// One workgroup per array: bins [qmin, qmax] of arrays inside the global
// window, packed array after array (arrays outside it contribute nothing).
__device__ __forceinline__ uint64_t compact_span(const ArrayAcc& a) {
    if (a.qmin > a.qmax || a.qmin < -kHistCentre || a.qmax >= kHistCentre) return 0;
    return (uint64_t)((int64_t)a.qmax - a.qmin + 1);
}

__global__ __launch_bounds__(256) void k_hist_compact(const DevAcc* __restrict__ acc,
                                                      const uint32_t* __restrict__ ghist, uint32_t* out) {
    const uint32_t a = blockIdx.x;
    uint64_t off = 0;
    for (uint32_t k = 0; k < a; k++) off += compact_span(acc->arr[k]);
    const uint64_t n = compact_span(acc->arr[a]);
    const uint32_t* src = ghist + (uint64_t)a * kHistGlobal + (acc->arr[a].qmin + kHistCentre);
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) out[off + i] = src[i];
}

// ------------------------------------------------------------------ encode

// book entry: code in bits 0..55, length in bits 56..63
__device__ __forceinline__ uint32_t book_len(uint64_t e) { return (uint32_t)(e >> 56); }
__device__ __forceinline__ uint64_t book_code(uint64_t e) { return e & ((1ull << 56) - 1); }

constexpr int kEncPerThread = kEncChunk / 256;  // 16
constexpr int kEncWords = kEncChunk + 2;        // LDS bit buffer for codes up to 32 bits

__global__ __launch_bounds__(256) void k_encode_chunks(const int32_t* __restrict__ sym,
                                                       const HuffArrayDev* __restrict__ arrays, uint32_t narr,
                                                       const uint64_t* __restrict__ book, uint32_t* scratch,
                                                       uint64_t slot_words, uint32_t* __restrict__ chunk_bits,
                                                       int use_lds) {
    __shared__ uint32_t lbuf[kEncWords];
    __shared__ uint32_t lbook[kLdsBook];
    __shared__ uint32_t s_scan[4];
    __shared__ uint32_t s_k;
    const uint32_t g = blockIdx.x;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (threadIdx.x == 0) {
        uint32_t k = 0;
        while (k + 1 < narr && !(g >= arrays[k].chunk0 && g < arrays[k].chunk0 + arrays[k].nchunk)) k++;
        s_k = k;
    }
    __syncthreads();
    const HuffArrayDev A = arrays[s_k];
    const uint64_t* b = book + A.book_off;
    if (A.lds_book) {
        for (uint32_t i = threadIdx.x; i < A.range; i += blockDim.x) {
            const uint64_t e = b[i];
            lbook[i] = (uint32_t)(e & 0xFFFFFFu) | (book_len(e) << 24);
        }
        __syncthreads();
    }
    const uint64_t first = (uint64_t)(g - A.chunk0) * kEncChunk;
    const uint64_t last = min<uint64_t>(first + kEncChunk, A.count);
    const int32_t* s = sym + A.src_off;
    // thread t codes symbols [first + 16 t, first + 16 t + 16)
    uint64_t e[kEncPerThread];
    uint32_t mybits = 0;
    const uint64_t my0 = first + (uint64_t)threadIdx.x * kEncPerThread;
#pragma unroll
    for (int j = 0; j < kEncPerThread; j++) {
        const uint64_t i = my0 + j;
        uint64_t v = 0;
        if (i < last) {
            const int64_t idx = (int64_t)s[i] - A.offset;
            if (A.lds_book) {
                const uint32_t l = lbook[idx];
                v = (uint64_t)(l & 0xFFFFFFu) | ((uint64_t)(l >> 24) << 56);
            } else {
                v = b[idx];
            }
        }
        e[j] = v;
        mybits += book_len(v);
    }
    // exclusive scan of the per-thread bit counts
    uint32_t incl = mybits;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o, 64);
        if (lane >= (uint32_t)o) incl += v;
    }
    if (lane == 63) s_scan[wave] = incl;
    __syncthreads();
    uint32_t wave_off = 0;
    for (uint32_t w = 0; w < wave; w++) wave_off += s_scan[w];
    const uint32_t total = s_scan[0] + s_scan[1] + s_scan[2] + s_scan[3];
    const uint32_t my_off = wave_off + incl - mybits;
    if (threadIdx.x == 0) chunk_bits[g] = total;
    uint32_t* out = scratch + (uint64_t)g * slot_words;
    if (use_lds) {
        const uint32_t nw = (total + 31) / 32;
        for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x) lbuf[w] = 0;
        __syncthreads();
        uint32_t p = my_off;
#pragma unroll
        for (int j = 0; j < kEncPerThread; j++) {
            const uint32_t len = book_len(e[j]);
            if (!len) continue;
            const uint64_t code = book_code(e[j]);
            const uint32_t w = p >> 5, sh = p & 31;
            atomicOr(&lbuf[w], (uint32_t)(code << sh));
            if (sh + len > 32) atomicOr(&lbuf[w + 1], (uint32_t)(code >> (32 - sh)));
            p += len;
        }
        __syncthreads();
        for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x) out[w] = lbuf[w];
    } else {
        // codes longer than 32 bits: OR straight into the zeroed slot
        uint64_t p = my_off;
        for (int j = 0; j < kEncPerThread; j++) {
            const uint32_t len = book_len(e[j]);
            if (!len) continue;
            const uint64_t code = book_code(e[j]);
            const uint32_t sh = (uint32_t)(p & 31);
            uint32_t* w = out + (p >> 5);
            atomicOr(&w[0], (uint32_t)(code << sh));
            if (sh + len > 32) atomicOr(&w[1], (uint32_t)(code >> (32 - sh)));
            if (sh + len > 64) atomicOr(&w[2], (uint32_t)(code >> (64 - sh)));
            p += len;
        }
    }
}

// One workgroup per array: chunk_off[c] = payload start + exclusive scan of
// the array's chunk bit counts.
__global__ __launch_bounds__(1024) void k_chunk_scan(const HuffArrayDev* __restrict__ arrays,
                                                     const uint32_t* __restrict__ chunk_bits, uint64_t* chunk_off) {
    __shared__ uint64_t buf[1024];
    const HuffArrayDev A = arrays[blockIdx.x];
    if (A.nchunk == 0) return;
    uint64_t carry = A.payload_bit;
    for (uint64_t base = 0; base < A.nchunk; base += 1024) {
        const uint64_t i = base + threadIdx.x;
        const uint64_t v = i < A.nchunk ? chunk_bits[A.chunk0 + i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const uint64_t add = threadIdx.x >= (uint32_t)o ? buf[threadIdx.x - o] : 0;
            __syncthreads();
            buf[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < A.nchunk) chunk_off[A.chunk0 + i] = carry + buf[threadIdx.x] - v;
        const uint64_t tot = buf[1023];
        __syncthreads();
        carry += tot;
    }
}

// One workgroup per chunk: the chunk's bits from its scratch slot to bit
// position chunk_off in the stream.
__global__ __launch_bounds__(256) void k_compact_bits(const uint32_t* __restrict__ scratch, uint64_t slot_words,
                                                      const uint32_t* __restrict__ chunk_bits,
                                                      const uint64_t* __restrict__ chunk_off, uint32_t* words) {
    const uint32_t g = blockIdx.x;
    const uint32_t nb = chunk_bits[g];
    if (nb == 0) return;
    const uint64_t o = chunk_off[g];
    const uint32_t sh = (uint32_t)(o & 31);
    const uint32_t nsrc = (nb + 31) / 32, nout = (sh + nb + 31) / 32;
    const uint32_t* src = scratch + (uint64_t)g * slot_words;
    uint32_t* dst = words + (o >> 5);
    for (uint32_t i = threadIdx.x; i < nout; i += blockDim.x) {
        const uint32_t cur = i < nsrc ? src[i] : 0u;
        uint32_t v = cur << sh;
        if (sh && i > 0) v |= src[i - 1] >> (32 - sh);
        if (i == 0 || i == nout - 1) atomicOr(&dst[i], v);
        else dst[i] = v;
    }
}

}  // namespace dev

void launch_hist(const int32_t* sym, const HuffArrayDev* arrays, const uint32_t* chunk_arr, uint64_t nchunks,
                 uint32_t* hist, hipStream_t st) {
    if (!nchunks) return;
    hipLaunchKernelGGL(dev::k_hist, dim3((uint32_t)nchunks), dim3(256), 0, st, sym, arrays, chunk_arr, hist);
}
void launch_hist_small(const int32_t* s0, uint64_t n0, const int32_t* s1, uint64_t n1, uint32_t* g0, uint32_t* g1,
                       hipStream_t st) {
    const uint64_t nc = (std::max(n0, n1) + kHistChunk - 1) / kHistChunk;
    if (!nc) return;
    hipLaunchKernelGGL(dev::k_hist_small, dim3((uint32_t)nc, 2), dim3(256), 0, st, s0, n0, s1, n1, g0, g1);
}
void launch_hist_reduce(const uint32_t* phist, const uint64_t* ranges, uint32_t rounds, uint32_t* ghist,
                        hipStream_t st) {
    if (!rounds) return;
    hipLaunchKernelGGL(dev::k_hist_reduce, dim3(dev::kHistReduceGroups, rounds), dim3(256), 0, st, phist, ranges,
                       ghist);
}
void launch_hist_compact(const DevAcc* acc, uint32_t narr, const uint32_t* ghist, uint32_t* out, hipStream_t st) {
    hipLaunchKernelGGL(dev::k_hist_compact, dim3(narr), dim3(256), 0, st, acc, ghist, out);
}
void launch_encode_chunks(const int32_t* sym, const HuffArrayDev* arrays, uint32_t narr, uint64_t nchunks,
                          const uint64_t* book, uint32_t* scratch, uint64_t slot_words, uint32_t* chunk_bits,
                          int max_len, hipStream_t st) {
    if (!nchunks) return;
    hipLaunchKernelGGL(dev::k_encode_chunks, dim3((uint32_t)nchunks), dim3(256), 0, st, sym, arrays, narr, book,
                       scratch, slot_words, chunk_bits, max_len <= 32 ? 1 : 0);
}
void launch_chunk_scan(const HuffArrayDev* arrays, uint32_t narrays, const uint32_t* chunk_bits, uint64_t* chunk_off,
                       hipStream_t st) {
    hipLaunchKernelGGL(dev::k_chunk_scan, dim3(narrays), dim3(1024), 0, st, arrays, chunk_bits, chunk_off);
}
void launch_compact_bits(const uint32_t* scratch, uint64_t slot_words, const uint32_t* chunk_bits,
                         const uint64_t* chunk_off, uint64_t nchunks, uint32_t* stream_words, hipStream_t st) {
    if (!nchunks) return;
    hipLaunchKernelGGL(dev::k_compact_bits, dim3((uint32_t)nchunks), dim3(256), 0, st, scratch, slot_words,
                       chunk_bits, chunk_off, stream_words);
}

}  // namespace asz
