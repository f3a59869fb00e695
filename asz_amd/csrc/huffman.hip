// GPU Huffman coder for the .asz stream (huffman.hpp).
//
// Encoding, per array: LDS-windowed histogram -> (host builds the reference
// tree codebook) -> per-chunk code-length sums -> per-array exclusive scan of
// chunk bit offsets -> each workgroup packs its chunk's codes LSB-first into
// an LDS bit buffer and writes it out (boundary words with atomicOr, interior
// words with plain stores).  The stream buffer already holds the header bytes
// written by the host; payload bits are ORed around them.
//
// Decoding is a self-synchronising parallel decode of the one bit stream each
// array has (the format has no sync points): every subsequence of kSubBits
// bits is decoded from a guessed start, then starts are repaired from the
// predecessor's exit position until no start changes; Huffman parses from
// different starts merge after a few codewords, so this converges in a
// couple of passes.  Counts are scanned and a final pass writes the symbols.
#include "device.hpp"
#include "launch.hpp"

namespace asz {
namespace dev {

// ------------------------------------------------------------------ histogram

constexpr int kHistWin = 4096;

__global__ __launch_bounds__(256) void k_hist(const int32_t* __restrict__ sym, const HuffArrayDev* __restrict__ arrays,
                                              const uint32_t* __restrict__ chunk_arr, uint32_t* __restrict__ hist) {
    __shared__ uint32_t bins[kHistWin];
    const uint32_t k = chunk_arr[blockIdx.x];
    const HuffArrayDev A = arrays[k];
    const uint64_t c = blockIdx.x - A.hchunk0;
    const uint64_t first = c * kHistChunk;
    const uint64_t last = min(first + kHistChunk, A.count);
    // window centred on the symbol 0 (QBs concentrate there)
    int64_t zero_idx = -A.offset;
    int64_t wlo = zero_idx - kHistWin / 2;
    if (wlo + kHistWin > (int64_t)A.range) wlo = (int64_t)A.range - kHistWin;
    if (wlo < 0) wlo = 0;
    for (int i = threadIdx.x; i < kHistWin; i += blockDim.x) bins[i] = 0;
    __syncthreads();
    const int32_t* s = sym + A.src_off;
    uint32_t* h = hist + A.hist_off;
    for (uint64_t i = first + threadIdx.x; i < last; i += blockDim.x) {
        uint64_t idx = (uint64_t)((int64_t)s[i] - A.offset);
        uint64_t w = idx - (uint64_t)wlo;
        if (w < kHistWin) atomicAdd(&bins[w], 1u);
        else atomicAdd(&h[idx], 1u);
    }
    __syncthreads();
    const uint64_t wn = min<uint64_t>(kHistWin, A.range - (uint64_t)wlo);
    for (uint64_t i = threadIdx.x; i < wn; i += blockDim.x)
        if (bins[i]) atomicAdd(&h[(uint64_t)wlo + i], bins[i]);
}

// ------------------------------------------------------------------ encode

// book entry: code in bits 0..55, length in bits 56..63
__device__ __forceinline__ uint32_t book_len(uint64_t e) { return (uint32_t)(e >> 56); }
__device__ __forceinline__ uint64_t book_code(uint64_t e) { return e & ((1ull << 56) - 1); }

__global__ __launch_bounds__(256) void k_chunk_bits(const int32_t* __restrict__ sym,
                                                    const HuffArrayDev* __restrict__ arrays,
                                                    const uint32_t* __restrict__ chunk_arr,
                                                    const uint64_t* __restrict__ book, uint32_t* chunk_bits) {
    const uint32_t k = chunk_arr[blockIdx.x];
    const HuffArrayDev A = arrays[k];
    const uint64_t first = (blockIdx.x - A.chunk0) * (uint64_t)kEncChunk;
    const uint64_t last = min(first + kEncChunk, A.count);
    const int32_t* s = sym + A.src_off;
    const uint64_t* b = book + A.hist_off;
    uint32_t bits = 0;
    for (uint64_t i = first + threadIdx.x; i < last; i += blockDim.x) bits += book_len(b[(int64_t)s[i] - A.offset]);
    __shared__ uint32_t red[4];
    bits = wave_sum(bits);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = bits;
    __syncthreads();
    if (threadIdx.x == 0) chunk_bits[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// One workgroup per array: exclusive scan of its chunks' bit counts.
__global__ __launch_bounds__(1024) void k_chunk_scan(const HuffArrayDev* __restrict__ arrays,
                                                     const uint32_t* __restrict__ chunk_bits, uint64_t* chunk_off) {
    __shared__ uint64_t buf[1024];
    const HuffArrayDev A = arrays[blockIdx.x];
    if (A.single || A.nchunk == 0) return;
    uint64_t carry = A.payload_bit;
    for (uint64_t base = 0; base < A.nchunk; base += 1024) {
        uint64_t i = base + threadIdx.x;
        uint64_t v = i < A.nchunk ? chunk_bits[A.chunk0 + i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            uint64_t add = threadIdx.x >= (uint32_t)o ? buf[threadIdx.x - o] : 0;
            __syncthreads();
            buf[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < A.nchunk) chunk_off[A.chunk0 + i] = carry + buf[threadIdx.x] - v;
        uint64_t tot = buf[1023];
        __syncthreads();
        carry += tot;
    }
}

constexpr int kEncPerThread = kEncChunk / 256;   // 16
constexpr int kEncWords = kEncChunk + 2;          // LDS words for codes up to 32 bits

__global__ __launch_bounds__(256) void k_encode(const int32_t* __restrict__ sym, const HuffArrayDev* __restrict__ arrays,
                                                const uint32_t* __restrict__ chunk_arr,
                                                const uint64_t* __restrict__ book,
                                                const uint64_t* __restrict__ chunk_off, uint32_t* words,
                                                int use_lds) {
    __shared__ uint32_t lbuf[kEncWords];
    __shared__ uint32_t scan[256];
    const uint32_t k = chunk_arr[blockIdx.x];
    const HuffArrayDev A = arrays[k];
    if (A.single) return;
    const uint64_t first = (blockIdx.x - A.chunk0) * (uint64_t)kEncChunk;
    const uint64_t last = min(first + kEncChunk, A.count);
    const int32_t* s = sym + A.src_off;
    const uint64_t* b = book + A.hist_off;
    // thread t encodes symbols [first + 16 t, first + 16 t + 16)
    uint64_t e[kEncPerThread];
    uint32_t mybits = 0;
    const uint64_t my0 = first + (uint64_t)threadIdx.x * kEncPerThread;
#pragma unroll
    for (int j = 0; j < kEncPerThread; j++) {
        uint64_t i = my0 + j;
        e[j] = i < last ? b[(int64_t)s[i] - A.offset] : 0;
        mybits += book_len(e[j]);
    }
    // exclusive scan of per-thread bits
    uint32_t incl = mybits;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t v = __shfl_up(incl, o, 64);
        if ((threadIdx.x & 63) >= (uint32_t)o) incl += v;
    }
    if ((threadIdx.x & 63) == 63) scan[threadIdx.x / 64] = incl;
    __syncthreads();
    uint32_t wave_off = 0;
    for (uint32_t w = 0; w < threadIdx.x / 64; w++) wave_off += scan[w];
    const uint32_t total = scan[0] + scan[1] + scan[2] + scan[3];
    const uint32_t my_off = wave_off + incl - mybits;
    const uint64_t base_bit = chunk_off[blockIdx.x];
    if (total == 0) return;
    if (use_lds) {
        const uint32_t bit0 = (uint32_t)(base_bit & 31);
        const uint32_t nw = (bit0 + total + 31) / 32;
        for (uint32_t w = threadIdx.x; w < nw; w += 256) lbuf[w] = 0;
        __syncthreads();
        uint32_t p = bit0 + my_off;
#pragma unroll
        for (int j = 0; j < kEncPerThread; j++) {
            uint32_t len = book_len(e[j]);
            if (!len) continue;
            uint64_t code = book_code(e[j]);
            uint32_t w = p >> 5, sh = p & 31;
            atomicOr(&lbuf[w], (uint32_t)(code << sh));
            if (sh + len > 32) atomicOr(&lbuf[w + 1], (uint32_t)(code >> (32 - sh)));
            p += len;
        }
        __syncthreads();
        uint32_t* out = words + (base_bit >> 5);
        for (uint32_t w = threadIdx.x; w < nw; w += 256) {
            if (w == 0 || w == nw - 1) atomicOr(&out[w], lbuf[w]);
            else out[w] = lbuf[w];
        }
    } else {
        uint64_t p = base_bit + my_off;
        for (int j = 0; j < kEncPerThread; j++) {
            uint32_t len = book_len(e[j]);
            if (!len) continue;
            uint64_t code = book_code(e[j]);
            uint32_t sh = (uint32_t)(p & 31);
            uint32_t* w = words + (p >> 5);
            atomicOr(&w[0], (uint32_t)(code << sh));
            if (sh + len > 32) atomicOr(&w[1], (uint32_t)(code >> (32 - sh)));
            if (sh + len > 64) atomicOr(&w[2], (uint32_t)(code >> (64 - sh)));
            p += len;
        }
    }
}

}  // namespace dev

void launch_hist(const int32_t* sym, const HuffArrayDev* arrays, const uint32_t* chunk_arr, uint64_t nchunks,
                 uint32_t* hist, hipStream_t st) {
    if (!nchunks) return;
    hipLaunchKernelGGL(dev::k_hist, dim3((uint32_t)nchunks), dim3(256), 0, st, sym, arrays, chunk_arr, hist);
}
void launch_chunk_bits(const int32_t* sym, const HuffArrayDev* arrays, const uint32_t* chunk_arr, uint64_t nchunks,
                       const uint64_t* book, uint32_t* chunk_bits, hipStream_t st) {
    if (!nchunks) return;
    hipLaunchKernelGGL(dev::k_chunk_bits, dim3((uint32_t)nchunks), dim3(256), 0, st, sym, arrays, chunk_arr, book,
                       chunk_bits);
}
void launch_chunk_scan(const HuffArrayDev* arrays, uint32_t narrays, const uint32_t* chunk_bits, uint64_t* chunk_off,
                       hipStream_t st) {
    hipLaunchKernelGGL(dev::k_chunk_scan, dim3(narrays), dim3(1024), 0, st, arrays, chunk_bits, chunk_off);
}
void launch_encode(const int32_t* sym, const HuffArrayDev* arrays, const uint32_t* chunk_arr, uint64_t nchunks,
                   const uint64_t* book, const uint64_t* chunk_off, uint32_t* stream_words, int max_len,
                   hipStream_t st) {
    if (!nchunks) return;
    hipLaunchKernelGGL(dev::k_encode, dim3((uint32_t)nchunks), dim3(256), 0, st, sym, arrays, chunk_arr, book,
                       chunk_off, stream_words, max_len <= 32 ? 1 : 0);
}

}  // namespace asz
