// Parallel Huffman decoding of the .asz bit streams (huffman.hpp:383-526).
//
// The format has one bit stream per array and no sync points.  Each array's
// payload is cut into subsequences of kSubBits bits; a subsequence owns the
// codewords that start inside it.  Workgroups hold 256 consecutive
// subsequences of one array, with that array's root lookup table in LDS.
//   1. k_dec_sync: decode every subsequence from its first bit (a guess),
//      recording the start, the exit (first codeword boundary past its end)
//      and the codeword count.
//   2. k_dec_fix (repeated until no exit changes): a subsequence whose start
//      differs from its predecessor's exit re-parses from that exit, walking
//      its old parse in lockstep; Huffman parses from different starts merge
//      after a few codewords, after which count and exit follow from the old
//      parse.  Subsequence 0 of each array starts at a true boundary, so the
//      chain converges.
//   3. k_dec_wgsum / k_dec_wgscan: output offsets (two-level scan).
//   4. k_dec_write: decode again from the exact starts, staging symbols in LDS
//      and storing them as contiguous runs.
#include "device.hpp"
#include "launch.hpp"

namespace asz {
namespace dev {

constexpr int kRootMax = 1 << kRootBits;

struct BitReader {
    const uint32_t* __restrict__ w;
    uint64_t idx;   // next word to load
    uint64_t buf;   // unread bits, LSB first
    int nb;         // valid bits in buf
    uint64_t pos;   // absolute bit position of buf bit 0

    __device__ __forceinline__ void init(const uint32_t* words, uint64_t p) {
        w = words;
        idx = p >> 5;
        const uint32_t sh = (uint32_t)(p & 31);
        buf = ((uint64_t)w[idx] | ((uint64_t)w[idx + 1] << 32)) >> sh;
        nb = 64 - (int)sh;
        idx += 2;
        pos = p;
    }
    __device__ __forceinline__ void refill() {
        if (nb <= 32) {
            buf |= (uint64_t)w[idx++] << nb;
            nb += 32;
        }
    }
    __device__ __forceinline__ void skip(uint32_t n) {
        buf >>= n;
        nb -= (int)n;
        pos += n;
    }
};

// Decode one codeword (codes <= 32 bits: buf holds >= 33 bits after refill).
__device__ __forceinline__ uint32_t decode_short(BitReader& br, const uint32_t* __restrict__ lds_tab,
                                                 const uint32_t* __restrict__ gtab, uint32_t root_bits,
                                                 uint32_t& sym) {
    br.refill();
    uint64_t b = br.buf;
    uint32_t e = lds_tab[b & ((1u << root_bits) - 1)];
    uint32_t consumed = 0, bits = root_bits;
    while (e & 0x80000000u) {
        consumed += bits;
        b >>= bits;
        const uint32_t* t = gtab + (e & 0xFFFFFF);
        bits = (e >> 24) & 0x1F;
        e = t[b & ((1u << bits) - 1)];
    }
    sym = e & 0xFFFFFF;
    const uint32_t len = consumed + ((e >> 24) & 0x7F);
    br.skip(len);
    return len;
}

__device__ __forceinline__ uint64_t window64(const uint32_t* __restrict__ w, uint64_t p) {
    const uint64_t i = p >> 5;
    const uint32_t sh = (uint32_t)(p & 31);
    uint64_t lo = (uint64_t)w[i] | ((uint64_t)w[i + 1] << 32);
    if (sh == 0) return lo;
    return (lo >> sh) | ((uint64_t)w[i + 2] << (64 - sh));
}

// Codes up to 64 bits: a fresh 64-bit window per codeword.
__device__ __forceinline__ uint32_t decode_long(const uint32_t* __restrict__ words, uint64_t p,
                                                const uint32_t* __restrict__ gtab, uint32_t root_bits,
                                                uint32_t& sym) {
    uint64_t b = window64(words, p);
    uint32_t e = gtab[b & ((1u << root_bits) - 1)];
    uint32_t consumed = 0, bits = root_bits;
    while (e & 0x80000000u) {
        consumed += bits;
        b >>= bits;
        const uint32_t* t = gtab + (e & 0xFFFFFF);
        bits = (e >> 24) & 0x1F;
        e = t[b & ((1u << bits) - 1)];
    }
    sym = e & 0xFFFFFF;
    return consumed + ((e >> 24) & 0x7F);
}

// Parser over one array: short codes through the LDS table and a bit buffer,
// long codes through global windows.
struct Parser {
    BitReader br;
    uint64_t p;
    bool lng;
    const uint32_t* lds;
    const uint32_t* gtab;
    uint32_t rb;
    __device__ __forceinline__ void init(const uint32_t* words, uint64_t start, bool long_codes, const uint32_t* l,
                                         const uint32_t* g, uint32_t root_bits) {
        lng = long_codes;
        lds = l;
        gtab = g;
        rb = root_bits;
        p = start;
        if (!lng) br.init(words, start);
        else br.w = words;
    }
    __device__ __forceinline__ uint32_t next() {
        uint32_t sym;
        if (!lng) {
            decode_short(br, lds, gtab, rb, sym);
            p = br.pos;
        } else {
            p += decode_long(br.w, p, gtab, rb, sym);
        }
        return sym;
    }
};

__device__ __forceinline__ void load_root(uint32_t* lds, const DecArrayDev& A, const uint32_t* __restrict__ tables) {
    const uint32_t n = 1u << A.root_bits;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) lds[i] = tables[A.table_off + i];
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_dec_sync(const uint32_t* __restrict__ words,
                                                  const DecArrayDev* __restrict__ arrays,
                                                  const uint32_t* __restrict__ wg_arr,
                                                  const uint32_t* __restrict__ tables, uint64_t* start,
                                                  uint64_t* exit_pos, uint32_t* count) {
    __shared__ uint32_t lds[kRootMax];
    const DecArrayDev A = arrays[wg_arr[blockIdx.x]];
    load_root(lds, A, tables);
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t u = j - A.sub0;
    if (u >= A.nsub) {
        start[j] = exit_pos[j] = 0;
        count[j] = 0;
        return;
    }
    const uint64_t lim = A.payload_bit + min((u + 1) * kSubBits, A.bits);
    const uint64_t s = A.payload_bit + u * kSubBits;
    Parser P;
    P.init(words, s, A.max_len > 32, lds, tables + A.table_off, A.root_bits);
    uint32_t n = 0;
    while (P.p < lim) {
        P.next();
        n++;
    }
    start[j] = s;
    exit_pos[j] = P.p;
    count[j] = n;
}

__global__ __launch_bounds__(256) void k_dec_fix(const uint32_t* __restrict__ words,
                                                 const DecArrayDev* __restrict__ arrays,
                                                 const uint32_t* __restrict__ wg_arr,
                                                 const uint32_t* __restrict__ tables, uint64_t* start,
                                                 const uint64_t* __restrict__ exit_in, uint64_t* exit_out,
                                                 uint32_t* count, unsigned int* changed) {
    __shared__ uint32_t lds[kRootMax];
    const DecArrayDev A = arrays[wg_arr[blockIdx.x]];
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t u = j - A.sub0;
    const bool settled = u >= A.nsub || u == 0 || start[j] == exit_in[j - 1];
    // the table only when some lane re-parses (after convergence: none)
    if (!__syncthreads_or(!settled)) {
        exit_out[j] = exit_in[j];
        return;
    }
    load_root(lds, A, tables);
    if (settled) {
        exit_out[j] = exit_in[j];
        return;
    }
    const uint64_t lim = A.payload_bit + min((u + 1) * kSubBits, A.bits);
    const uint64_t ns = exit_in[j - 1];
    const bool lng = A.max_len > 32;
    Parser pn, po;
    pn.init(words, ns, lng, lds, tables + A.table_off, A.root_bits);
    po.init(words, start[j], lng, lds, tables + A.table_off, A.root_bits);
    uint32_t dn = 0, dold = 0;
    uint64_t new_exit;
    uint32_t new_count;
    for (;;) {
        if (pn.p >= lim) {  // the new parse left the subsequence without meeting the old one
            new_exit = pn.p;
            new_count = dn;
            break;
        }
        if (pn.p == po.p) {  // merged: the rest is the old parse
            new_exit = exit_in[j];
            new_count = dn + (count[j] - dold);
            break;
        }
        if (po.p < pn.p && po.p < lim) {
            po.next();
            dold++;
        } else {
            pn.next();
            dn++;
        }
    }
    start[j] = ns;
    count[j] = new_count;
    exit_out[j] = new_exit;
    if (new_exit != exit_in[j]) atomicOr(changed, 1u);
}

// Symbols per workgroup.
__global__ __launch_bounds__(256) void k_dec_wgsum(const uint32_t* __restrict__ count, uint64_t* wg_sum) {
    __shared__ uint32_t red[4];
    uint32_t v = count[(uint64_t)blockIdx.x * blockDim.x + threadIdx.x];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) wg_sum[blockIdx.x] = (uint64_t)red[0] + red[1] + red[2] + red[3];
}

// Exclusive scan of workgroup sums per array (arrays own contiguous workgroup ranges).
__global__ __launch_bounds__(1024) void k_dec_wgscan(const DecArrayDev* __restrict__ arrays, uint32_t narrays,
                                                     const uint64_t* __restrict__ wg_sum, uint64_t* wg_off,
                                                     unsigned int* err) {
    __shared__ uint64_t buf[1024];
    for (uint32_t k = 0; k < narrays; k++) {
        const DecArrayDev A = arrays[k];
        uint64_t carry = A.out_off;
        for (uint64_t base = 0; base < A.nwg; base += 1024) {
            const uint64_t i = base + threadIdx.x;
            const uint64_t v = i < A.nwg ? wg_sum[A.wg0 + i] : 0;
            buf[threadIdx.x] = v;
            __syncthreads();
            for (int o = 1; o < 1024; o <<= 1) {
                const uint64_t add = threadIdx.x >= (uint32_t)o ? buf[threadIdx.x - o] : 0;
                __syncthreads();
                buf[threadIdx.x] += add;
                __syncthreads();
            }
            if (i < A.nwg) wg_off[A.wg0 + i] = carry + buf[threadIdx.x] - v;
            const uint64_t tot = buf[1023];
            __syncthreads();
            carry += tot;
        }
        if (threadIdx.x == 0 && carry - A.out_off != A.count) atomicOr(err, 1u);
        __syncthreads();
    }
}

constexpr int kStage = 32;

__global__ __launch_bounds__(256) void k_dec_write(const uint32_t* __restrict__ words,
                                                   const DecArrayDev* __restrict__ arrays,
                                                   const uint32_t* __restrict__ wg_arr,
                                                   const uint32_t* __restrict__ tables,
                                                   const uint64_t* __restrict__ start,
                                                   const uint32_t* __restrict__ count,
                                                   const uint64_t* __restrict__ wg_off, int32_t* out,
                                                   unsigned int* err) {
    __shared__ uint32_t lds[kRootMax];
    __shared__ int32_t stage[256][kStage + 1];
    __shared__ uint32_t nh[256];
    __shared__ uint64_t pos[256];
    __shared__ uint32_t scan[256];
    const DecArrayDev A = arrays[wg_arr[blockIdx.x]];
    load_root(lds, A, tables);
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t u = j - A.sub0;
    const bool live = u < A.nsub;
    uint32_t remaining = live ? count[j] : 0;
    // exclusive scan of counts inside the workgroup
    scan[threadIdx.x] = remaining;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        uint32_t add = threadIdx.x >= (uint32_t)o ? scan[threadIdx.x - o] : 0;
        __syncthreads();
        scan[threadIdx.x] += add;
        __syncthreads();
    }
    pos[threadIdx.x] = wg_off[blockIdx.x] + scan[threadIdx.x] - remaining;
    const uint64_t oend = A.out_off + A.count;
    Parser P;
    if (live) P.init(words, start[j], A.max_len > 32, lds, tables + A.table_off, A.root_bits);
    uint64_t written = 0;
    for (;;) {
        const uint32_t n = min<uint32_t>(remaining, kStage);
        for (uint32_t k = 0; k < n; k++) stage[threadIdx.x][k] = (int32_t)(A.offset + (int64_t)P.next());
        nh[threadIdx.x] = n;
        remaining -= n;
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < 256 * kStage; e += 256) {
            const uint32_t l = e / kStage, k = e - l * kStage;
            if (k < nh[l]) {
                const uint64_t o = pos[l] + written + k;
                if (o < oend) out[o] = stage[l][k];
                else atomicOr(err, 2u);
            }
        }
        written += kStage;
        if (!__syncthreads_or(remaining > 0)) break;
    }
    if (live && u + 1 == A.nsub && P.p != A.payload_bit + A.bits) atomicOr(err, 4u);  // ends inside a code
}

}  // namespace dev

void launch_dec_sync(const uint32_t* words, const DecArrayDev* arrays, const uint32_t* wg_arr, uint64_t nwg,
                     const uint32_t* tables, uint64_t* start, uint64_t* exit_pos, uint32_t* count, hipStream_t st) {
    if (!nwg) return;
    hipLaunchKernelGGL(dev::k_dec_sync, dim3((uint32_t)nwg), dim3(256), 0, st, words, arrays, wg_arr, tables, start,
                       exit_pos, count);
}
void launch_dec_fix(const uint32_t* words, const DecArrayDev* arrays, const uint32_t* wg_arr, uint64_t nwg,
                    const uint32_t* tables, uint64_t* start, const uint64_t* exit_in, uint64_t* exit_out,
                    uint32_t* count, unsigned int* changed, hipStream_t st) {
    if (!nwg) return;
    hipLaunchKernelGGL(dev::k_dec_fix, dim3((uint32_t)nwg), dim3(256), 0, st, words, arrays, wg_arr, tables, start,
                       exit_in, exit_out, count, changed);
}
void launch_dec_offsets(const DecArrayDev* arrays, uint32_t narrays, uint64_t nwg, const uint32_t* count,
                        uint64_t* wg_sum, uint64_t* wg_off, unsigned int* err, hipStream_t st) {
    if (!nwg) return;
    hipLaunchKernelGGL(dev::k_dec_wgsum, dim3((uint32_t)nwg), dim3(256), 0, st, count, wg_sum);
    hipLaunchKernelGGL(dev::k_dec_wgscan, dim3(1), dim3(1024), 0, st, arrays, narrays, wg_sum, wg_off, err);
}
void launch_dec_write(const uint32_t* words, const DecArrayDev* arrays, const uint32_t* wg_arr, uint64_t nwg,
                      const uint32_t* tables, const uint64_t* start, const uint32_t* count, const uint64_t* wg_off,
                      int32_t* out, unsigned int* err, hipStream_t st) {
    if (!nwg) return;
    hipLaunchKernelGGL(dev::k_dec_write, dim3((uint32_t)nwg), dim3(256), 0, st, words, arrays, wg_arr, tables, start,
                       count, wg_off, out, err);
}

}  // namespace asz
