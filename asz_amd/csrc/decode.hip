// Parallel Huffman decoding of the .asz bit streams (huffman.hpp:383-526).
//
// The format has one bit stream per array and no sync points.  Each array's
// payload is cut into subsequences of kSubBits bits; a subsequence owns the
// codewords that start inside it.  Workgroups hold 256 consecutive
// subsequences of one array, with that array's root lookup table in LDS.
//   1. k_dec_sync: decode every subsequence from a little before its first
//      bit, taking the first codeword boundary at or past that bit as its
//      start (a guess, nearly always right), recording the start, the exit
//      (first codeword boundary past its end) and the codeword count.
//   2. k_dec_fix (repeated until no exit changes): a subsequence whose start
//      differs from its predecessor's exit re-parses from that exit, walking
//      its old parse in lockstep; Huffman parses from different starts merge
//      after a few codewords, after which count and exit follow from the old
//      parse.  Subsequence 0 of each array starts at a true boundary, so the
//      chain converges.
//   3. k_dec_wgsum / k_dec_wgscan: output offsets (two-level scan).
//   4. k_dec_write: decode again from the exact starts, each thread storing
//      whole 128-byte lines of the output straight from registers.
#include "device.hpp"
#include "launch.hpp"

namespace asz {
namespace dev {

constexpr int kRootMax = 1 << kRootBits;
constexpr uint64_t kSyncLead = 96;  // bits the first parse of a subsequence starts early

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
    const uint64_t s0 = A.payload_bit + u * kSubBits;
    // the parse starts kSyncLead bits early, inside the predecessor: by the
    // subsequence's first bit it has almost always fallen into step with the
    // true codeword boundaries, and the first boundary at or past that bit is
    // the start (the fix passes correct the rest)
    const uint64_t lead = min<uint64_t>(kSyncLead, u * kSubBits);
    Parser P;
    P.init(words, s0 - lead, A.max_len > 32, lds, tables + A.table_off, A.root_bits);
    while (P.p < s0) P.next();
    const uint64_t s = P.p;
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

// Output slot of every decoder workgroup's first symbol.  One workgroup
// scans the workgroup sums of all arrays at once (thread t owns a run of
// consecutive sums: it sums the run, the workgroup scans the 1024 run sums,
// and the thread walks its run again), then rebases them per array (arrays
// own contiguous workgroup ranges): wg_off[w] = out_off + X[w] - X[wg0] with
// X the exclusive scan over all arrays.  The last workgroup of each array
// checks the array's total (error bit 1).
constexpr uint32_t kDecMaxArrays = 128;
__global__ __launch_bounds__(1024) void k_dec_wgscan(const DecArrayDev* __restrict__ arrays,
                                                     const uint32_t* __restrict__ wg_arr, uint64_t nwg,
                                                     const uint64_t* __restrict__ wg_sum, uint64_t* wg_off,
                                                     unsigned int* err) {
    __shared__ uint64_t wsum[16];
    __shared__ uint64_t base[kDecMaxArrays];  // X at each array's first workgroup
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t per = (nwg + 1023) / 1024;
    const uint64_t r0 = min<uint64_t>(nwg, threadIdx.x * per), r1 = min<uint64_t>(nwg, r0 + per);
    uint64_t mine = 0;
    for (uint64_t w = r0; w < r1; w++) mine += wg_sum[w];
    uint64_t incl = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t t = __shfl_up(incl, o, 64);
        if (lane >= (uint32_t)o) incl += t;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint64_t x0 = incl - mine;
    for (uint32_t w = 0; w < wave; w++) x0 += wsum[w];
    uint64_t x = x0;
    for (uint64_t w = r0; w < r1; w++) {
        const uint32_t k = wg_arr[w];
        if (w == arrays[k].wg0) base[k] = x;
        x += wg_sum[w];
    }
    __syncthreads();
    x = x0;
    for (uint64_t w = r0; w < r1; w++) {
        const uint32_t k = wg_arr[w];
        const DecArrayDev A = arrays[k];
        wg_off[w] = A.out_off + x - base[k];
        x += wg_sum[w];
        if (w + 1 == A.wg0 + A.nwg && x - base[k] != A.count) atomicOr(err, 1u);
    }
}

// Write pass.  Each thread decodes its subsequence again from the exact start;
// its symbols' output slots start at pos (the workgroup's slot plus an
// exclusive scan of the counts inside the workgroup).  Every 128-byte line of
// the output (kLineSyms slots; arrays start on line boundaries) is stored by
// one thread, whole: thread t stores the slots [roundup(pos_t),
// roundup(pos_t + n_t)), so it skips its symbols before its first line
// boundary (the previous owner stores them) and, past its own symbols, parses
// on into its successors' first symbols up to the next boundary (the bit
// stream is continuous and the starts are exact), at most kLineSyms - 1 more
// codewords.  A line is decoded into the thread's own LDS row and then stored
// with eight 16-byte stores back to back (a line written piecemeal over a
// long decode leaves half-written lines in L2, which reach HBM as partial
// writes).  Errors: a run past its array's end (2), the parse of an
// array's last subsequence not ending at the payload's end (4).
constexpr uint32_t kLineSyms = 32;
__global__ __launch_bounds__(256) void k_dec_write(const uint32_t* __restrict__ words,
                                                   const DecArrayDev* __restrict__ arrays,
                                                   const uint32_t* __restrict__ wg_arr,
                                                   const uint32_t* __restrict__ tables,
                                                   const uint64_t* __restrict__ start,
                                                   const uint32_t* __restrict__ count,
                                                   const uint64_t* __restrict__ wg_off, int32_t* out,
                                                   unsigned int* err) {
    __shared__ uint32_t lds[kRootMax];
    __shared__ uint32_t wsum[4];
    __shared__ int32_t lrow[256][kLineSyms + 1];
    const DecArrayDev A = arrays[wg_arr[blockIdx.x]];
    load_root(lds, A, tables);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t u = j - A.sub0;
    const bool live = u < A.nsub;
    const uint32_t n = live ? count[j] : 0;
    // exclusive scan of the counts inside the workgroup
    uint32_t incl = n;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o, 64);
        if (lane >= (uint32_t)o) incl += t;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint64_t pos = wg_off[blockIdx.x] + incl - n;
    for (uint32_t w = 0; w < wave; w++) pos += wsum[w];
    if (!live) return;
    const uint64_t oend = A.out_off + A.count;
    if (pos + n > oend) {
        atomicOr(err, 2u);
        return;
    }
    auto up = [](uint64_t x) { return (x + kLineSyms - 1) / kLineSyms * kLineSyms; };
    const uint64_t a = min(up(pos), oend), b = min(up(pos + n), oend);  // this thread's slots [a, b)
    Parser P;
    P.init(words, start[j], A.max_len > 32, lds, tables + A.table_off, A.root_bits);
    if (b <= a) {  // no line starts among this thread's slots: nothing to store
        if (u + 1 == A.nsub) {
            for (uint32_t k = 0; k < n; k++) P.next();
            if (P.p != A.payload_bit + A.bits) atomicOr(err, 4u);
        }
        return;
    }
    const int64_t off = A.offset;
    for (uint64_t k = pos; k < a; k++) P.next();
    int32_t* row = &lrow[threadIdx.x][0];  // this thread's line (padded rows: no bank conflicts)
    uint64_t k = a;
    for (; k + kLineSyms <= b; k += kLineSyms) {
        for (uint32_t i = 0; i < kLineSyms; i++) row[i] = (int32_t)(off + (int64_t)P.next());
        int4* d = reinterpret_cast<int4*>(out + k);
#pragma unroll
        for (uint32_t i = 0; i < kLineSyms / 4; i++) d[i] = make_int4(row[4 * i], row[4 * i + 1], row[4 * i + 2], row[4 * i + 3]);
    }
    for (; k < b; k++) out[k] = (int32_t)(off + (int64_t)P.next());  // the partial line at an array's end
    // the array's last run ends at the payload's end (b = oend = pos + n)
    if (u + 1 == A.nsub && P.p != A.payload_bit + A.bits) atomicOr(err, 4u);
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
void launch_dec_offsets(const DecArrayDev* arrays, uint32_t narrays, const uint32_t* wg_arr, uint64_t nwg,
                        const uint32_t* count, uint64_t* wg_sum, uint64_t* wg_off, unsigned int* err, hipStream_t st) {
    if (!nwg) return;
    if (narrays > dev::kDecMaxArrays) fail(ASZ_ERR_DATA, "too many Huffman arrays in one stream");
    hipLaunchKernelGGL(dev::k_dec_wgsum, dim3((uint32_t)nwg), dim3(256), 0, st, count, wg_sum);
    hipLaunchKernelGGL(dev::k_dec_wgscan, dim3(1), dim3(1024), 0, st, arrays, wg_arr, nwg, wg_sum, wg_off, err);
}
void launch_dec_write(const uint32_t* words, const DecArrayDev* arrays, const uint32_t* wg_arr, uint64_t nwg,
                      const uint32_t* tables, const uint64_t* start, const uint32_t* count, const uint64_t* wg_off,
                      int32_t* out, unsigned int* err, hipStream_t st) {
    if (!nwg) return;
    hipLaunchKernelGGL(dev::k_dec_write, dim3((uint32_t)nwg), dim3(256), 0, st, words, arrays, wg_arr, tables, start,
                       count, wg_off, out, err);
}

}  // namespace asz
