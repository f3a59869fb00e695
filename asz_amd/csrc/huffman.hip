// GPU side of the Huffman coder for the .asz stream (huffman.hpp).
//
// Histogram.  The round kernels leave one LDS window histogram per workgroup
// (symbols [-win/2, win/2), win chosen per round) and add the rest of their
// QBs to the array's global window (kHistGlobal bins around 0) with atomics.
// k_fold_round, launched right after each round, folds that round's windows
// into the global window (each workgroup of the fold owns a slice of the
// windows, so the global atomics are few); k_hist_small bins the two short
// arrays (t values and pivots) the same way.  The workgroup that finishes an
// array writes its QB range, errors and used bins straight into pinned host
// memory and raises a flag, so the host builds that codebook while later
// rounds run, without a copy or an API call in between.  An array whose range
// leaves the global window is binned again by k_hist over its exact range.
//
// Encoding.  The host builds the reference-tree codebooks.  k_encode codes
// the arrays in one pass: chunks of kEncChunk symbols are taken by ticket,
// coded from an LDS copy of the codebook when it is small, and each chunk
// finds its bit position in the stream by a decoupled look-back over the bit
// counts its predecessors publish, then writes its codes there directly
// (the stream buffer already holds the header bytes): interior words are
// stored, the two boundary words shared with the neighbours are ORed in.
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
constexpr int kEncPad = kEncChunk + kEncChunk / kEncPerThread;  // staging with one pad word per 16
constexpr int kEncWords = kEncChunk + 2;        // LDS bit buffer for codes up to 32 bits

// Look-back status of a chunk: flag in bits 62..63 (0 not yet published,
// 1 the chunk's own bit count, 2 its inclusive bit prefix within its array),
// the count in bits 0..61.
constexpr uint64_t kLbAggregate = 1ull << 62, kLbPrefix = 2ull << 62, kLbValue = (1ull << 62) - 1;
// Polls of one predecessor's status before the encoder gives up and raises
// its error flag (far beyond any real wait: a predecessor publishes its count
// right after its own scan).
constexpr uint32_t kLbSpinLimit = 1u << 22;

__device__ __forceinline__ void lb_publish(unsigned long long* status, uint64_t c, uint64_t w) {
    __hip_atomic_store(status + c, (unsigned long long)w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t lb_load(unsigned long long* status, uint64_t c) {
    return __hip_atomic_load(status + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back (one wave): publishes chunk c's bit count, sums the
// counts of its predecessors within its array (first = the array's first
// chunk) back to the nearest one that has published its inclusive prefix,
// publishes c's own inclusive prefix and returns its exclusive one.  Lane l
// inspects predecessor j - l, 64 predecessors per step.  Every predecessor
// holds a smaller ticket, so it belongs to a workgroup that is already
// running and publishes its count before it looks back itself: the wait
// always ends.
__device__ uint64_t lookback(unsigned long long* status, uint64_t c, uint64_t first, uint64_t total, uint32_t lane,
                             unsigned int* err) {
    if (c == first) {
        if (lane == 0) lb_publish(status, c, kLbPrefix | total);
        return 0;
    }
    if (lane == 0) lb_publish(status, c, kLbAggregate | total);
    uint64_t excl = 0;
    int64_t j = (int64_t)c - 1;
    while (true) {
        const int64_t idx = j - (int64_t)lane;
        uint64_t w = kLbPrefix;  // before the array's first chunk: an inclusive prefix of 0
        if (idx >= (int64_t)first) {
            w = lb_load(status, (uint64_t)idx);
            for (uint32_t it = 0; (w >> 62) == 0; it++) {
                if (it >= kLbSpinLimit) {
                    atomicOr(err, 1u);
                    w = kLbPrefix;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                w = lb_load(status, (uint64_t)idx);
            }
        }
        const uint64_t pm = __ballot((w >> 62) == 2);
        const uint64_t val = w & kLbValue;
        if (pm) {
            const uint32_t lp = (uint32_t)__ffsll((unsigned long long)pm) - 1;
            excl += wave_sum<uint64_t>(lane <= lp ? val : 0ull);
            break;
        }
        excl += wave_sum<uint64_t>(val);
        j -= kWave;
    }
    if (lane == 0) lb_publish(status, c, kLbPrefix | (excl + total));
    return excl;
}

// Single-pass encoder.  Each workgroup takes chunks of kEncChunk symbols by
// ticket (an atomic counter, so a chunk's predecessors always belong to
// running workgroups) until the tickets run out, with the codebook of the
// current array in LDS when it is small; the symbols of its next chunk are
// loaded (coalesced: lane l of step j reads symbol 256 j + l) while the
// current one is coded.  Per chunk: codebook entries are staged in LDS so
// that thread t owns the 16 consecutive symbols 16 t .. 16 t + 15 (padded
// rows, no bank conflicts); a workgroup scan of the per-thread bit counts
// gives the chunk's bit count, and the look-back its bit position in the
// stream.  Each thread then packs its codes into a 64-bit accumulator at
// that position's offset within a 32-bit word and emits whole words into an
// LDS bit buffer (only the words it shares with its neighbours are ORed in);
// the buffer goes to the stream buffer, interior words stored and the two
// boundary words (shared with the neighbouring chunks or the array's header
// bytes) ORed in.  Codes longer than 32 bits (short_codes false) are ORed
// straight into the stream.  The kernel is held to 128 VGPRs so that four
// workgroups share a CU.  lb: [0] ticket counter, [1] error flag, [2 + c]
// chunk c's status; zeroed, like the stream buffer, before the launch.
template <bool short_codes>
__global__ __launch_bounds__(256, 4) void k_encode(const int32_t* __restrict__ sym,
                                                   const HuffArrayDev* __restrict__ arrays, uint32_t narr,
                                                   uint64_t nchunks, const uint64_t* __restrict__ book,
                                                   uint32_t* words, unsigned long long* lb) {
    __shared__ uint32_t lbuf[kEncPad > kEncWords ? kEncPad : kEncWords];  // staging, then the bit buffer
    __shared__ uint32_t lbook[kLdsBook];
    __shared__ uint32_t s_scan[4];
    __shared__ unsigned long long s_tk[2];
    __shared__ unsigned long long s_excl;
    unsigned long long* status = lb + 2;
    unsigned int* err = reinterpret_cast<unsigned int*>(lb + 1);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // tickets of one workgroup increase, so its array index only moves forward
    // (the table stays in global memory: one call may code many fields' arrays)
    auto array_of = [&](uint64_t c, uint32_t k) {
        while (k + 1 < narr && c >= arrays[k].chunk0 + arrays[k].nchunk) k++;
        return k;
    };
    auto load_chunk = [&](uint64_t c, uint32_t k, int32_t v[kEncPerThread]) {
        const HuffArrayDev A = arrays[k];
        const uint64_t first = (c - A.chunk0) * kEncChunk;
        const uint32_t n = (uint32_t)min<uint64_t>(kEncChunk, A.count - first);
        const int32_t* s = sym + A.src_off + first;
#pragma unroll
        for (int j = 0; j < kEncPerThread; j++) {
            const uint32_t i = j * 256 + threadIdx.x;
            v[j] = i < n ? s[i] : 0;
        }
    };
    if (threadIdx.x == 0) s_tk[0] = atomicAdd(lb, 1ull);
    __syncthreads();
    uint64_t c = s_tk[0];
    if (c >= nchunks) return;
    uint32_t k = array_of(c, 0);
    int32_t v[kEncPerThread];
    if (short_codes) load_chunk(c, k, v);
    int book_k = -1;
    while (true) {
        if (threadIdx.x == 0) s_tk[1] = atomicAdd(lb, 1ull);  // the next chunk, read after the next barrier
        const HuffArrayDev A = arrays[k];
        const uint64_t* b = book + A.book_off;
        const uint64_t first = (c - A.chunk0) * kEncChunk;
        const uint32_t n = (uint32_t)min<uint64_t>(kEncChunk, A.count - first);
        uint64_t e[kEncPerThread];
        uint32_t mybits = 0;
        uint64_t cn;
        uint32_t kn;
        if (short_codes) {
            if (A.lds_book && book_k != (int)k) {
                for (uint32_t i = threadIdx.x; i < A.range; i += blockDim.x) {
                    const uint64_t be = b[i];
                    lbook[i] = (uint32_t)(be & 0xFFFFFFu) | (book_len(be) << 24);
                }
                book_k = (int)k;
            }
            lds_barrier();  // LDS codebook complete; the previous chunk's buffer is free; next ticket in LDS
            cn = s_tk[1];
            kn = cn < nchunks ? array_of(cn, k) : k;
#pragma unroll
            for (int j = 0; j < kEncPerThread; j++) {
                const uint32_t i = j * 256 + threadIdx.x;
                uint32_t ent = 0;
                if (i < n) {
                    const int64_t idx = (int64_t)v[j] - A.offset;
                    if (A.lds_book) {
                        ent = lbook[idx];
                    } else {
                        // codes of 25..32 bits keep the symbol index, resolved after staging
                        const uint64_t be = b[idx];
                        ent = book_len(be) <= 24 ? ((uint32_t)(be & 0xFFFFFFu) | (book_len(be) << 24))
                                                 : (0x80000000u | (uint32_t)idx);
                    }
                }
                lbuf[i + i / kEncPerThread] = ent;
            }
            // the next chunk's symbols stay in flight while this one is packed
            // (LDS-only barriers below do not wait for them)
            if (cn < nchunks) load_chunk(cn, kn, v);
            lds_barrier();
#pragma unroll
            for (int j = 0; j < kEncPerThread; j++) {
                const uint32_t i = threadIdx.x * kEncPerThread + j;
                const uint32_t ent = lbuf[i + i / kEncPerThread];
                uint64_t v64;
                if (ent & 0x80000000u) v64 = b[ent & 0x7FFFFFFFu];
                else v64 = (uint64_t)(ent & 0xFFFFFFu) | ((uint64_t)(ent >> 24) << 56);
                e[j] = v64;
                mybits += book_len(v64);
            }
        } else {
            lds_barrier();
            cn = s_tk[1];
            kn = cn < nchunks ? array_of(cn, k) : k;
            const uint32_t my0 = threadIdx.x * kEncPerThread;
            const int32_t* s = sym + A.src_off + first;
#pragma unroll
            for (int j = 0; j < kEncPerThread; j++) {
                const uint32_t i = my0 + j;
                uint64_t v64 = 0;
                if (i < n) v64 = b[(int64_t)s[i] - A.offset];
                e[j] = v64;
                mybits += book_len(v64);
            }
        }
        // exclusive scan of the per-thread bit counts
        uint32_t incl = mybits;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, 64);
            if (lane >= (uint32_t)o) incl += t;
        }
        if (lane == 63) s_scan[wave] = incl;
        lds_barrier();  // also: every thread has read its staged entries
        uint32_t wave_off = 0;
        for (uint32_t w = 0; w < wave; w++) wave_off += s_scan[w];
        const uint32_t total = s_scan[0] + s_scan[1] + s_scan[2] + s_scan[3];
        const uint32_t my_off = wave_off + incl - mybits;
        if (wave == 0) {
            const uint64_t ex = lookback(status, c, A.chunk0, total, lane, err);
            if (lane == 0) s_excl = ex;
        }
        lds_barrier();
        const uint64_t gbit = A.payload_bit + s_excl;  // the chunk's first bit in the stream
        const uint32_t sh = (uint32_t)(gbit & 31);
        uint32_t* out = words + (gbit >> 5);
        if (short_codes) {
            const uint32_t nw = (sh + total + 31) / 32;
            for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x) lbuf[w] = 0;
            lds_barrier();
            if (mybits) {
                const uint32_t p0 = sh + my_off;
                uint32_t word = p0 >> 5;
                uint32_t nbits = p0 & 31;
                const bool shared_first = nbits != 0;
                bool first_word = true;
                uint64_t accb = 0;
#pragma unroll
                for (int j = 0; j < kEncPerThread; j++) {
                    accb |= book_code(e[j]) << nbits;
                    nbits += book_len(e[j]);
                    if (nbits >= 32) {  // at most one whole word per code (nbits < 64)
                        const uint32_t lo = (uint32_t)accb;
                        if (first_word && shared_first) atomicOr(&lbuf[word], lo);
                        else lbuf[word] = lo;
                        first_word = false;
                        word++;
                        accb >>= 32;
                        nbits -= 32;
                    }
                }
                if (nbits) atomicOr(&lbuf[word], (uint32_t)accb);
            }
            lds_barrier();
            for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x) {
                const uint32_t val = lbuf[w];
                if (w == 0 || w + 1 == nw) atomicOr(&out[w], val);
                else out[w] = val;
            }
        } else {
            uint64_t p = (uint64_t)sh + my_off;
            for (int j = 0; j < kEncPerThread; j++) {
                const uint32_t len = book_len(e[j]);
                if (!len) continue;
                const uint64_t code = book_code(e[j]);
                const uint32_t s2 = (uint32_t)(p & 31);
                uint32_t* w = out + (p >> 5);
                atomicOr(&w[0], (uint32_t)(code << s2));
                if (s2 + len > 32) atomicOr(&w[1], (uint32_t)(code >> (32 - s2)));
                if (s2 + len > 64) atomicOr(&w[2], (uint32_t)(code >> (64 - s2)));
                p += len;
            }
        }
        if (cn >= nchunks) break;
        c = cn;
        k = kn;
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
uint64_t encode_lb_words(uint64_t nchunks) { return 2 * (nchunks + 2); }

void launch_encode(const int32_t* sym, const HuffArrayDev* arrays, uint32_t narr, uint64_t nchunks,
                   const uint64_t* book, uint32_t* stream_words, uint32_t* lb_words, int max_len, hipStream_t st) {
    if (!nchunks) return;
    // one resident wave of workgroups: CUs x workgroups that fit on a CU
    int dev_id = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev_id);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id);
    if (max_len <= 32)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dev::k_encode<true>, 256, 0);
    else
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dev::k_encode<false>, 256, 0);
    const uint64_t resident = (uint64_t)std::max(cus, 1) * (uint64_t)std::max(per_cu, 1);
    const uint64_t groups = std::min<uint64_t>(nchunks, std::min<uint64_t>(resident, kEncGroups));
    unsigned long long* lb = reinterpret_cast<unsigned long long*>(lb_words);
    if (max_len <= 32)
        hipLaunchKernelGGL(dev::k_encode<true>, dim3((uint32_t)groups), dim3(256), 0, st, sym, arrays, narr, nchunks,
                           book, stream_words, lb);
    else
        hipLaunchKernelGGL(dev::k_encode<false>, dim3((uint32_t)groups), dim3(256), 0, st, sym, arrays, narr, nchunks,
                           book, stream_words, lb);
}

}  // namespace asz
