// Parallel Huffman decoding of the .asz bit streams (huffman.hpp:383-526).
//
// The format has one bit stream per array and no sync points.  Each array's
// payload is cut into subsequences of kSubBits bits; a subsequence owns the
// codewords that start inside it.  Workgroups hold 256 consecutive
// subsequences of one array, with that array's root lookup table in LDS.
//   1. k_dec_sync: parse every subsequence from a little before its first
//      bit (two codewords per lookup where the root bits hold both), taking
//      the first codeword boundary at or past that bit as its start (a
//      guess, nearly always right), recording the start, the exit (first
//      codeword boundary past its end) and the codeword count.
//   2. k_dec_fix (repeated until no exit changes; inside a workgroup each
//      pass repeats until its subsequences agree): a subsequence whose start
//      differs from its predecessor's exit re-parses from that exit, walking
//      its old parse in lockstep; Huffman parses from different starts merge
//      after a few codewords, after which count and exit follow from the old
//      parse.  Subsequence 0 of each array starts at a true boundary, so the
//      chain converges.
//   3. k_dec_wgsum / k_dec_wgscan: output offsets (two-level scan).
//   4. k_dec_write: decode again from the exact starts, each thread storing
//      whole 128-byte lines of the output straight from registers.
// Every pass reads the workgroup's stretch of the stream from a window in LDS
// (arrays whose codes fit 32 bits).
#include <type_traits>

#include "device.hpp"
#include "launch.hpp"

namespace asz {
namespace dev {

constexpr int kRootMax = 1 << kRootBits;

// Where a parse reads: the stream words, the workgroup's window of them in
// LDS (arrays with codes of at most 32 bits), the root table in LDS and every
// table in global memory.
struct ParseSrc {
    const uint32_t* __restrict__ words;
    const uint32_t* win;  // window rows (kWinRowWords words each)
    uint64_t w0;          // stream word of window word 0
    uint32_t nwin;        // window words loaded
    const uint32_t* lds;  // root table
    const uint32_t* __restrict__ gtab;
    uint32_t rb;          // root bits
};

// The window: the stream words a workgroup's 256 subsequences read, from the
// sync pass's lead before the first one to kWinTail words past the last one
// (a write-pass thread parses up to kLineSyms - 1 codewords of at most 32
// bits past its own).  Row r holds words 16r .. 16r + 16 (the next row's
// first word repeated), so the two words under any bit position lie in one
// row and one ds_read2 fetches them; neighbouring lanes, 16 words apart in
// the stream, land 17 banks apart.
constexpr uint32_t kLineSyms = 32;
constexpr uint32_t kWinTail = 40;
constexpr uint32_t kWinMaxWords = 256 * (kSubBits / 32) + kMaxSyncLead / 32 + kWinTail + 8;
constexpr uint32_t kWinRowWords = 17;
constexpr uint32_t kWinLds = (kWinMaxWords / 16) * kWinRowWords;
static_assert(kWinTail * 32 >= (kLineSyms + 2) * 32, "window tail shorter than a write-pass overrun");

__device__ __forceinline__ void load_window(uint32_t* win, const uint32_t* __restrict__ words, const DecArrayDev& A,
                                            uint64_t& w0, uint32_t& nwin) {
    const uint64_t u0 = (uint64_t)blockIdx.x * blockDim.x - A.sub0;
    const uint64_t b0 = A.payload_bit + u0 * kSubBits;
    w0 = (b0 - min<uint64_t>(A.lead, u0 * kSubBits)) >> 5;
    const uint64_t eb = A.payload_bit + min((u0 + blockDim.x) * kSubBits, A.bits);
    nwin = (uint32_t)min<uint64_t>((eb + 31) / 32 + kWinTail - w0, kWinMaxWords);
    for (uint32_t k = threadIdx.x; k < nwin; k += blockDim.x) {
        const uint32_t v = words[w0 + k];
        const uint32_t r = k >> 4, c = k & 15;
        win[r * kWinRowWords + c] = v;
        if (c == 0 && r) win[(r - 1) * kWinRowWords + 16] = v;
    }
}

// Parsers over one array's bit stream.  Positions are 32-bit, counted from a
// base bit the caller picks (a thread parses a few thousand bits at most).
//
// WinParser (codes of at most 32 bits): the 32 bits at the position come from
// the window (one ds_read2 and a funnel shift), then one lookup in the root
// table in LDS; longer codes continue through the subtables in global memory.
// There is no bit buffer: every step reads at its position, so lanes never
// branch to refill one.  A position past the window (only in a malformed
// stream) reads the window's last words instead.
struct WinParser {
    const uint32_t* win;
    const uint32_t* lds;
    const uint32_t* __restrict__ gtab;
    uint32_t o;     // window bit of the base
    uint32_t rel;   // position, bits past the base
    uint32_t imax;  // last window word a read may start at
    uint32_t rb;

    __device__ __forceinline__ void init(const ParseSrc& S, uint64_t p, uint64_t base) {
        win = S.win;
        lds = S.lds;
        gtab = S.gtab;
        rb = S.rb;
        o = (uint32_t)(base - S.w0 * 32);
        rel = (uint32_t)(p - base);
        imax = S.nwin - 2;
    }
    __device__ __forceinline__ uint32_t next() {
        const uint32_t q = o + rel;
        const uint32_t i = min(q >> 5, imax);
        const uint32_t a = (i >> 4) * kWinRowWords + (i & 15);
        const uint32_t b = __builtin_amdgcn_alignbit(win[a + 1], win[a], q & 31);
        uint32_t e = lds[b & ((1u << rb) - 1)];
        uint32_t len = e >> 24;  // a leaf: bit 31 clear, the length above the symbol
        if (e & 0x80000000u) {
            uint32_t consumed = rb, bb = b >> rb;
            for (;;) {
                const uint32_t bits = (e >> 24) & 0x1F;
                e = gtab[(e & 0xFFFFFF) + (bb & ((1u << bits) - 1))];
                if (!(e & 0x80000000u)) break;
                consumed += bits;
                bb >>= bits;
            }
            len = consumed + (e >> 24);
        }
        rel += len;
        return e & 0xFFFFFF;
    }
    // With the sync pass's root table in lds: one codeword, or two when the
    // root bits hold two whole codes and the second starts before bound.
    // Returns the codewords taken.
    __device__ __forceinline__ uint32_t skip(uint32_t bound) {
        const uint32_t q = o + rel;
        const uint32_t i = min(q >> 5, imax);
        const uint32_t a = (i >> 4) * kWinRowWords + (i & 15);
        const uint32_t b = __builtin_amdgcn_alignbit(win[a + 1], win[a], q & 31);
        uint32_t e = lds[b & ((1u << rb) - 1)];
        if (e & kPairFlag) {
            const uint32_t l1 = (e >> 8) & 0xF;
            const bool both = rel + l1 < bound;
            rel += both ? (e & 0xFF) : l1;
            return both ? 2u : 1u;
        }
        uint32_t len = e >> 24;
        if (e & 0x80000000u) {
            uint32_t consumed = rb, bb = b >> rb;
            for (;;) {
                const uint32_t bits = (e >> 24) & 0x1F;
                e = gtab[(e & 0xFFFFFF) + (bb & ((1u << bits) - 1))];
                if (!(e & 0x80000000u)) break;
                consumed += bits;
                bb >>= bits;
            }
            len = consumed + (e >> 24);
        }
        rel += len;
        return 1u;
    }
};

__device__ __forceinline__ uint64_t window64(const uint32_t* __restrict__ w, uint64_t p) {
    const uint64_t i = p >> 5;
    const uint32_t sh = (uint32_t)(p & 31);
    uint64_t lo = (uint64_t)w[i] | ((uint64_t)w[i + 1] << 32);
    if (sh == 0) return lo;
    return (lo >> sh) | ((uint64_t)w[i + 2] << (64 - sh));
}

// LongParser (codes up to 64 bits): a fresh 64-bit window per codeword from
// the stream words and every table in global memory.
struct LongParser {
    const uint32_t* __restrict__ w;
    const uint32_t* __restrict__ gtab;
    uint64_t base;
    uint32_t rel;
    uint32_t rb;

    __device__ __forceinline__ void init(const ParseSrc& S, uint64_t p, uint64_t b) {
        w = S.words;
        gtab = S.gtab;
        rb = S.rb;
        base = b;
        rel = (uint32_t)(p - b);
    }
    __device__ __forceinline__ uint32_t next() {
        uint64_t b = window64(w, base + rel);
        uint32_t e = gtab[b & ((1u << rb) - 1)];
        uint32_t consumed = 0, bits = rb;
        while (e & 0x80000000u) {
            consumed += bits;
            b >>= bits;
            bits = (e >> 24) & 0x1F;
            e = gtab[(e & 0xFFFFFF) + ((uint32_t)b & ((1u << bits) - 1))];
        }
        rel += consumed + ((e >> 24) & 0x7F);
        return e & 0xFFFFFF;
    }
};

// The array's root table (sync: its two-code form, codes of at most 32 bits)
// and (codes of at most 32 bits) the workgroup's window into LDS, then one
// barrier.
__device__ __forceinline__ ParseSrc load_src(uint32_t* lds, uint32_t* win, const uint32_t* __restrict__ words,
                                             const DecArrayDev& A, const uint32_t* __restrict__ tables,
                                             bool sync = false) {
    ParseSrc S;
    S.words = words;
    S.win = win;
    S.w0 = 0;
    S.nwin = 2;
    S.lds = lds;
    S.gtab = tables + A.table_off;
    S.rb = A.root_bits;
    const uint32_t n = 1u << A.root_bits;
    const uint32_t* root = sync && A.max_len <= 32 ? tables + A.pair_off : S.gtab;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) lds[i] = root[i];
    if (A.max_len <= 32) load_window(win, words, A, S.w0, S.nwin);
    __syncthreads();
    return S;
}

// Re-parse of a subsequence from its predecessor's exit ns, in lockstep with
// its current parse from os (count oc codewords, exit oe); both positions
// count from the lower of the two (a predecessor's exit and every start lie
// at or past the subsequence's first bit).  Returns the new exit; nc the new
// count.  PAIRS: the two-code root table in LDS (one codeword per step).
template <class PR, bool PAIRS = false>
__device__ __forceinline__ uint64_t fix_run(const ParseSrc& S, const DecArrayDev& A, uint64_t u, uint64_t ns,
                                            uint64_t os, uint32_t oc, uint64_t oe, uint32_t& nc) {
    const uint64_t base = min(ns, os);
    const uint64_t lim_abs = A.payload_bit + min((u + 1) * kSubBits, A.bits);
    const uint32_t lim = lim_abs > base ? (uint32_t)(lim_abs - base) : 0u;
    PR pn, po;
    pn.init(S, ns, base);
    po.init(S, os, base);
    uint32_t dn = 0, dold = 0;
    for (;;) {
        if (pn.rel >= lim) {  // the new parse left the subsequence without meeting the old one
            nc = dn;
            return base + pn.rel;
        }
        if (pn.rel == po.rel) {  // merged: the rest is the old parse
            nc = dn + (oc - dold);
            return oe;
        }
        if (po.rel < pn.rel && po.rel < lim) {
            if constexpr (PAIRS) po.skip(0u);
            else po.next();
            dold++;
        } else {
            if constexpr (PAIRS) pn.skip(0u);
            else pn.next();
            dn++;
        }
    }
}

// The parse starts A.lead bits early, inside the predecessor: by the
// subsequence's first bit it has almost always fallen into step with the true
// codeword boundaries, and the first boundary at or past that bit is the
// start (the fix passes correct the rest).
template <class PR>
__device__ __forceinline__ void sync_run(const ParseSrc& S, const DecArrayDev& A, uint64_t u, uint64_t& st,
                                         uint64_t& ex, uint32_t& cnt) {
    const uint32_t lead = (uint32_t)min<uint64_t>(A.lead, u * kSubBits);
    const uint64_t base = A.payload_bit + u * kSubBits - lead;
    const uint32_t lim = lead + (uint32_t)min<uint64_t>(kSubBits, A.bits - u * kSubBits);
    PR P;
    P.init(S, base, base);
    uint32_t n = 0;
    if constexpr (std::is_same<PR, WinParser>::value) {
        while (P.rel < lead) P.skip(lead);
    } else {
        while (P.rel < lead) P.next();
    }
    const uint32_t s = P.rel;
    if constexpr (std::is_same<PR, WinParser>::value) {
        while (P.rel < lim) n += P.skip(lim);
    } else {
        while (P.rel < lim) {
            P.next();
            n++;
        }
    }
    st = base + s;
    ex = base + P.rel;
    cnt = n;
}

__global__ __launch_bounds__(256) void k_dec_sync(const uint32_t* __restrict__ words,
                                                  const DecArrayDev* __restrict__ arrays,
                                                  const uint32_t* __restrict__ wg_arr,
                                                  const uint32_t* __restrict__ tables, uint64_t* start,
                                                  uint64_t* exit_pos, uint32_t* count) {
    __shared__ uint32_t lds[kRootMax];
    __shared__ uint32_t win[kWinLds];
    __shared__ uint64_t lex[256];
    const DecArrayDev A = arrays[wg_arr[blockIdx.x]];
    const ParseSrc S = load_src(lds, win, words, A, tables, true);
    const uint32_t t = threadIdx.x;
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + t;
    const uint64_t u = j - A.sub0;
    const bool in = u < A.nsub;
    uint64_t st = 0, ex = 0;
    uint32_t cnt = 0;
    if (in) {
        if (A.max_len > 32) sync_run<LongParser>(S, A, u, st, ex, cnt);
        else sync_run<WinParser>(S, A, u, st, ex, cnt);
    }
    // The workgroup's own chains settle here, while its tables and window are
    // in LDS (as in a fix pass); a workgroup's first subsequence, whose
    // predecessor another workgroup parses, is left to the fix passes.
    lex[t] = ex;
    __syncthreads();
    const bool live = in && t != 0;
    for (uint32_t it = 0; it < blockDim.x; it++) {
        const uint64_t prev = t ? lex[t - 1] : 0;
        const bool need = live && st != prev;
        __syncthreads();  // every lane has read lex before any writes it
        if (need) {
            uint32_t nc;
            const uint64_t ne = A.max_len > 32 ? fix_run<LongParser>(S, A, u, prev, st, cnt, lex[t], nc)
                                               : fix_run<WinParser, true>(S, A, u, prev, st, cnt, lex[t], nc);
            st = prev;
            cnt = nc;
            lex[t] = ne;
        }
        if (!__syncthreads_or(need)) break;
    }
    start[j] = st;
    exit_pos[j] = lex[t];
    count[j] = cnt;
}

// One fix pass.  Inside a workgroup the pass repeats until no subsequence's
// start differs from its predecessor's exit (the exits in LDS, a barrier per
// step), so a chain of changed exits inside the workgroup settles in one
// launch; only a workgroup's first subsequence waits for the next pass to
// see its predecessor's new exit.  changed: some exit differs from exit_in.
__global__ __launch_bounds__(256) void k_dec_fix(const uint32_t* __restrict__ words,
                                                 const DecArrayDev* __restrict__ arrays,
                                                 const uint32_t* __restrict__ wg_arr,
                                                 const uint32_t* __restrict__ tables, uint64_t* start,
                                                 const uint64_t* __restrict__ exit_in, uint64_t* exit_out,
                                                 uint32_t* count, unsigned int* changed) {
    __shared__ uint32_t lds[kRootMax];
    __shared__ uint32_t win[kWinLds];
    __shared__ uint64_t lex[256];
    const DecArrayDev A = arrays[wg_arr[blockIdx.x]];
    const uint32_t t = threadIdx.x;
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + t;
    const uint64_t u = j - A.sub0;
    const bool live = u < A.nsub && u != 0;
    const uint64_t ex0 = exit_in[j];
    const uint64_t prev0 = live ? exit_in[j - 1] : 0;
    uint64_t st = start[j];
    // the tables only when some lane re-parses (after convergence: none)
    if (!__syncthreads_or(live && st != prev0)) {
        exit_out[j] = ex0;
        return;
    }
    const ParseSrc S = load_src(lds, win, words, A, tables);
    lex[t] = ex0;
    uint32_t cnt = count[j];
    __syncthreads();
    for (uint32_t it = 0; it < blockDim.x; it++) {
        // the predecessor's exit: from the previous workgroup as this pass
        // found it, or from this workgroup's current exits
        const uint64_t prev = t == 0 ? prev0 : lex[t - 1];
        const bool need = live && st != prev;
        __syncthreads();  // every lane has read lex before any writes it
        if (need) {
            uint32_t nc;
            const uint64_t ne = A.max_len > 32 ? fix_run<LongParser>(S, A, u, prev, st, cnt, lex[t], nc)
                                               : fix_run<WinParser>(S, A, u, prev, st, cnt, lex[t], nc);
            st = prev;
            cnt = nc;
            lex[t] = ne;
        }
        if (!__syncthreads_or(need)) break;
    }
    start[j] = st;
    count[j] = cnt;
    exit_out[j] = lex[t];
    if (lex[t] != ex0) atomicOr(changed, 1u);
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
// codewords.  A line is decoded into registers and then stored with eight
// 16-byte stores back to back (a line written piecemeal over a long decode
// leaves half-written lines in L2, which reach HBM as partial writes).
// Errors: a run past its array's end (2), the parse of an array's last
// subsequence not ending at the payload's end (4).
template <class PR>
__device__ __forceinline__ void write_run(const ParseSrc& S, const DecArrayDev& A, uint64_t u, uint64_t st,
                                          uint64_t pos, uint32_t n, uint64_t a, uint64_t b, int32_t* out,
                                          unsigned int* err) {
    PR P;
    P.init(S, st, st);
    const uint64_t end = A.payload_bit + A.bits;
    if (b <= a) {  // no line starts among this thread's slots: nothing to store
        if (u + 1 == A.nsub) {
            for (uint32_t k = 0; k < n; k++) P.next();
            if (st + P.rel != end) atomicOr(err, 4u);
        }
        return;
    }
    // symbols as int32: the offset added modulo 2^32
    const uint32_t off = (uint32_t)A.offset;
    for (uint64_t k = pos; k < a; k++) P.next();
    uint64_t k = a;
    for (; k + kLineSyms <= b; k += kLineSyms) {
        int32_t v[kLineSyms];
#pragma unroll
        for (uint32_t i = 0; i < kLineSyms; i++) v[i] = (int32_t)(off + P.next());
        int4* d = reinterpret_cast<int4*>(out + k);
#pragma unroll
        for (uint32_t i = 0; i < kLineSyms / 4; i++) d[i] = make_int4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
    }
    for (; k < b; k++) out[k] = (int32_t)(off + P.next());  // the partial line at an array's end
    // the array's last run ends at the payload's end (b = oend = pos + n)
    if (u + 1 == A.nsub && st + P.rel != end) atomicOr(err, 4u);
}

__global__ __launch_bounds__(256) void k_dec_write(const uint32_t* __restrict__ words,
                                                   const DecArrayDev* __restrict__ arrays,
                                                   const uint32_t* __restrict__ wg_arr,
                                                   const uint32_t* __restrict__ tables,
                                                   const uint64_t* __restrict__ start,
                                                   const uint32_t* __restrict__ count,
                                                   const uint64_t* __restrict__ wg_off, int32_t* out,
                                                   unsigned int* err) {
    __shared__ uint32_t lds[kRootMax];
    __shared__ uint32_t win[kWinLds];
    __shared__ uint32_t wsum[4];
    const DecArrayDev A = arrays[wg_arr[blockIdx.x]];
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
    const ParseSrc S = load_src(lds, win, words, A, tables);  // its barrier also publishes wsum
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
    if (A.max_len > 32) write_run<LongParser>(S, A, u, start[j], pos, n, a, b, out, err);
    else write_run<WinParser>(S, A, u, start[j], pos, n, a, b, out, err);
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
