// Host Zstd wrapping of the .asz frame (compress.tpp:294-299, decompress.tpp:83-86).
// Links the image's libzstd.so.1; only its stable public C entry points are used.
//
// Threaded form (no reference counterpart).  The pre-Zstd stream is Huffman
// output, which Zstd stores as raw blocks, so a single-threaded level-3 call
// spends its time scanning for matches it does not find.  With T threads the
// stream is cut into T slices at multiples of the 128 KiB block size and the
// slices are compressed concurrently, each by its own context:
//   * slice 0 by a streaming context that was told the whole stream's size,
//     so its frame header and its blocks are the ones ZSTD_compress writes
//     for the start of the stream;
//   * every later slice as a frame of its own, whose blocks are then spliced
//     in after those of slice 0 (frame header dropped, last-block bits
//     cleared except on the final block).
// A raw or RLE block carries no state from earlier blocks, so the spliced
// frame is valid whenever the blocks of slices 1 .. T-1 are all raw or RLE.
// If a later slice holds a compressed block (its sequences could use repeat
// offsets that depend on the blocks before it), the stream is compressed
// again by one single-threaded call.  When the single-threaded call would
// also store those slices raw (no match reaching back across a slice
// boundary), the spliced frame is byte for byte the single-threaded frame.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <thread>
#include <vector>

#include "internal.hpp"

extern "C" {
typedef struct ZSTD_CCtx_s ZSTD_CCtx;
typedef struct {
    const void* src;
    size_t size, pos;
} ZSTD_inBuffer;
typedef struct {
    void* dst;
    size_t size, pos;
} ZSTD_outBuffer;
size_t ZSTD_compressBound(size_t srcSize);
size_t ZSTD_compress(void* dst, size_t dstCapacity, const void* src, size_t srcSize, int compressionLevel);
size_t ZSTD_decompress(void* dst, size_t dstCapacity, const void* src, size_t compressedSize);
unsigned long long ZSTD_getFrameContentSize(const void* src, size_t srcSize);
unsigned ZSTD_isError(size_t code);
ZSTD_CCtx* ZSTD_createCCtx(void);
size_t ZSTD_freeCCtx(ZSTD_CCtx* cctx);
size_t ZSTD_CCtx_setParameter(ZSTD_CCtx* cctx, int param, int value);
size_t ZSTD_CCtx_setPledgedSrcSize(ZSTD_CCtx* cctx, unsigned long long pledgedSrcSize);
size_t ZSTD_compressStream2(ZSTD_CCtx* cctx, ZSTD_outBuffer* output, ZSTD_inBuffer* input, int endOp);
}

namespace {

constexpr int kZstdLevelParam = 100;     // ZSTD_c_compressionLevel
constexpr int kZstdFlush = 1;            // ZSTD_e_flush
constexpr uint64_t kBlock = 1ull << 17;  // ZSTD_BLOCKSIZE_MAX
constexpr uint64_t kMinSlice = 8 * kBlock;
constexpr uint32_t kMaxThreads = 64;

uint32_t rd24(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16); }

// Bytes of a frame header (RFC 8878 section 3.1.1.1) at p, or 0 if malformed.
size_t frame_header_bytes(const uint8_t* p, size_t n) {
    if (n < 6 || p[0] != 0x28 || p[1] != 0xB5 || p[2] != 0x2F || p[3] != 0xFD) return 0;
    const uint8_t fhd = p[4];
    const bool single = (fhd >> 5) & 1;
    static const size_t did[4] = {0, 1, 2, 4};
    static const size_t fcs[4] = {0, 2, 4, 8};
    size_t fcs_bytes = fcs[fhd >> 6];
    if (single && (fhd >> 6) == 0) fcs_bytes = 1;
    const size_t h = 5 + (single ? 0 : 1) + did[fhd & 3] + fcs_bytes;
    return h <= n ? h : 0;
}

// Walks the blocks of a frame body [p, p + n): returns false if any block is
// compressed or the walk does not end exactly at the last block (a content
// checksum is never requested here).
bool body_is_stateless(const uint8_t* p, size_t n) {
    size_t o = 0;
    while (o + 3 <= n) {
        const uint32_t bh = rd24(p + o);
        const uint32_t type = (bh >> 1) & 3, size = bh >> 3;
        if (type == 2 || type == 3) return false;
        o += 3 + (type == 1 ? 1 : size);
        if (bh & 1) return o == n;
    }
    return false;
}

// Clears the last-block bit of the final block of the body [p, p + n).
void clear_last_bit(uint8_t* p, size_t n) {
    size_t o = 0;
    while (o + 3 <= n) {
        const uint32_t bh = rd24(p + o);
        const uint32_t type = (bh >> 1) & 3, size = bh >> 3;
        if (bh & 1) {
            p[o] &= 0xFE;
            return;
        }
        o += 3 + (type == 1 ? 1 : size);
    }
}

int compress_single(const uint8_t* src, uint64_t n, int level, uint8_t** dst, uint64_t* dst_bytes) {
    size_t cap = ZSTD_compressBound(n);
    uint8_t* out = static_cast<uint8_t*>(std::malloc(cap ? cap : 1));
    if (!out) return ASZ_ERR_NOMEM;
    size_t r = ZSTD_compress(out, cap, src, n, level);
    if (ZSTD_isError(r)) {
        std::free(out);
        return ASZ_ERR_DATA;
    }
    *dst = out;
    *dst_bytes = r;
    return ASZ_OK;
}

// Runs f(0) .. f(k-1) on k threads (the caller runs f(0)).
template <typename F>
void parallel(uint32_t k, F&& f) {
    std::vector<std::thread> th;
    th.reserve(k);
    for (uint32_t i = 1; i < k; i++) th.emplace_back([&f, i] { f(i); });
    f(0);
    for (auto& t : th) t.join();
}

}  // namespace

extern "C" int asz_zstd_compress_mt(const uint8_t* src, uint64_t n, int level, uint32_t threads, uint8_t** dst,
                                    uint64_t* dst_bytes, uint32_t* threads_used) {
    if (!src || !dst || !dst_bytes) return ASZ_ERR_ARG;
    if (threads == 0) threads = std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
    threads = std::min(threads, kMaxThreads);
    // slices of whole blocks, none shorter than kMinSlice
    uint64_t slice = (n + threads - 1) / threads;
    slice = std::max<uint64_t>(kMinSlice, (slice + kBlock - 1) / kBlock * kBlock);
    const uint32_t k = (uint32_t)std::min<uint64_t>(threads, (n + slice - 1) / slice);
    if (threads_used) *threads_used = 1;
    if (k <= 1) return compress_single(src, n, level, dst, dst_bytes);

    struct Part {
        std::unique_ptr<uint8_t[]> out;  // not value-initialised: the slice is compressed into it
        size_t cap = 0, body = 0, len = 0;  // capacity, first body byte and body bytes in out
        bool ok = false;
    };
    std::vector<Part> parts(k);
    parallel(k, [&](uint32_t i) {
        Part& P = parts[i];
        const uint64_t lo = i * slice, len = std::min<uint64_t>(slice, n - lo);
        ZSTD_CCtx* cc = ZSTD_createCCtx();
        if (!cc) return;
        P.cap = ZSTD_compressBound(len) + 64;
        P.out.reset(new (std::nothrow) uint8_t[P.cap]);
        if (!P.out) {
            ZSTD_freeCCtx(cc);
            return;
        }
        if (i == 0) {
            // the whole stream's frame header, then the blocks of slice 0
            ZSTD_inBuffer in{src, len, 0};
            ZSTD_outBuffer ob{P.out.get(), P.cap, 0};
            bool good = !ZSTD_isError(ZSTD_CCtx_setParameter(cc, kZstdLevelParam, level)) &&
                        !ZSTD_isError(ZSTD_CCtx_setPledgedSrcSize(cc, n));
            size_t rem = 1;
            while (good && rem != 0) {
                rem = ZSTD_compressStream2(cc, &ob, &in, kZstdFlush);
                good = !ZSTD_isError(rem) && (rem == 0 || ob.pos < ob.size);
            }
            if (good && in.pos == len) {
                P.body = 0;
                P.len = ob.pos;
                P.ok = true;
            }
        } else {
            const size_t r = ZSTD_compress(P.out.get(), P.cap, src + lo, len, level);
            if (!ZSTD_isError(r)) {
                const size_t h = frame_header_bytes(P.out.get(), r);
                if (h && body_is_stateless(P.out.get() + h, r - h)) {
                    P.body = h;
                    P.len = r - h;
                    if (i + 1 < k) clear_last_bit(P.out.get() + h, P.len);
                    P.ok = true;
                }
            }
        }
        ZSTD_freeCCtx(cc);
    });
    bool all = true;
    for (const Part& P : parts) all = all && P.ok;
    if (!all) return compress_single(src, n, level, dst, dst_bytes);
    std::vector<uint64_t> off(k + 1, 0);
    for (uint32_t i = 0; i < k; i++) off[i + 1] = off[i] + parts[i].len;
    uint8_t* out = static_cast<uint8_t*>(std::malloc(off[k] ? off[k] : 1));
    if (!out) return ASZ_ERR_NOMEM;
    parallel(k, [&](uint32_t i) { std::memcpy(out + off[i], parts[i].out.get() + parts[i].body, parts[i].len); });
    *dst = out;
    *dst_bytes = off[k];
    if (threads_used) *threads_used = k;
    return ASZ_OK;
}

extern "C" int asz_zstd_compress(const uint8_t* src, uint64_t n, uint8_t** dst, uint64_t* dst_bytes) {
    return asz_zstd_compress_mt(src, n, 3, 1, dst, dst_bytes, nullptr);
}

extern "C" int asz_zstd_decompress(const uint8_t* src, uint64_t n, uint8_t** dst, uint64_t* dst_bytes) {
    unsigned long long size = ZSTD_getFrameContentSize(src, n);
    if (size >= (1ull << 63)) return ASZ_ERR_DATA;
    uint8_t* out = static_cast<uint8_t*>(std::malloc(size ? size : 1));
    if (!out) return ASZ_ERR_NOMEM;
    size_t r = ZSTD_decompress(out, size, src, n);
    if (ZSTD_isError(r) || r != size) {
        std::free(out);
        return ASZ_ERR_DATA;
    }
    *dst = out;
    *dst_bytes = size;
    return ASZ_OK;
}

extern "C" void asz_free(void* p) { std::free(p); }
