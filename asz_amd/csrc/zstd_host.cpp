// Host Zstd wrapping of the .asz frame (compress.tpp:294-299, decompress.tpp:83-86).
// Links the image's libzstd.so.1; only its stable public C entry points are used.
#include <cstdlib>
#include <cstring>

#include "internal.hpp"

extern "C" {
size_t ZSTD_compressBound(size_t srcSize);
size_t ZSTD_compress(void* dst, size_t dstCapacity, const void* src, size_t srcSize, int compressionLevel);
size_t ZSTD_decompress(void* dst, size_t dstCapacity, const void* src, size_t compressedSize);
unsigned long long ZSTD_getFrameContentSize(const void* src, size_t srcSize);
unsigned ZSTD_isError(size_t code);
}

extern "C" int asz_zstd_compress(const uint8_t* src, uint64_t n, uint8_t** dst, uint64_t* dst_bytes) {
    size_t cap = ZSTD_compressBound(n);
    uint8_t* out = static_cast<uint8_t*>(std::malloc(cap ? cap : 1));
    if (!out) return ASZ_ERR_NOMEM;
    size_t r = ZSTD_compress(out, cap, src, n, 3);
    if (ZSTD_isError(r)) {
        std::free(out);
        return ASZ_ERR_DATA;
    }
    *dst = out;
    *dst_bytes = r;
    return ASZ_OK;
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
