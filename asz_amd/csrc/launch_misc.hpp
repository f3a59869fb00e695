#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace asz {

// Copy job: len bytes from src[src_off] to the stream at byte dst.
struct ByteSeg {
    uint64_t dst, src_off, len;
};
void launch_scatter_bytes(const ByteSeg* segs, uint32_t nsegs, const uint8_t* src, uint8_t* dst, hipStream_t st);

// nfields > 1: field f's values v_stride and its result out_stride int32 further.
void launch_minmax_i32(const int32_t* v, uint64_t n, int32_t* out2, hipStream_t st, uint32_t nfields = 1,
                       uint64_t v_stride = 0, uint64_t out_stride = 0);
void launch_fill_i32(int32_t* v, uint64_t n, int32_t val, hipStream_t st);
void launch_t_float(const int32_t* t_int, uint64_t nblocks, uint32_t L, float prec, float* tblk, hipStream_t st);

// The byte ranges of a stream in device memory that its host parser reads
// (the preamble, every array's count and Huffman header with its payload
// length), copied back to back into out (cap bytes).  meta (3 + 2
// kStreamRegions words): status (0 ok, 1 truncated or malformed, 2 out too
// small), the bytes the ranges take, the range count, then (start, length)
// per range.  One workgroup; a malformed stream stops the walk, never a read
// past n.
constexpr uint32_t kStreamRegions = 2 * (63 + 2) + 1;
void launch_stream_regions(const uint8_t* s, uint64_t n, uint8_t* out, uint64_t cap, uint64_t* meta, hipStream_t st);

}  // namespace asz
