// Small helper kernels: int32 min/max, fill, t_int -> per-round float t.
#include "device.hpp"
#include "launch_misc.hpp"

namespace asz {
namespace dev {

__global__ __launch_bounds__(256) void k_minmax_i32(const int32_t* __restrict__ v, uint64_t n, int32_t* out2,
                                                    uint64_t v_stride, uint64_t out_stride) {
    v += blockIdx.y * v_stride;  // field blockIdx.y
    out2 += blockIdx.y * out_stride;
    int32_t mn = INT32_MAX, mx = INT32_MIN;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        mn = min(mn, v[i]);
        mx = max(mx, v[i]);
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&out2[0], mn);
        atomicMax(&out2[1], mx);
    }
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t* __restrict__ v, uint64_t n, int32_t val) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        v[i] = val;
}

// tblk[r * nblocks + b] = (float) t_int[b * L + r] / (float) t_precision  (decompress.tpp:185)
__global__ __launch_bounds__(256) void k_t_float(const int32_t* __restrict__ t_int, uint64_t nblocks, uint32_t L,
                                                 float prec, float* tblk) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nblocks * L) return;
    uint64_t b = i / L, r = i % L;
    tblk[r * nblocks + b] = (float)t_int[i] / prec;
}

// One workgroup per segment: copy header bytes into the (zeroed) stream buffer.
__global__ __launch_bounds__(256) void k_scatter_bytes(const ByteSeg* __restrict__ segs, const uint8_t* __restrict__ src,
                                                       uint8_t* __restrict__ dst) {
    const ByteSeg S = segs[blockIdx.x];
    for (uint64_t i = threadIdx.x; i < S.len; i += blockDim.x) dst[S.dst + i] = src[S.src_off + i];
}

// Walker of launch_stream_regions.  The layout (the reference's, huffman.hpp
// and compress.tpp:243-299): 4-byte min and max, a plan byte (0x40 | L, 0x80
// for quintic; a plain L in older streams) and 2 L predictor bytes, then per
// array an 8-byte little-endian count and, unless the count is 0 (the t
// array always has a header), a Huffman header: a flag byte (low 6 bits: the
// symbol width), the 8-byte offset, the big-endian leaf count and maximum,
// then for a one-symbol array 8 bytes, otherwise the tree (2 nleaf - 1 +
// nleaf width bits, whole bytes) and the big-endian payload bit length (xor
// 0x1234abcd), then the payload.
__global__ __launch_bounds__(256) void k_stream_regions(const uint8_t* __restrict__ s, uint64_t n, uint8_t* out,
                                                        uint64_t cap, uint64_t* meta) {
    __shared__ uint64_t rs[kStreamRegions], rl[kStreamRegions], ro[kStreamRegions];
    __shared__ uint32_t s_nreg, s_status;
    __shared__ uint64_t s_need;
    if (threadIdx.x == 0) {
        uint32_t nreg = 0, status = 0;
        auto add = [&](uint64_t a, uint64_t l) {
            rs[nreg] = a;
            rl[nreg] = l;
            nreg++;
        };
        // 8 bytes at off (all loads issued before use)
        auto get8 = [&](uint64_t off, bool big) {
            uint8_t b[8];
#pragma unroll
            for (int i = 0; i < 8; i++) b[i] = s[off + i];
            uint64_t v = 0;
#pragma unroll
            for (int i = 0; i < 8; i++) v |= (uint64_t)b[i] << (8 * (big ? 7 - i : i));
            return v;
        };
        uint64_t L = 0;
        if (n < 9) status = 1;
        else {
            const uint8_t f = s[8];
            L = (f & 0x40) ? (f & 0x3Fu) : f;
            if (L == 0 || L > 63 || 9 + 2 * L + 8 > n) status = 1;
        }
        if (!status) {
            add(0, 9 + 2 * L);
            uint64_t off = 9 + 2 * L;
            for (uint64_t k = 0; k < L + 2; k++) {
                if (off + 8 > n) { status = 1; break; }
                const uint64_t cnt = get8(off, false);
                add(off, 8);
                off += 8;
                if (cnt == 0 && k > 0) continue;
                if (off + 25 > n) { status = 1; break; }
                const uint32_t width = s[off] & 0x3Fu;
                const uint64_t nleaf = get8(off + 9, true), maxval = get8(off + 17, true);
                if (nleaf == 0) { status = 1; break; }
                if (maxval == 1 || nleaf == 1) {
                    if (off + 33 > n) { status = 1; break; }
                    add(off, 33);
                    off += 33;
                    continue;
                }
                if (width > 31 || nleaf > n * 8) { status = 1; break; }
                const uint64_t tree = (2 * nleaf - 1 + nleaf * width + 7) / 8;
                if (tree > n || off + 33 + tree > n) { status = 1; break; }
                const uint64_t bits = get8(off + 25 + tree, true) ^ 0x1234abcdull;
                add(off, 33 + tree);
                off += 33 + tree;
                const uint64_t pay = (bits + 7) / 8;
                if (bits > n * 8 || pay > n - off) { status = 1; break; }
                off += pay;
            }
        }
        uint64_t need = 0;
        for (uint32_t r = 0; r < nreg; r++) {
            ro[r] = need;
            need += rl[r];
        }
        if (!status && need > cap) status = 2;
        s_nreg = nreg;
        s_status = status;
        s_need = need;
        meta[0] = status;
        meta[1] = need;
        meta[2] = nreg;
        for (uint32_t r = 0; r < nreg; r++) {
            meta[3 + 2 * r] = rs[r];
            meta[4 + 2 * r] = rl[r];
        }
    }
    __syncthreads();
    if (s_status) return;
    for (uint32_t r = 0; r < s_nreg; r++)
        for (uint64_t i = threadIdx.x; i < rl[r]; i += blockDim.x) out[ro[r] + i] = s[rs[r] + i];
}

}  // namespace dev

void launch_scatter_bytes(const ByteSeg* segs, uint32_t nsegs, const uint8_t* src, uint8_t* dst, hipStream_t st) {
    if (!nsegs) return;
    hipLaunchKernelGGL(dev::k_scatter_bytes, dim3(nsegs), dim3(256), 0, st, segs, src, dst);
}

void launch_minmax_i32(const int32_t* v, uint64_t n, int32_t* out2, hipStream_t st, uint32_t nfields,
                       uint64_t v_stride, uint64_t out_stride) {
    if (!n) return;
    uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(dev::k_minmax_i32, dim3((uint32_t)blocks, nfields), dim3(256), 0, st, v, n, out2, v_stride,
                       out_stride);
}
void launch_fill_i32(int32_t* v, uint64_t n, int32_t val, hipStream_t st) {
    if (!n) return;
    uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(dev::k_fill_i32, dim3((uint32_t)blocks), dim3(256), 0, st, v, n, val);
}
void launch_t_float(const int32_t* t_int, uint64_t nblocks, uint32_t L, float prec, float* tblk, hipStream_t st) {
    uint64_t n = nblocks * L;
    if (!n) return;
    hipLaunchKernelGGL(dev::k_t_float, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, t_int, nblocks, L, prec,
                       tblk);
}

void launch_stream_regions(const uint8_t* s, uint64_t n, uint8_t* out, uint64_t cap, uint64_t* meta, hipStream_t st) {
    hipLaunchKernelGGL(dev::k_stream_regions, dim3(1), dim3(256), 0, st, s, n, out, cap, meta);
}

}  // namespace asz
