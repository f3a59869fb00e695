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

}  // namespace asz
