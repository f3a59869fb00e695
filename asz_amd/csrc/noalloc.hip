// Allocation-free range scan (asz_range_f32_noalloc, no reference
// counterpart): the scan of asz_range_f32 (load_input's finite check and
// minmax_element, s1_compress.cpp:128-137) for callers that cannot allocate
// device memory.  Its only device memory beyond the caller's field is the
// scan's kRangeWords result words, which live in a file-scope device array of
// this library (32 bytes on every GPU the library runs on, held until the
// process ends) instead of the context's growable buffers.  The array's size
// does not depend on the field's length, so no length is too large for it.
// Every call on one GPU shares that GPU's array: calls that overlap on one GPU
// (several contexts or threads at once) race on it.
#include <mutex>

#include "context.hpp"
#include "launch.hpp"

namespace asz {
namespace dev {

// The scan's result words.  Host code never uses this name as an address:
// noalloc_setup resolves each device's copy with hipGetSymbolAddress.
static __device__ __align__(256) uint32_t scratch[kRangeWords];

// The preset of the result words that launch_range expects (min key ~0, max
// key 0, flag 0, first zero index ~0, last zero index 0), written by a
// kernel: every call starts from it, whatever the last call left.
__global__ __launch_bounds__(64) void k_range_preset(uint32_t* r) {
    if (threadIdx.x < kRangeWords)
        r[threadIdx.x] = (threadIdx.x == 0 || threadIdx.x == 4 || threadIdx.x == 5) ? 0xFFFFFFFFu : 0u;
}

}  // namespace dev

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices];  // per device: the device address of dev::scratch
}  // namespace

void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) fail(ASZ_ERR_ARG, "device index beyond the library's limit of 64");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(dev::scratch)));
    g_scratch[device] = p;
}

}  // namespace asz

extern "C" int asz_range_f32_noalloc(asz_ctx* c, const float* d_field, uint64_t n, float* min_val,
                                     float* max_val) {
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        if (n == 0) asz::fail(ASZ_ERR_ARG, "empty field");
        void* p = nullptr;
        {
            std::lock_guard<std::mutex> lk(asz::g_mu);
            p = asz::g_scratch[c->device];
        }
        if (!p) asz::fail(ASZ_ERR_HIP, "the range scan's scratch array is not set up on this device");
        uint32_t* r = static_cast<uint32_t*>(p);
        hipLaunchKernelGGL(asz::dev::k_range_preset, dim3(1), dim3(64), 0, c->stream, r);
        asz::launch_range(d_field, n, r, c->stream);
        uint32_t w[asz::kRangeWords];
        HIPCHECK(hipMemcpyAsync(w, r, sizeof(w), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        asz::range_result(c, w, d_field, min_val, max_val, std::string());
    });
}
