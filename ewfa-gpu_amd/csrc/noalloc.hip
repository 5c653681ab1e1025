// noalloc.hip -- the allocation-free variant of ewfa_align_device.
//
// Its scratch is one file-scope device array per GPU, sized
// EWFA_NOALLOC_SCRATCH_BYTES.  Host code never uses the array's name as an
// address (on the host it names a stand-in, not device memory): the device
// address is looked up once per device with hipGetSymbolAddress, with that
// device current, and kept per device.  Every call on a GPU shares that GPU's
// array, so calls that overlap on one GPU race; the plan clears the control
// words it needs at zero with fill kernels at its start, and every other byte
// a call reads it has written itself.
#include <mutex>

#include "ewfa_internal.h"
#include "runtime.h"

namespace {

__device__ __align__(256) uint8_t scratch[EWFA_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};

// The array's device address on the current device `device`, looked up on
// the first call for that device.
int scratch_ptr(int device, void** out) {
    if (device < 0 || device >= MAX_DEVICES)
        return ewfa::set_error(EWFA_ERR_ARG, "device %d out of range for the allocation-free scratch", device);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return ewfa::set_error(EWFA_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                                   hipGetErrorString(e));
        g_ptr[device] = p;
    }
    *out = g_ptr[device];
    return EWFA_OK;
}

}  // namespace

extern "C" int ewfa_align_device_noalloc(ewfa_ctx_t* ctx, const uint64_t* d_words, const ewfa_pair_t* d_pairs,
                                         size_t n, uint32_t max_len, ewfa_result_t* d_results, int8_t* d_cigars,
                                         void* stream) {
    if (!ctx) return ewfa::set_error(EWFA_ERR_ARG, "ctx is NULL");
    const int device = ewfa::context_device(ctx);
    if (hipSetDevice(device) != hipSuccess) return ewfa::set_error(EWFA_ERR_HIP, "hipSetDevice(%d) failed", device);
    void* base = nullptr;
    const int rc = scratch_ptr(device, &base);
    if (rc != EWFA_OK) return rc;
    return ewfa::align_device_in(ctx, d_words, d_pairs, n, max_len, d_results, d_cigars, stream, base,
                                 EWFA_NOALLOC_SCRATCH_BYTES);
}
