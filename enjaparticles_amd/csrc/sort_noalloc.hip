// rtps_sort_pairs_noalloc: the key/value sort of rtps_sort_pairs for callers
// that cannot allocate device memory (code under stream capture, serving
// loops, harnesses that forbid hipMalloc).  The sort's scratch (digit totals,
// tile counters, look-back status words, alternate key/value arrays) lives in
// one file-scope device array of fixed size instead of a caller workspace.
//
// Each GPU holds its own copy of the array, at its own address: the first
// call on a device (with that device current) looks the address up with
// hipGetSymbolAddress and keeps it per device; host code uses only that
// pointer, never the array's name.  The array keeps the last call's contents;
// the sort clears the parts that must start at zero (its digit totals, tile
// counters and look-back status words) with kernels at every call.  Calls
// that overlap on one GPU (two streams, two threads) share the array and race.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "rtps_internal.h"

namespace rtps {

// Largest sort: 2^20 keys (the alternate key/value arrays, 8 MiB, plus the
// digit totals and look-back status words: sort_temp_bytes(2^20, 32) < 11 MiB).
constexpr int kNoAllocMaxKeys = 1 << 20;
constexpr size_t kNoAllocScratchBytes = (size_t)11 << 20;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
int scratch_of_current_device(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_error(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return set_error(RTPS_ERR_ARG, "rtps_sort_pairs_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(scratch)");
        if (!p) return set_error(RTPS_ERR_HIP, "hipGetSymbolAddress(scratch) gave a null address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return RTPS_OK;
}

}  // namespace
}  // namespace rtps

using namespace rtps;

extern "C" int rtps_sort_pairs_noalloc(int n, int key_bits, uint32_t* keys, uint32_t* vals, void* stream) {
    if (n < 0 || key_bits < 0 || key_bits > 32)
        return set_error(RTPS_ERR_ARG, "rtps_sort_pairs_noalloc: bad n or key_bits");
    if (n > kNoAllocMaxKeys)
        return set_error(RTPS_ERR_ARG, "rtps_sort_pairs_noalloc: n = " + std::to_string(n) + " exceeds the maximum of " +
                                           std::to_string(kNoAllocMaxKeys) + " keys");
    if (n <= 1 || key_bits == 0) return RTPS_OK;
    if (!keys || !vals) return set_error(RTPS_ERR_ARG, "rtps_sort_pairs_noalloc: null buffer");
    if (sort_temp_bytes(n, key_bits) > kNoAllocScratchBytes)
        return set_error(RTPS_ERR_ARG, "rtps_sort_pairs_noalloc: scratch too small for this sort");
    void* tmp = nullptr;
    int rc = scratch_of_current_device(&tmp);
    if (rc) return rc;
    hipError_t e = sort_pairs_inplace(n, key_bits, keys, vals, tmp, (hipStream_t)stream);
    if (e != hipSuccess) return hip_error(e, "rtps_sort_pairs_noalloc");
    return RTPS_OK;
}
