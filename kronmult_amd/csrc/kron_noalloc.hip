// kron_noalloc.hip -- kron_xbatched_noalloc_f64: the xbatched update with no
// allocation anywhere in the call.
//
// kron_xbatched_f64 keeps the scratch of its Y-target sort (calls with
// batchCount >= 2048) in a per-stream workspace that the library allocates
// on first use and grows on demand.  Callers that cannot allocate (stream
// capture, serving loops, harnesses that forbid hipMalloc) get the same
// update with that scratch in a fixed device array instead.  Each GPU holds
// its own copy of the array (a module global) until the process ends; its
// address on the current device is looked up once per device with
// hipGetSymbolAddress and kept.  Every call clears the part of the array
// that must start at zero with a fill kernel, so whatever an earlier call
// left there does not matter.  Calls that overlap on one GPU race for the
// array; calls ordered on one stream do not.
#include <mutex>

#include "kron_internal.hpp"

namespace kron {

constexpr size_t kNoallocScratchBytes = group_workspace_bytes(KRON_NOALLOC_MAX_BATCH);
static_assert(kNoallocScratchBytes <= ((size_t)32 << 20), "no-allocation scratch stays modest");

// The sort scratch of kron_xbatched_noalloc_f64 (16.75 MiB per GPU).  In host
// code this name is only a handle for hipGetSymbolAddress, never an address.
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;

// Device address of `scratch` on the current device (nullptr on failure).
static void *scratch_on_current_device()
{
    static std::mutex mu;
    static void *addr[kMaxDevices] = {nullptr};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_last_hip_error((int)e);
        return nullptr;
    }
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_last_hip_error((int)e);
            return nullptr;
        }
        addr[dev] = ptr;
    }
    return addr[dev];
}

}  // namespace kron

using namespace kron;

extern "C" kron_status kron_xbatched_noalloc_f64(int d, int n, const double *const *Aarray,
                                                 int lda, const double *const *pX,
                                                 double *const *pY, double *const *pW,
                                                 int batchCount, int subbatchCount,
                                                 kron_stream_t stream)
{
    if (batchCount > KRON_NOALLOC_MAX_BATCH) return KRON_ETOOBIG;
    KronArgs<double> a;
    kron_status st = xbatched_args(a, d, n, Aarray, lda, pX, pY, pW, batchCount, subbatchCount,
                                   stream);
    if (st != KRON_OK || batchCount == 0) return st;
    void *ws = scratch_on_current_device();
    if (!ws) return KRON_EHIP;
    a.sort_ws = ws;
    a.sort_ws_bytes = kNoallocScratchBytes;
    return launch_kron(a);
}
