// fk_noalloc.hip -- device scratch of the entry point that allocates nothing
// (fkgpu_eig_logz_device_noalloc).
//
// The HBM-workspace eigensolvers (N > 256) keep one N x N working matrix per
// workgroup.  The ordinary entry points grow a hipMalloc'd buffer for it; callers
// that cannot allocate (code under HIP graph capture, serving loops, harnesses
// that forbid hipMalloc) use this file-scope array instead.  Every GPU the library
// runs on holds its own copy, FK_NOALLOC_SCRATCH_BYTES in size, from the moment the
// code object is loaded until the process ends.  Host code must never pass the
// array's name to a kernel (on the host it is only a stand-in symbol): its device
// address is resolved once per device with hipGetSymbolAddress, with that device
// current, and only that address is used.
#include <hip/hip_runtime.h>

#include <mutex>

#include "fk_internal.h"

static __device__ __align__(256) double scratch[FK_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
bool g_resolved[kMaxDevices];
hipError_t g_status[kMaxDevices];
double* g_ptr[kMaxDevices];

}  // namespace

// Device address of `device`'s copy of the scratch array.  Resolved on the first
// call for that device (with the device made current), cached afterwards; the
// status of that resolution is returned on every call.
hipError_t fk_noalloc_scratch(int device, double** ptr, size_t* doubles) {
    *ptr = nullptr;
    *doubles = 0;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_resolved[device]) {
        int prev = 0;
        hipError_t e = hipGetDevice(&prev);
        if (e == hipSuccess) e = hipSetDevice(device);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != device) (void)hipSetDevice(prev);
        g_status[device] = e;
        g_ptr[device] = e == hipSuccess ? static_cast<double*>(p) : nullptr;
        g_resolved[device] = true;
    }
    if (g_status[device] != hipSuccess) return g_status[device];
    *ptr = g_ptr[device];
    *doubles = FK_NOALLOC_SCRATCH_BYTES / sizeof(double);
    return hipSuccess;
}
