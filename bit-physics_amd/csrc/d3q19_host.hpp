// d3q19_host.hpp -- host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "d3q19_launch.hpp"

namespace d3q19 {

#define D3Q19_HIDDEN __attribute__((visibility("hidden")))

// Record the calling thread's error message (read by d3q19_last_error) and
// return `code`.
D3Q19_HIDDEN int set_error(int code, const std::string& msg);

// Copy n elements of dtype `dt` (D3Q19_F64 / D3Q19_F32) from host or device
// memory into the device buffer `dst` of double (dst_double) or float.
D3Q19_HIDDEN int stage_in(const void* src, int dt, int on_device, size_t n, void* dst,
                          bool dst_double, hipStream_t s);

// Copy n elements from the device buffer `src` (double or float) to `dst`,
// host or device memory of dtype `dt`.  Synchronises the stream.
D3Q19_HIDDEN int stage_out(const void* src, bool src_double, size_t n, void* dst, int dt,
                           int on_device, hipStream_t s);

inline uint32_t grid_for(uint64_t n, int block) { return (uint32_t)((n + block - 1) / block); }

// LDS bytes per CU of `device`, 0 if the runtime does not report it.
inline uint32_t device_cu_count(int device) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0) return 0;
    return (uint32_t)n;
}

inline uint32_t device_lds_per_cu(int device) {
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess ||
        lds <= 0)
        return 0;
    return (uint32_t)lds;
}

}  // namespace d3q19

#define D3Q19_HIPCHK(expr)                                                                    \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return ::d3q19::set_error(D3Q19_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
