// d3q19_host.hpp -- host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

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
inline uint32_t device_lds_per_cu(int device) {
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess ||
        lds <= 0)
        return 0;
    return (uint32_t)lds;
}

// Dynamic LDS that a block reserves, and does not use, so that at most k
// blocks fit on a CU with lds_per_cu bytes: lds_per_cu / (k + 1) plus 1 KiB
// leaves room for k blocks and not for k + 1.  0 (no reservation) for k = 0,
// an unknown LDS size, or a reservation beyond 64 KiB per block.
inline uint32_t reserve_lds(uint32_t lds_per_cu, uint32_t k) {
    if (k == 0 || lds_per_cu == 0) return 0;
    const uint32_t b = lds_per_cu / (k + 1) + 1024;
    return (b <= lds_per_cu / k && b <= 65536) ? b : 0;
}

}  // namespace d3q19

#define D3Q19_HIPCHK(expr)                                                                    \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return ::d3q19::set_error(D3Q19_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
