// d3q19_calibration.hip -- measurement helpers.
//  * calibration copy kernel: a plain 16-byte-per-lane device copy of known
//    size, used to measure the achievable HBM bandwidth and to scale the
//    FETCH_SIZE / WRITE_SIZE counters against a known byte count;
//  * shader-clock probe: the device's current shader clock, read between
//    launches to see power-management transients in a timed window.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/d3q19.h"
#include "d3q19_host.hpp"

namespace {

typedef unsigned int v4u __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_copy16(const v4u* __restrict__ a, v4u* __restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256;
    for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(&a[i]), &b[i]);
}

// One wavefront counts shader-clock ticks (s_memtime) against the constant
// real-time counter (s_memrealtime) over about 2000 real-time ticks (20 us
// at 100 MHz); the loop is bounded in iterations as well.
__global__ void __launch_bounds__(64) k_clock(unsigned long long* __restrict__ out) {
    const unsigned long long c0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime();
    unsigned long long c1 = c0, r1 = r0;
    for (int it = 0; it < 200000 && r1 - r0 < 2000ull; ++it) {
        c1 = __builtin_amdgcn_s_memtime();
        r1 = __builtin_amdgcn_s_memrealtime();
    }
    if (threadIdx.x == 0) {
        out[0] = c1 - c0;
        out[1] = r1 - r0;
    }
}

}  // namespace

extern "C" int d3q19_shader_clock(void* ticks, double* mhz, void* stream) {
    if (!ticks || !mhz) return d3q19::set_error(D3Q19_EINVAL, "null argument");
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, khz = 0;
    D3Q19_HIPCHK(hipGetDevice(&dev));
    D3Q19_HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    if (khz <= 0) return d3q19::set_error(D3Q19_EHIP, "device reports no wall-clock rate");
    k_clock<<<1, 64, 0, s>>>((unsigned long long*)ticks);
    D3Q19_HIPCHK(hipGetLastError());
    unsigned long long t[2] = {0, 0};
    D3Q19_HIPCHK(hipMemcpyAsync(t, ticks, sizeof(t), hipMemcpyDeviceToHost, s));
    D3Q19_HIPCHK(hipStreamSynchronize(s));
    if (t[1] == 0) return d3q19::set_error(D3Q19_EHIP, "real-time counter did not advance");
    *mhz = (double)t[0] / (double)t[1] * (khz / 1000.0);
    return D3Q19_OK;
}

extern "C" int d3q19_copy_kernel(const void* src, void* dst, size_t bytes, void* stream) {
    if (!src || !dst) return d3q19::set_error(D3Q19_EINVAL, "null array");
    if (bytes % 16) return d3q19::set_error(D3Q19_EINVAL, "bytes must be a multiple of 16");
    const size_t n = bytes / 16;
    const size_t blocks = (n + 255) / 256;
    const unsigned grid = (unsigned)(blocks < 65536 * 16 ? blocks : 65536 * 16);
    k_copy16<<<grid, 256, 0, (hipStream_t)stream>>>((const v4u*)src, (v4u*)dst, n);
    D3Q19_HIPCHK(hipGetLastError());
    return D3Q19_OK;
}
