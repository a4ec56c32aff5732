// d3q19_calibration.hip -- calibration copy kernel: a plain 16-byte-per-lane device
// copy of known size, used to measure the achievable HBM bandwidth and to
// scale the FETCH_SIZE / WRITE_SIZE counters against a known byte count.
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

}  // namespace

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
