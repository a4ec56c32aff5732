// d3q19_launch.hpp -- launch-shape arithmetic of the host code.  No HIP
// dependency, so host-only tests compile it with a plain C++ compiler.
#pragma once

#include <stdint.h>

namespace d3q19 {

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
