// d3q19_launch.hpp -- launch-shape and memory-plan arithmetic of the host code.  No HIP
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

// Launches one fused step is split into along x (plane chunks, each a run of
// whole x planes).  Within one long launch of the step the wavefronts slow
// down as it proceeds (256^3 FP32, per-wavefront timestamps: 3.77 us in its
// first tenth, 3.97 us in its last); a new launch starts fast again.  The
// step is therefore launched in chunks of about target_us of work, from the
// rate est_glups the storage steps at, once the whole step would take
// min_us or more; at most one chunk per plane.  override_n > 0 forces that
// many chunks (clamped to nx).
inline uint32_t step_chunks(uint64_t ncell, uint32_t nx, double est_glups, double min_us, double target_us,
                            uint32_t override_n = 0) {
    if (nx == 0) return 1;
    if (override_n > 0) return override_n < nx ? override_n : nx;
    if (est_glups <= 0.0 || target_us <= 0.0) return 1;
    const double us = (double)ncell / (est_glups * 1e3);
    if (us < min_us) return 1;
    uint32_t n = (uint32_t)(us / target_us + 0.5);
    if (n < 1) n = 1;
    return n < nx ? n : nx;
}

// Planes per chunk for n chunks of nx planes (the last chunk takes the rest).
inline uint32_t chunk_planes(uint32_t nx, uint32_t n) { return n <= 1 ? nx : (nx + n - 1) / n; }

// Gap, in x planes, between the state a moment-coded ring step reads and the
// one it writes in place (me_row_in / me_row_out): the wavefronts of plane x
// wait for the pulls of planes x - gap - 1 .. x - gap + 1, which should be
// long done -- twice the planes that the blocks resident at once cover, plus
// two.  At most nx (then every plane is written into a free slot).
inline uint32_t ring_gap(uint32_t nx, uint64_t blocks_per_plane, uint64_t resident_blocks) {
    if (blocks_per_plane == 0) return nx;
    const uint64_t planes = (resident_blocks + blocks_per_plane - 1) / blocks_per_plane;
    const uint64_t g = 2 * planes + 2;
    return g < nx ? (uint32_t)g : nx;
}

// Device bytes of the adjoint with checkpoint interval K over T steps:
// ceil(T / K) checkpoints of ckpt_bytes, K moment slots of mom_bytes (rho - 1
// and m of each state a segment recomputes), four working states of
// state_bytes (the forward's and the recomputation's two, the two adjoint
// states) and `fields` bytes of per-cell fields.
inline uint64_t adjoint_plan_bytes(int32_t T, int32_t K, uint64_t ckpt_bytes, uint64_t mom_bytes, uint64_t state_bytes,
                                   uint64_t fields) {
    const uint64_t nck = (uint64_t)((T + K - 1) / K);
    return nck * ckpt_bytes + (uint64_t)K * mom_bytes + 4 * state_bytes + fields;
}

// The checkpoint interval in 1 .. T that needs the least memory (the smallest
// such K).
inline int32_t adjoint_least_memory_K(int32_t T, uint64_t ckpt_bytes, uint64_t mom_bytes, uint64_t state_bytes,
                                      uint64_t fields) {
    int32_t K = 1;
    uint64_t best = adjoint_plan_bytes(T, 1, ckpt_bytes, mom_bytes, state_bytes, fields);
    for (int32_t k = 2; k <= T; ++k) {
        const uint64_t b = adjoint_plan_bytes(T, k, ckpt_bytes, mom_bytes, state_bytes, fields);
        if (b < best) {
            best = b;
            K = k;
        }
    }
    return K;
}

}  // namespace d3q19
