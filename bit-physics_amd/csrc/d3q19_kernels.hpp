// d3q19_kernels.hpp -- HIP kernels of the D3Q19 engine (gfx950).
#pragma once

#include "d3q19_device.hpp"

namespace d3q19 {

constexpr int BLOCK = 256;

// rho and u of the step's pre-collision state, when asked for (FP32).
template <class C>
__device__ __forceinline__ void step_moments_out(float* rho_out, float* u_out, uint32_t ncell, uint32_t cell, C rho,
                                                 C ux, C uy, C uz) {
    if (rho_out) rho_out[cell] = (float)rho;
    if (u_out) {
        u_out[cell] = (float)ux;
        u_out[(size_t)ncell + cell] = (float)uy;
        u_out[2 * (size_t)ncell + cell] = (float)uz;
    }
}

// In-place step of a moment-coded ring (me_row_in / me_row_out): the
// blocks of plane x write into the slot of old plane q = x - gap, which
// planes q - 1, q and q + 1 pull from; they store once every block of those
// planes has consumed its pulls.  ring_done[x * ME_RING_STRIDE] counts the
// blocks of plane x whose wavefronts have all consumed their codes
// (monotonic over steps: ring_target after this step), one counter per
// 256-byte line; within a block an LDS count finds the last live wavefront,
// which adds the block.  The first gap planes write into free slots, and
// the slot of old plane 0 is free as well (its pulls read the copy
// ring_saved).  The block's thread (0, 0) loads the three counters when the
// block starts and checks them when its own cells are collided (they are
// long done: the gap spans more planes than the blocks resident at once),
// then raises an LDS flag that the other wavefronts wait on; a short count
// makes it spin, and give up (ring_err) after about 10 ms.  Waits only
// ever point at lower planes, which are dispatched earlier, and a block adds
// itself before it waits.  Measured at 256^3 (GLUPS): without any wait
// 44.8 (two lattices: 46-47); counted per wavefront on packed counters 2.3,
// per block on packed counters with every wavefront reading them 8.8 -- the
// operations on one counter line serialise; a counter per line with one
// wavefront per block reading them when it got there, 41.8.
constexpr uint32_t ME_RING_STRIDE = 64;  // uint32 per counter: one 256-byte line each

__device__ __forceinline__ bool me_ring_short(uint32_t count, uint32_t target) {
    return (int32_t)(count - target) < 0;
}

// The counts of the three planes a block of plane x waits for, loaded by the
// block's thread (0, 0) -- a cell of every block -- when the block starts,
// so that their latency hides behind the pull (0 where there is no wait).
struct MeRingDeps {
    uint32_t c[3] = {0u, 0u, 0u};
};
__device__ __forceinline__ MeRingDeps me_ring_deps(const Params<float>& p, uint32_t x) {
    MeRingDeps d;
    if (p.ring_done && x > p.ring_gap && threadIdx.x == 0 && threadIdx.y == 0) {
#pragma unroll
        for (uint32_t k = 0; k < 3; ++k)
            d.c[k] = __hip_atomic_load(p.ring_done + (size_t)(x - p.ring_gap - 1 + k) * ME_RING_STRIDE,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return d;
}

// lds[0]: wavefronts of the block done with their pulls; lds[1]: dependencies checked.
__device__ __forceinline__ void me_ring_sync(const Params<float>& p, uint32_t x, uint32_t live_waves, uint32_t* lds,
                                             const MeRingDeps& dep) {
    asm volatile("" ::: "memory");  // the pulls above are consumed by the collision
    if (__lane_id() == 0) {
        const uint32_t n = __hip_atomic_fetch_add(lds, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (n + 1 == live_waves)
            __hip_atomic_fetch_add(p.ring_done + (size_t)x * ME_RING_STRIDE, 1u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        if (x > p.ring_gap) {
            if (threadIdx.x == 0 && threadIdx.y == 0) {
                for (uint32_t k = 0; k < 3; ++k) {
                    const uint32_t* c = p.ring_done + (size_t)(x - p.ring_gap - 1 + k) * ME_RING_STRIDE;
                    uint32_t seen = dep.c[k], spins = 0;
                    while (me_ring_short(seen, p.ring_target)) {
                        if (++spins > p.ring_spins) {
                            __hip_atomic_store(p.ring_err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(4);
                        seen = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
                __hip_atomic_store(lds + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            } else {
                uint32_t spins = 0;
                while (__hip_atomic_load(lds + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0 &&
                       ++spins < 4 * p.ring_spins + 4)
                    __builtin_amdgcn_s_sleep(2);
            }
        }
    }
    asm volatile("" ::: "memory");  // no store of this wavefront above the wait
}

// The adjoint's recomputation: rho - 1 and m of the pulled f^n of a cell, the
// only part of the state its reverse step reads (k_adj_step), computed as the
// reverse step computed them from the pulled state; one 4-vector per cell,
// stored nontemporally (read once, by the reverse pass; backward 264.3 ->
// 263.1 ms at 128^3 x 2000 steps against a plain store and load).
typedef float MomF4 __attribute__((ext_vector_type(4)));
typedef double MomD4 __attribute__((ext_vector_type(4)));
template <class C>
using Vec4 = typename std::conditional<std::is_same<C, double>::value, MomD4, MomF4>::type;
template <class C>
__device__ __forceinline__ void store_adjoint_moments(const Params<C>& p, uint32_t cell, const C d[Q]) {
    C drho, mx, my, mz;
    moments(d, drho, mx, my, mz);
    const Vec4<C> m = {drho, mx, my, mz};
    __builtin_nontemporal_store(m, (Vec4<C>*)p.mom + cell);
}

// ---------------------------------------------------------------------------
// The hot path: one fused pull-stream / boundary / moments / BGK+Guo pass.
// Reads the post-collision state of step n-1, writes that of step n, and
// (optionally) rho and u of the reference state f^n from the same pass.
// ---------------------------------------------------------------------------
// MOM: also store the moments of the pulled f^n (store_adjoint_moments).
template <class Tag, int BC, bool FORCED, bool ROWS = true, bool MOM = false>
static __global__ void __launch_bounds__(fused_block<Tag>(), fused_min_waves<Tag>())
    k_step(const typename Codec<Tag>::raw* __restrict__ in, typename Codec<Tag>::raw* __restrict__ out,
           Params<typename Codec<Tag>::C> p, float* rho_out, float* u_out) {
    typedef typename Codec<Tag>::C C;
    constexpr bool RING = std::is_same<Tag, TME16>::value && ROWS;  // a moment-coded ring may step in place
    __shared__ uint32_t ring_lds[RING ? 2 : 1];
    if constexpr (RING) {
        if (p.ring_done) {  // uniform: every thread of the block, before any returns
            if (threadIdx.x == 0 && threadIdx.y == 0) ring_lds[0] = ring_lds[1] = 0u;
            __syncthreads();
        }
    }
    uint32_t x, y, z;
    if (!cell_of<ROWS>(p.nx, p.ny, p.nz, p.ncell, x, y, z, p.x0)) return;
    MeRingDeps ring_deps;
    if constexpr (RING) ring_deps = me_ring_deps(p, x);
    const uint32_t cell = (x * p.ny + y) * p.nz + z;
    // Wavefront-uniform segment bases (seg_base) for 16-bit storage on the
    // row grid; for FP32 the same store form made the periodic step slower
    // (256^3: 40.4 -> 38.4 GLUPS), so 4- and 8-byte storage keep the
    // compiler's addressing.
    constexpr bool U = ROWS && sizeof(typename Codec<Tag>::raw) == 2;
    C d[Q];
    C rho = C(1), ux = C(0), uy = C(0), uz = C(0);
    if constexpr (BC == BC_HALFWAY && sizeof(typename Codec<Tag>::raw) == 2 && !std::is_same<Tag, TME16>::value) {
        // 16-bit storage without solids: interior rows and edge rows each
        // with a collision and store of their own, compiled apart from each
        // other and from the general path below (fewer live scalar registers
        // in the interior rows' code).  Measured against one shared path,
        // 256^3 FP16, GLUPS: lid-driven cavity 73.0 -> 75.9, half-way z walls
        // 73.2 -> 76.4; with the uniform segment bases, reference y-walls
        // 79.6 -> 82.9.  (FP32 gained nothing at 256^3 and lost 11% on the
        // latency-bound 64^3 cavity: 4- and 8-byte storage keep one path.)
        if (!p.solid) {
            if (!halfway_edge_row(p, x, y)) {
                pull_halfway_fast<Tag, false, U>(in, p, x, y, z, d);
                collide<C, FORCED>(d, p, cell, rho, ux, uy, uz);
                store_local<Tag, U>(out, p, x, y, z, d);
                step_moments_out(rho_out, u_out, p.ncell, cell, rho, ux, uy, uz);
                return;
            }
            pull_halfway_edge<Tag, false, U>(in, p, x, y, z, d);
            collide<C, FORCED>(d, p, cell, rho, ux, uy, uz);
            store_local<Tag, U>(out, p, x, y, z, d);
            step_moments_out(rho_out, u_out, p.ncell, cell, rho, ux, uy, uz);
            return;
        }
    }
    if constexpr (RING) {
        if (BC == BC_HALFWAY && p.solid && p.solid[cell]) {
#pragma unroll
            for (int i = 0; i < Q; ++i) d[i] = C(0);
        } else {
            pull<Tag, BC, false, U>(in, p, x, y, z, d);
            collide<C, FORCED>(d, p, cell, rho, ux, uy, uz);
        }
        if (p.ring_done) {
            // live wavefronts of the block: its rows below ny times the wavefronts of its z segment below nz
            const uint32_t rows = min(blockDim.y, p.ny - blockIdx.y * blockDim.y);
            const uint32_t zs = min(blockDim.x, p.nz - blockIdx.x * blockDim.x);
            me_ring_sync(p, x, rows * ((zs + 63u) / 64u), ring_lds, ring_deps);
        }
    } else {
        if (BC == BC_HALFWAY && p.solid && p.solid[cell]) {
#pragma unroll
            for (int i = 0; i < Q; ++i) d[i] = C(0);
        } else {
            pull<Tag, BC, false, U>(in, p, x, y, z, d);
            if constexpr (MOM) store_adjoint_moments(p, cell, d);
            collide<C, FORCED>(d, p, cell, rho, ux, uy, uz);
        }
    }
    store_local<Tag, U>(out, p, x, y, z, d);
    step_moments_out(rho_out, u_out, p.ncell, cell, rho, ux, uy, uz);
}

// Load a reference-layout population field f^n as the lattice state, in
// place (pre-collision form): stored deviations d_i = f_i - w_i.
template <class Tag, class In, bool ROWS = true>
static __global__ void __launch_bounds__(step_block<Tag>())
    k_load(const In* __restrict__ f, typename Codec<Tag>::raw* __restrict__ out, Params<typename Codec<Tag>::C> p) {
    typedef typename Codec<Tag>::C C;
    uint32_t x, y, z;
    if (!cell_of<ROWS>(p.nx, p.ny, p.nz, p.ncell, x, y, z)) return;
    const uint32_t cell = (x * p.ny + y) * p.nz + z;
    const bool solid = p.solid && p.solid[cell];
    C d[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i)
        d[i] = solid ? C(0) : (C)((double)f[(size_t)i * p.ncell + cell] - wgt<double>(i));
    store_local<Tag>(out, p, x, y, z, d);
}

// First step after a load: collide the in-place f^n (no streaming) into the
// post-collision storage the fused step consumes; optionally write rho/u of f^n.
// MOM: also store the moments of f^n (store_adjoint_moments).
template <class Tag, bool ROWS = true, bool MOM = false>
static __global__ void __launch_bounds__(step_block<Tag>())
    k_collide_local(const typename Codec<Tag>::raw* __restrict__ in, typename Codec<Tag>::raw* __restrict__ out,
                    Params<typename Codec<Tag>::C> p, float* rho_out, float* u_out) {
    typedef typename Codec<Tag>::C C;
    constexpr bool U = ROWS && sizeof(typename Codec<Tag>::raw) == 2;  // as k_step
    uint32_t x, y, z;
    if (!cell_of<ROWS>(p.nx, p.ny, p.nz, p.ncell, x, y, z)) return;
    const uint32_t cell = (x * p.ny + y) * p.nz + z;
    C d[Q];
    C rho = C(1), ux = C(0), uy = C(0), uz = C(0);
    if (p.solid && p.solid[cell]) {
#pragma unroll
        for (int i = 0; i < Q; ++i) d[i] = C(0);
    } else {
        load_local<Tag, U>(in, p, x, y, z, d);
        if constexpr (MOM) store_adjoint_moments(p, cell, d);
        collide(d, p, cell, rho, ux, uy, uz);
    }
    store_local<Tag, U>(out, p, x, y, z, d);
    if (rho_out) rho_out[cell] = (float)rho;
    if (u_out) {
        u_out[cell] = (float)ux;
        u_out[(size_t)p.ncell + cell] = (float)uy;
        u_out[2 * (size_t)p.ncell + cell] = (float)uz;
    }
}

// Pre-collision deviations of the current state at a cell: in place after a
// load, else pull + boundary of the stored post-collision state.
template <class Tag, int BC>
__device__ __forceinline__ void state_at(const void* in, const Params<typename Codec<Tag>::C>& p, uint32_t x,
                                         uint32_t y, uint32_t z, typename Codec<Tag>::C d[Q]) {
    if (p.local)
        load_local<Tag>(in, p, x, y, z, d);
    else
        pull<Tag, BC>(in, p, x, y, z, d);
}

// Read-back: the reference's pre-collision f^n and/or its moments, in the
// reference's (19, Nx, Ny, Nz) / (Nx, Ny, Nz) / (3, Nx, Ny, Nz) layouts.
template <class Tag, int BC, class Out, bool ROWS = true>
static __global__ void __launch_bounds__(step_block<Tag>())
    k_readback(const typename Codec<Tag>::raw* __restrict__ in, Params<typename Codec<Tag>::C> p,
               Out* __restrict__ f_out, Out* __restrict__ rho_out, Out* __restrict__ u_out) {
    typedef typename Codec<Tag>::C C;
    uint32_t x, y, z;
    if (!cell_of<ROWS>(p.nx, p.ny, p.nz, p.ncell, x, y, z)) return;
    const uint32_t cell = (x * p.ny + y) * p.nz + z;
    C d[Q];
    const bool solid = (BC == BC_HALFWAY) && p.solid && p.solid[cell];
    if (solid) {
#pragma unroll
        for (int i = 0; i < Q; ++i) d[i] = C(0);
    } else {
        state_at<Tag, BC>(in, p, x, y, z, d);
    }
    if (f_out) {
#pragma unroll
        for (int i = 0; i < Q; ++i) f_out[(size_t)i * p.ncell + cell] = (Out)(wgt<double>(i) + (double)d[i]);
    }
    if (rho_out || u_out) {
        C rho, ux, uy, uz;
        readback_moments(d, p, cell, rho, ux, uy, uz);
        if (solid) ux = uy = uz = C(0);
        if (rho_out) rho_out[cell] = (Out)rho;
        if (u_out) {
            u_out[cell] = (Out)ux;
            u_out[(size_t)p.ncell + cell] = (Out)uy;
            u_out[2 * (size_t)p.ncell + cell] = (Out)uz;
        }
    }
}

// FP64 diagnostics of f^n: per-block partial sums of
// [mass, px, py, pz, kinetic energy], summed in a fixed order afterwards.
constexpr int NDIAG = 5;

template <class Tag, int BC, bool ROWS = true>
static __global__ void __launch_bounds__(step_block<Tag>())
    k_diag(const typename Codec<Tag>::raw* __restrict__ in, Params<typename Codec<Tag>::C> p,
           double* __restrict__ partials) {
    typedef typename Codec<Tag>::C C;
    constexpr int NW = step_block<Tag>() / 64;
    __shared__ double red[NW][NDIAG];
    double v[NDIAG] = {0, 0, 0, 0, 0};
    uint32_t x, y, z;
    const bool live = cell_of<ROWS>(p.nx, p.ny, p.nz, p.ncell, x, y, z);
    const uint32_t cell = (x * p.ny + y) * p.nz + z;
    if (live && !((BC == BC_HALFWAY) && p.solid && p.solid[cell])) {
        C d[Q];
        state_at<Tag, BC>(in, p, x, y, z, d);
        C drho, mx, my, mz, rho, ux, uy, uz;
        moments(d, drho, mx, my, mz);
        readback_moments(d, p, cell, rho, ux, uy, uz);
        v[0] = 1.0 + (double)drho;
        v[1] = (double)mx;
        v[2] = (double)my;
        v[3] = (double)mz;
        v[4] = 0.5 * (double)rho * ((double)ux * ux + (double)uy * uy + (double)uz * uz);
    }
    const int t = threadIdx.x + blockDim.x * threadIdx.y;
    const int lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int k = 0; k < NDIAG; ++k) {
        double s = wave_sum(v[k]);
        if (lane == 0) red[wave][k] = s;
    }
    __syncthreads();
    if (t < NDIAG) {
        double s = 0.0;
        for (int w = 0; w < NW; ++w) s += red[w][t];
        partials[(size_t)t * block_count() + block_linear_id()] = s;
    }
}

// Health of f^n (the runners' tier-1 check, NP/sim.py:135-145, on the GPU):
// per-block [non-finite cells, min rho, max rho, max |u|]; a cell is
// non-finite when any population, rho or u is NaN or infinite (its rho and u
// are then left out of the extremes).
constexpr int NHEALTH = 4;

template <class Tag, int BC, bool ROWS = true>
static __global__ void __launch_bounds__(step_block<Tag>())
    k_health(const typename Codec<Tag>::raw* __restrict__ in, Params<typename Codec<Tag>::C> p,
             double* __restrict__ partials) {
    typedef typename Codec<Tag>::C C;
    constexpr int NW = step_block<Tag>() / 64;
    __shared__ double red[NW][NHEALTH];
    double v[NHEALTH] = {0.0, 1e300, -1e300, 0.0};
    uint32_t x, y, z;
    const bool live = cell_of<ROWS>(p.nx, p.ny, p.nz, p.ncell, x, y, z);
    const uint32_t cell = live ? (x * p.ny + y) * p.nz + z : 0u;
    if (live && !((BC == BC_HALFWAY) && p.solid && p.solid[cell])) {
        C d[Q];
        state_at<Tag, BC>(in, p, x, y, z, d);
        bool finite = true;
#pragma unroll
        for (int i = 0; i < Q; ++i) finite = finite && isfinite((double)d[i]);
        C rho, ux, uy, uz;
        readback_moments(d, p, cell, rho, ux, uy, uz);
        const double u2 = (double)ux * ux + (double)uy * uy + (double)uz * uz;
        finite = finite && isfinite((double)rho) && isfinite(u2);
        if (finite) {
            v[1] = v[2] = (double)rho;
            v[3] = sqrt(u2);
        } else {
            v[0] = 1.0;
        }
    }
    const int t = threadIdx.x + blockDim.x * threadIdx.y;
    const int lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v[0] += __shfl_xor(v[0], o, 64);
        v[1] = fmin(v[1], __shfl_xor(v[1], o, 64));
        v[2] = fmax(v[2], __shfl_xor(v[2], o, 64));
        v[3] = fmax(v[3], __shfl_xor(v[3], o, 64));
    }
    if (lane == 0)
        for (int k = 0; k < NHEALTH; ++k) red[wave][k] = v[k];
    __syncthreads();
    if (t == 0) {
        const int nw = (int)(blockDim.x * blockDim.y) / 64;
        for (int w = 1; w < nw; ++w) {
            v[0] += red[w][0];
            v[1] = fmin(v[1], red[w][1]);
            v[2] = fmax(v[2], red[w][2]);
            v[3] = fmax(v[3], red[w][3]);
        }
        for (int k = 0; k < NHEALTH; ++k) partials[(size_t)k * block_count() + block_linear_id()] = v[k];
    }
}

// Final reduction of the health partials: one 256-thread block per slot.
[[maybe_unused]] static __global__ void __launch_bounds__(256)
    k_health_final(const double* __restrict__ partials, uint32_t nblocks, double* __restrict__ out) {
    __shared__ double red[4];
    const int k = blockIdx.x;
    double v = k == 1 ? 1e300 : (k == 2 ? -1e300 : 0.0);
    for (uint32_t b = threadIdx.x; b < nblocks; b += 256) {
        const double q = partials[(size_t)k * nblocks + b];
        v = k == 0 ? v + q : (k == 1 ? fmin(v, q) : fmax(v, q));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double q = __shfl_xor(v, o, 64);
        v = k == 0 ? v + q : (k == 1 ? fmin(v, q) : fmax(v, q));
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) v = k == 0 ? v + red[w] : (k == 1 ? fmin(v, red[w]) : fmax(v, red[w]));
        out[k] = v;
    }
}

// Final sum of per-block partials: one 256-thread block per component, each
// summing in a fixed order (strided per thread, then a fixed tree), so the
// result is the same bit for bit on every run.
[[maybe_unused]] static __global__ void __launch_bounds__(BLOCK)
    k_sum_partials(const double* __restrict__ partials, uint32_t nblocks, int ncomp,
                   double* __restrict__ out) {
    __shared__ double red[BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = blockIdx.x;
    if (k >= ncomp) return;
    double s = 0.0;
    for (uint32_t b = threadIdx.x; b < nblocks; b += BLOCK) s += partials[(size_t)k * nblocks + b];
    s = wave_sum(s);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < BLOCK / 64; ++w) t += red[w];
        out[k] = t;
    }
}

// FP64 enstrophy partials 0.5 |curl u|^2 with periodic central differences,
// from the read-back velocity u (3, ncell) in the lattice's compute type T.
template <class T>
[[maybe_unused]] static __global__ void __launch_bounds__(BLOCK)
    k_enstrophy(const T* __restrict__ u, uint32_t nx, uint32_t ny, uint32_t nz,
                double* __restrict__ partials) {
    __shared__ double red[BLOCK / 64];
    const uint32_t ncell = nx * ny * nz;
    const uint32_t cell = blockIdx.x * BLOCK + threadIdx.x;
    double v = 0.0;
    if (cell < ncell) {
        uint32_t x, y, z;
        cell_coords(cell, ny, nz, x, y, z);
        const uint32_t nyz = ny * nz;
        const uint32_t xm = (x == 0 ? nx - 1 : x - 1), xp = (x + 1 == nx ? 0 : x + 1);
        const uint32_t ym = (y == 0 ? ny - 1 : y - 1), yp = (y + 1 == ny ? 0 : y + 1);
        const uint32_t zm = (z == 0 ? nz - 1 : z - 1), zp = (z + 1 == nz ? 0 : z + 1);
        const T* ux = u;
        const T* uy = u + ncell;
        const T* uz = u + 2 * (size_t)ncell;
        auto at = [&](const T* a, uint32_t i, uint32_t j, uint32_t k) {
            return (double)a[i * nyz + j * nz + k];
        };
        const double dwy_dz = 0.5 * (at(uy, x, y, zp) - at(uy, x, y, zm));
        const double dwz_dy = 0.5 * (at(uz, x, yp, z) - at(uz, x, ym, z));
        const double dwx_dz = 0.5 * (at(ux, x, y, zp) - at(ux, x, y, zm));
        const double dwz_dx = 0.5 * (at(uz, xp, y, z) - at(uz, xm, y, z));
        const double dwy_dx = 0.5 * (at(uy, xp, y, z) - at(uy, xm, y, z));
        const double dwx_dy = 0.5 * (at(ux, x, yp, z) - at(ux, x, ym, z));
        const double ox = dwz_dy - dwy_dz, oy = dwx_dz - dwz_dx, oz = dwy_dx - dwx_dy;
        v = 0.5 * (ox * ox + oy * oy + oz * oz);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double s = wave_sum(v);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < BLOCK / 64; ++w) t += red[w];
        partials[blockIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------
// Stateless drop-ins on reference-layout arrays (unshifted populations).
// FP64 instantiations follow the NumPy reference's operation order term by
// term with contraction off, so they agree with it to the last bit.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void ref_moments(const T fc[Q], T& r, T& mx, T& my, T& mz) {
#pragma clang fp contract(off)
    T a = T(0), bx = T(0), by = T(0), bz = T(0);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        a = a + fc[i];
        bx = bx + T(CX[i]) * fc[i];
        by = by + T(CY[i]) * fc[i];
        bz = bz + T(CZ[i]) * fc[i];
    }
    r = a;
    mx = bx;
    my = by;
    mz = bz;
}

template <class T> struct RefConst {
    T cs2, inv_cs2, inv_cs4, inv_two_cs4, inv_two_cs2;
};

template <class T>
__host__ __device__ inline RefConst<T> ref_const() {
    RefConst<T> k;
    k.cs2 = T(1) / T(3);
    k.inv_cs2 = T(1) / k.cs2;
    k.inv_cs4 = T(1) / (k.cs2 * k.cs2);
    k.inv_two_cs4 = T(1) / (T(2) * k.cs2 * k.cs2);
    k.inv_two_cs2 = T(1) / (T(2) * k.cs2);
    return k;
}

template <class T>
__device__ __forceinline__ T rho_floor(T r) {
    return (r != r) ? r : (r > T(1e-30) ? r : T(1e-30));
}

// reference/equilibrium.py:90-98, one direction
template <class T>
__device__ __forceinline__ T ref_feq(int i, T r, T ux, T uy, T uz, T u_sq, const RefConst<T>& k) {
#pragma clang fp contract(off)
    T cu = T(CX[i]) * ux + T(CY[i]) * uy + T(CZ[i]) * uz;
    return wgt<T>(i) * r * (T(1) + cu * k.inv_cs2 + cu * cu * k.inv_two_cs4 - u_sq * k.inv_two_cs2);
}

// reference/equilibrium.py:34-53 (the point function feq), one direction: the
// same polynomial as ref_feq with the reference's divisions by c_s^2,
// 2 c_s^4 and 2 c_s^2 instead of reciprocal multiplies; the two forms differ
// in the last bit for about a quarter of random (rho, u).
template <class T>
__device__ __forceinline__ T ref_feq_point(int i, T r, T ux, T uy, T uz, T u_sq) {
#pragma clang fp contract(off)
    const T cs2 = T(1) / T(3);
    T cu = T(CX[i]) * ux + T(CY[i]) * uy + T(CZ[i]) * uz;
    return wgt<T>(i) * r * (T(1) + cu / cs2 + (cu * cu) / (T(2) * cs2 * cs2) - u_sq / (T(2) * cs2));
}

// reference/bgk.py:72-110 collision of one cell.  POINT: the equilibrium in
// the point function's division form (ref_feq_point), which is the
// moment-encoded variant's collision shader's (shaders/lbm_collide.comp:58-65;
// its u_eq and Guo term are bgk.py's with the force (force_x, 0, 0)).
template <class T, bool POINT = false>
__device__ __forceinline__ void ref_collide(T fc[Q], T tau, bool has_force, T fx, T fy, T fz) {
#pragma clang fp contract(off)
    const RefConst<T> k = ref_const<T>();
    T r, mx, my, mz;
    ref_moments(fc, r, mx, my, mz);
    T rs = rho_floor(r);
    T ux = mx / rs, uy = my / rs, uz = mz / rs;
    if (has_force) {
        ux = ux + T(0.5) * fx / rs;
        uy = uy + T(0.5) * fy / rs;
        uz = uz + T(0.5) * fz / rs;
    }
    T u_sq = ux * ux + uy * uy + uz * uz;
    T pref = T(1) - T(0.5) / tau;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        T feq = POINT ? ref_feq_point(i, r, ux, uy, uz, u_sq) : ref_feq(i, r, ux, uy, uz, u_sq, k);
        T fp = fc[i] - (fc[i] - feq) / tau;
        if (has_force) {
            T c0 = T(CX[i]), c1 = T(CY[i]), c2 = T(CZ[i]);
            T cdu = c0 * ux + c1 * uy + c2 * uz;
            T tx = (c0 - ux) * k.inv_cs2 + cdu * c0 * k.inv_cs4;
            T ty = (c1 - uy) * k.inv_cs2 + cdu * c1 * k.inv_cs4;
            T tz = (c2 - uz) * k.inv_cs2 + cdu * c2 * k.inv_cs4;
            T ft = pref * wgt<T>(i) * (tx * fx + ty * fy + tz * fz);
            fp = fp + ft;
        }
        fc[i] = fp;
    }
}

// bgk_step: collide, then push along c_i (== np.roll by +c_i).
template <class T>
static __global__ void __launch_bounds__(BLOCK)
    k_fn_bgk_step(const T* __restrict__ f, T* __restrict__ out, const T* __restrict__ force,
                  uint32_t nx, uint32_t ny, uint32_t nz, T tau) {
    const uint32_t ncell = nx * ny * nz;
    const uint32_t cell = blockIdx.x * BLOCK + threadIdx.x;
    if (cell >= ncell) return;
    uint32_t x, y, z;
    cell_coords(cell, ny, nz, x, y, z);
    T fc[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) fc[i] = f[(size_t)i * ncell + cell];
    T fx = T(0), fy = T(0), fz = T(0);
    if (force) {
        fx = force[cell];
        fy = force[(size_t)ncell + cell];
        fz = force[2 * (size_t)ncell + cell];
    }
    ref_collide(fc, tau, force != nullptr, fx, fy, fz);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const uint32_t dx = (x + nx + CX[i]) % nx, dy = (y + ny + CY[i]) % ny, dz = (z + nz + CZ[i]) % nz;
        out[(size_t)i * ncell + (dx * ny + dy) * nz + dz] = fc[i];
    }
}

// stream: out[i, x] = in[i, x - c_i] (periodic).
template <class T>
static __global__ void __launch_bounds__(BLOCK)
    k_fn_stream(const T* __restrict__ f, T* __restrict__ out, uint32_t nx, uint32_t ny, uint32_t nz) {
    const uint32_t ncell = nx * ny * nz;
    const uint32_t cell = blockIdx.x * BLOCK + threadIdx.x;
    if (cell >= ncell) return;
    uint32_t x, y, z;
    cell_coords(cell, ny, nz, x, y, z);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const uint32_t sx = (x + nx - CX[i]) % nx, sy = (y + ny - CY[i]) % ny, sz = (z + nz - CZ[i]) % nz;
        out[(size_t)i * ncell + cell] = f[(size_t)i * ncell + (sx * ny + sy) * nz + sz];
    }
}

// apply_bounce_back_y_walls (reference/bgk.py:145-216).
template <class T>
static __global__ void __launch_bounds__(BLOCK)
    k_fn_bb_ywalls(const T* __restrict__ f, T* __restrict__ out, uint32_t nx, uint32_t ny,
                   uint32_t nz, T t0, T t1, T t2, T b0, T b1, T b2) {
#pragma clang fp contract(off)
    const uint32_t ncell = nx * ny * nz;
    const uint32_t cell = blockIdx.x * BLOCK + threadIdx.x;
    if (cell >= ncell) return;
    uint32_t x, y, z;
    cell_coords(cell, ny, nz, x, y, z);
    T fc[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) fc[i] = f[(size_t)i * ncell + cell];
    T oc[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) oc[i] = fc[i];
    const T cs2 = T(1) / T(3);
    if (y == 0 || y + 1 == ny) {
        T rw = T(0);
#pragma unroll
        for (int d = 0; d < Q; ++d) rw = rw + fc[d];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (CY[i] > 0 && y == 0) {
                T cdu = T(CX[i]) * b0 + T(CY[i]) * b1 + T(CZ[i]) * b2;
                T inj = T(-2) * wgt<T>(i) * rw * cdu / cs2;
                oc[i] = fc[OPP[i]] + inj;
            }
            if (CY[i] < 0 && y + 1 == ny) {
                T cdu = T(CX[i]) * t0 + T(CY[i]) * t1 + T(CZ[i]) * t2;
                T inj = T(-2) * wgt<T>(i) * rw * cdu / cs2;
                oc[i] = fc[OPP[i]] + inj;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) out[(size_t)i * ncell + cell] = oc[i];
}

// feq_field (reference/equilibrium.py:71-98); rho == nullptr means rho = 1.
// POINT: the point function's operation order (equilibrium.py:34-53) at each
// of the ncell (rho, u) pairs.
template <class T, bool POINT = false>
static __global__ void __launch_bounds__(BLOCK)
    k_fn_feq(const T* __restrict__ rho, const T* __restrict__ u, T* __restrict__ out, uint32_t ncell) {
#pragma clang fp contract(off)
    const uint32_t cell = blockIdx.x * BLOCK + threadIdx.x;
    if (cell >= ncell) return;
    const RefConst<T> k = ref_const<T>();
    const T ux = u[cell], uy = u[(size_t)ncell + cell], uz = u[2 * (size_t)ncell + cell];
    const T u_sq = ux * ux + uy * uy + uz * uz;
    const T r = rho ? rho[cell] : T(1);
#pragma unroll
    for (int i = 0; i < Q; ++i)
        out[(size_t)i * ncell + cell] = POINT ? ref_feq_point(i, r, ux, uy, uz, u_sq)
                                              : ref_feq(i, r, ux, uy, uz, u_sq, k);
}

// set_equilibrium in one pass: the state k_fn_feq followed by k_load would
// give (each f_i^eq rounded to In, then f_i - w_i), with no (19, ncell)
// staging array -- which for the largest lattices would not fit beside them.
template <class Tag, class In, bool ROWS = true>
static __global__ void __launch_bounds__(step_block<Tag>())
    k_load_eq(const In* __restrict__ rho, const In* __restrict__ u, typename Codec<Tag>::raw* __restrict__ out,
              Params<typename Codec<Tag>::C> p) {
#pragma clang fp contract(off)
    typedef typename Codec<Tag>::C C;
    uint32_t x, y, z;
    if (!cell_of<ROWS>(p.nx, p.ny, p.nz, p.ncell, x, y, z)) return;
    const uint32_t cell = (x * p.ny + y) * p.nz + z;
    const bool solid = p.solid && p.solid[cell];
    const RefConst<In> k = ref_const<In>();
    const In ux = u[cell], uy = u[(size_t)p.ncell + cell], uz = u[2 * (size_t)p.ncell + cell];
    const In u_sq = ux * ux + uy * uy + uz * uz;
    const In r = rho ? rho[cell] : In(1);
    C d[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const In feq = ref_feq(i, r, ux, uy, uz, u_sq, k);
        d[i] = solid ? C(0) : (C)((double)feq - wgt<double>(i));
    }
    store_local<Tag>(out, p, x, y, z, d);
}

// density_field / momentum_field / macroscopic_velocity; any output may be null.
template <class T>
static __global__ void __launch_bounds__(BLOCK)
    k_fn_moments(const T* __restrict__ f, const T* __restrict__ force, T* __restrict__ rho_out,
                 T* __restrict__ mom_out, T* __restrict__ u_out, uint32_t ncell) {
#pragma clang fp contract(off)
    const uint32_t cell = blockIdx.x * BLOCK + threadIdx.x;
    if (cell >= ncell) return;
    T fc[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) fc[i] = f[(size_t)i * ncell + cell];
    T r, mx, my, mz;
    ref_moments(fc, r, mx, my, mz);
    if (rho_out) rho_out[cell] = r;
    if (mom_out) {
        mom_out[cell] = mx;
        mom_out[(size_t)ncell + cell] = my;
        mom_out[2 * (size_t)ncell + cell] = mz;
    }
    if (u_out) {
        const T rs = rho_floor(r);
        if (force) {
            mx = mx + T(0.5) * force[cell];
            my = my + T(0.5) * force[(size_t)ncell + cell];
            mz = mz + T(0.5) * force[2 * (size_t)ncell + cell];
        }
        u_out[cell] = mx / rs;
        u_out[(size_t)ncell + cell] = my / rs;
        u_out[2 * (size_t)ncell + cell] = mz / rs;
    }
}

}  // namespace d3q19
