// d3q19_moment_codec.hpp -- moment-coded storage kernels (gfx950).
//
// The moments are the frontier-moment-encoded variant's: per cell, the 19
// moments m = M f over the monomials of the velocity set
//     {1, cx, cy, cz, cx^2, cy^2, cz^2, cx cy, cx cz, cy cz,
//      cx^2 cy, cx^2 cz, cy^2 cx, cy^2 cz, cz^2 cx, cz^2 cy,
//      cx^2 cy^2, cx^2 cz^2, cy^2 cz^2}
// each quantized to 16 bits over a per-moment range [lo_k, hi_k]
// (packages/lattice-boltzmann-d3q19-frontier-moment-encoded:
//  include/bit_physics/lbm_d3q19_me/lbm_me.hpp:10-20, src/lbm_me.cpp:83-150,
//  shaders/lbm_encode.comp:37-51, shaders/lbm_decode.comp:33-43).
//
// D3Q19_ME16 (tag TME16, d3q19_device.hpp) is a population storage of the
// engine: the codes of the stored state's deviation moments, stepped by the
// fused step like every other storage (pull decodes, store encodes).
// D3Q19_ME16_REF (below) keeps the variant's own quantization point and
// FP64 arithmetic.  This file also holds the moment-range reduction both
// calibrations use.
#pragma once

#include "d3q19_device.hpp"
#include "d3q19_kernels.hpp"

namespace d3q19 {

// ---------------------------------------------------------------------------
// The moment-encoded variant's step in its own FP64 arithmetic (storage
// D3Q19_ME16_REF), on codes in the variant's moment-major order (k * ncell +
// cell) and FP64 populations in the reference layout (19, ncell):
//   decode  m_k = lo_k + code_k (hi_k - lo_k) / 65535,  f_i = sum_k Minv[i][k] m_k
//           (shaders/lbm_decode.comp:33-43),
//   collide BGK + Guo with the division-form equilibrium (lbm_collide.comp),
//   stream  the pull f_i(x) = f_i^post(x - c_i), then the half-way y-walls
//           (lbm_stream_bounce.comp; moving walls take apply_bounce_back_y_walls'
//           momentum term, zero for the variant's resting walls),
//   encode  m_k = sum_i M[k][i] f_i,  t = clamp((m_k - lo_k) / (hi_k - lo_k), 0, 1),
//           code_k = round(t 65535), halves away from zero
//           (shaders/lbm_encode.comp:37-51),
// every sum over all 19 terms in index order, contraction off.  M^-1 holds
// quarters, so its products are exact.  One step is two kernels: decode +
// collide (codes -> FP64 post-collision populations) and stream + walls +
// encode (-> codes).
// ---------------------------------------------------------------------------
struct MeRef {
    double lo[Q], hi[Q];
};

__device__ __forceinline__ void me_ref_decode_cell(const uint16_t* __restrict__ codes, const MeRef& r,
                                                   uint32_t ncell, uint32_t cell, double f[Q]) {
#pragma clang fp contract(off)
    double m[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k)
        m[k] = r.lo[k] + (double)codes[(size_t)k * ncell + cell] * (r.hi[k] - r.lo[k]) / 65535.0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < Q; ++k) s += ((double)ME_MINV4[i][k] * 0.25) * m[k];
        f[i] = s;
    }
}

__device__ __forceinline__ void me_ref_encode_cell(const double f[Q], const MeRef& r, uint32_t ncell,
                                                   uint32_t cell, uint16_t* __restrict__ codes) {
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        double mk = 0.0;
#pragma unroll
        for (int i = 0; i < Q; ++i) mk += (double)me_mono(k, i) * f[i];
        double t = (mk - r.lo[k]) / (r.hi[k] - r.lo[k]);
        t = t > 0.0 ? (t > 1.0 ? 1.0 : t) : 0.0;  // a NaN moment takes code 0
        codes[(size_t)k * ncell + cell] = (uint16_t)round(t * 65535.0);
    }
}

static __global__ void __launch_bounds__(256)
    k_me_ref_decode(const uint16_t* __restrict__ codes, double* __restrict__ f, MeRef r, uint32_t ncell) {
    const uint32_t cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= ncell) return;
    double fc[Q];
    me_ref_decode_cell(codes, r, ncell, cell, fc);
#pragma unroll
    for (int i = 0; i < Q; ++i) f[(size_t)i * ncell + cell] = fc[i];
}

static __global__ void __launch_bounds__(256)
    k_me_ref_encode(const double* __restrict__ f, uint16_t* __restrict__ codes, MeRef r, uint32_t ncell) {
    const uint32_t cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= ncell) return;
    double fc[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) fc[i] = f[(size_t)i * ncell + cell];
    me_ref_encode_cell(fc, r, ncell, cell, codes);
}

// decode + collide: post = collide(decode(codes)); force (3, ncell) or null.
static __global__ void __launch_bounds__(256)
    k_me_ref_decode_collide(const uint16_t* __restrict__ codes, double* __restrict__ post,
                            const double* __restrict__ force, MeRef r, uint32_t ncell, double tau) {
    const uint32_t cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= ncell) return;
    double fc[Q];
    me_ref_decode_cell(codes, r, ncell, cell, fc);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (force) {
        fx = force[cell];
        fy = force[(size_t)ncell + cell];
        fz = force[2 * (size_t)ncell + cell];
    }
    ref_collide<double, true>(fc, tau, force != nullptr, fx, fy, fz);
#pragma unroll
    for (int i = 0; i < Q; ++i) post[(size_t)i * ncell + cell] = fc[i];
}

// stream + (y-walls) + encode: codes = encode(walls(pull(post))).
struct MeRefWalls {
    int on;
    double top[3], bot[3];
};

static __global__ void __launch_bounds__(256)
    k_me_ref_stream_encode(const double* __restrict__ post, uint16_t* __restrict__ codes, MeRef r, MeRefWalls w,
                           uint32_t nx, uint32_t ny, uint32_t nz) {
#pragma clang fp contract(off)
    const uint32_t ncell = nx * ny * nz;
    const uint32_t cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= ncell) return;
    uint32_t x, y, z;
    cell_coords(cell, ny, nz, x, y, z);
    double g[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const uint32_t sx = (x + nx - CX[i]) % nx, sy = (y + ny - CY[i]) % ny, sz = (z + nz - CZ[i]) % nz;
        g[i] = post[(size_t)i * ncell + (sx * ny + sy) * nz + sz];
    }
    double o[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) o[i] = g[i];
    if (w.on && (y == 0 || y + 1 == ny)) {
        // apply_bounce_back_y_walls (reference/bgk.py:145-216) on the pulled values
        const double cs2 = 1.0 / 3.0;
        double rw = 0.0;
#pragma unroll
        for (int d = 0; d < Q; ++d) rw = rw + g[d];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (CY[i] > 0 && y == 0) {
                const double cdu = (double)CX[i] * w.bot[0] + (double)CY[i] * w.bot[1] + (double)CZ[i] * w.bot[2];
                o[i] = g[OPP[i]] + -2.0 * wgt<double>(i) * rw * cdu / cs2;
            }
            if (CY[i] < 0 && y + 1 == ny) {
                const double cdu = (double)CX[i] * w.top[0] + (double)CY[i] * w.top[1] + (double)CZ[i] * w.top[2];
                o[i] = g[OPP[i]] + -2.0 * wgt<double>(i) * rw * cdu / cs2;
            }
        }
    }
    me_ref_encode_cell(o, r, ncell, cell, codes);
}

// Per-block minimum and maximum of the 19 moments m = M f (FP64) of the
// current state f^n, for calibrating the quantization ranges.  Partials:
// [0..18][block] minima, [19..37][block] maxima.  Each thread folds `per`
// cells (x planes blockIdx.z * per .. on the row grid, cell runs of the block
// on the linear one) before the block's reduction: minima and maxima are
// exact, so the grouping does not change the result, and the 38 wavefront
// reductions (six FP64 shuffles each) are paid once per `per` cells.
template <class Tag, int BC, bool ROWS>
static __global__ void __launch_bounds__(step_block<Tag>())
    k_moment_range(const typename Codec<Tag>::raw* __restrict__ in, Params<typename Codec<Tag>::C> p,
                   double* __restrict__ partials, uint32_t per) {
    typedef typename Codec<Tag>::C C;
    constexpr int NW = step_block<Tag>() / 64;
    __shared__ double red[NW][2 * Q];
    double lo[Q], hi[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        lo[k] = 1e300;
        hi[k] = -1e300;
    }
    for (uint32_t f = 0; f < per; ++f) {
        uint32_t x, y, z;
        bool live;
        if constexpr (ROWS) {
            live = grid_cell(p.nx, p.ny, p.nz, x, y, z, blockIdx.z * (per - 1u) + f);  // x = blockIdx.z per + f
        } else {
            const uint32_t c = (blockIdx.x * per + f) * blockDim.x + threadIdx.x;
            live = c < p.ncell;
            z = live ? c % p.nz : 0u;
            y = live ? (c / p.nz) % p.ny : 0u;
            x = live ? c / p.nz / p.ny : 0u;
        }
        const uint32_t cell = live ? (x * p.ny + y) * p.nz + z : 0u;
        if (live && !((BC == BC_HALFWAY) && p.solid && p.solid[cell])) {
            C d[Q];
            if (p.local)
                load_local<Tag>(in, p, x, y, z, d);
            else
                pull<Tag, BC>(in, p, x, y, z, d);
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                double s = 0.0;
#pragma unroll
                for (int i = 0; i < Q; ++i)
                    if (me_mono(k, i) != 0) s += (double)me_mono(k, i) * (wgt<double>(i) + (double)d[i]);
                lo[k] = fmin(lo[k], s);
                hi[k] = fmax(hi[k], s);
            }
        }
    }
    const int t = threadIdx.x + blockDim.x * threadIdx.y;
    const int lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        double a = lo[k], b = hi[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            a = fmin(a, __shfl_xor(a, o, 64));
            b = fmax(b, __shfl_xor(b, o, 64));
        }
        if (lane == 0) {
            red[wave][k] = a;
            red[wave][Q + k] = b;
        }
    }
    __syncthreads();
    const int nw = (int)(blockDim.x * blockDim.y) / 64;
    if (t < 2 * Q) {
        double v = red[0][t];
        for (int w = 1; w < nw; ++w) v = t < Q ? fmin(v, red[w][t]) : fmax(v, red[w][t]);
        partials[(size_t)t * block_count() + block_linear_id()] = v;
    }
}

// Final reduction of the per-block partials: one 256-thread block per slot
// (38 blocks: 19 minima, 19 maxima).
static __global__ void __launch_bounds__(256)
    k_moment_range_final(const double* __restrict__ partials, uint32_t nblocks, double* __restrict__ out) {
    __shared__ double red[4];
    const int t = blockIdx.x;  // slot: 0..18 minima, 19..37 maxima
    const bool is_min = t < Q;
    double v = is_min ? 1e300 : -1e300;
    for (uint32_t b = threadIdx.x; b < nblocks; b += 256) {
        const double p = partials[(size_t)t * nblocks + b];
        v = is_min ? fmin(v, p) : fmax(v, p);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = is_min ? fmin(v, w) : fmax(v, w);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) v = is_min ? fmin(v, red[w]) : fmax(v, red[w]);
        out[t] = v;
    }
}

}  // namespace d3q19
