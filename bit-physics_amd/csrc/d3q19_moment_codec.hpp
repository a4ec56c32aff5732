// d3q19_moment_codec.hpp -- 16-bit moment-encoded population storage (gfx950).
//
// The persistent state of a D3Q19_ME16 lattice is, per cell, the 19 moments
// m = M f over the monomials of the velocity set
//     {1, cx, cy, cz, cx^2, cy^2, cz^2, cx cy, cx cz, cy cz,
//      cx^2 cy, cx^2 cz, cy^2 cx, cy^2 cz, cz^2 cx, cz^2 cy,
//      cx^2 cy^2, cx^2 cz^2, cy^2 cz^2}
// each quantized to 16 bits over a per-moment range [lo_k, hi_k]:
//     code_k = round((m_k - lo_k) / (hi_k - lo_k) * 65535) clamped to [0, 65535],
//     m_k    = lo_k + code_k (hi_k - lo_k) / 65535,   f = M^-1 m
// (packages/lattice-boltzmann-d3q19-frontier-moment-encoded:
//  include/bit_physics/lbm_d3q19_me/lbm_me.hpp:10-20, src/lbm_me.cpp:83-150,
//  shaders/lbm_encode.comp:37-51, shaders/lbm_decode.comp:33-43).
// The codes hold the post-stream (pre-collision) state f^n, the point where
// the reference quantizes.
//
// Here the moments are taken of the deviations d = f - w, m' = M d = m - M w,
// with the ranges shifted by M w on the host, so FP32 resolves them (m_0 is
// rho ~ 1 with a range ~1e-12 wide).  M has entries 0 and +-1 and M^-1 only
// 0, +-1/4, +-1/2, +-1 (91 non-zeros), so encoding and decoding are short
// exact sums over compile-time coefficients.
//
// A step is two passes: decode + collide (codes -> post-collision FP32 work
// lattice), then pull + boundary + encode (work lattice -> codes).  The work
// lattice uses the FP32 layout of the lattice; the codes use the same layout
// with 2-byte elements (moment k in place of direction k).
#pragma once

#include "d3q19_device.hpp"
#include "d3q19_kernels.hpp"

namespace d3q19 {

struct TME16 {};  // storage tag of the moment-encoded lattice

// Monomial k of velocity i (an integer: 0 or +-1).
__host__ __device__ constexpr int me_mono(int k, int i) {
    const int x = CX[i], y = CY[i], z = CZ[i];
    switch (k) {
        case 0: return 1;
        case 1: return x;
        case 2: return y;
        case 3: return z;
        case 4: return x * x;
        case 5: return y * y;
        case 6: return z * z;
        case 7: return x * y;
        case 8: return x * z;
        case 9: return y * z;
        case 10: return x * x * y;
        case 11: return x * x * z;
        case 12: return y * y * x;
        case 13: return y * y * z;
        case 14: return z * z * x;
        case 15: return z * z * y;
        case 16: return x * x * y * y;
        case 17: return x * x * z * z;
        default: return y * y * z * z;
    }
}

// 4 M^-1 (exact integers): f_i = sum_k ME_MINV4[i][k] m_k / 4.  Solved from
// the basis above; the host checks M M^-1 = I exactly (d3q19_moment_basis).
constexpr int ME_MINV4[Q][Q] = {
    {4, 0, 0, 0, -4, -4, -4, 0, 0, 0, 0, 0, 0, 0, 0, 0, 4, 4, 4},
    {0, 2, 0, 0, 2, 0, 0, 0, 0, 0, 0, 0, -2, 0, -2, 0, -2, -2, 0},
    {0, -2, 0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 2, 0, 2, 0, -2, -2, 0},
    {0, 0, 2, 0, 0, 2, 0, 0, 0, 0, -2, 0, 0, 0, 0, -2, -2, 0, -2},
    {0, 0, -2, 0, 0, 2, 0, 0, 0, 0, 2, 0, 0, 0, 0, 2, -2, 0, -2},
    {0, 0, 0, 2, 0, 0, 2, 0, 0, 0, 0, -2, 0, -2, 0, 0, 0, -2, -2},
    {0, 0, 0, -2, 0, 0, 2, 0, 0, 0, 0, 2, 0, 2, 0, 0, 0, -2, -2},
    {0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 1, 0, 0, 0, 1, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 1, 0, 0, -1, 0, -1, 0, 0, 0, 1, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, -1, 0, 0, -1, 0, 1, 0, 0, 0, 1, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, -1, 0, 0, 1, 0, -1, 0, 0, 0, 1, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, -1, 0, 0, -1, 0, 0, 1, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, -1, 0, 0, -1, 0, 0, 1, 0, 0, 1, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, -1, 0, 0, 1, 0, 0, -1, 0, 0, 1, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 0, 0, 1},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, -1, 0, -1, 0, 0, 1},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, -1, 0, 0, 0, -1, 0, 1, 0, 0, 1},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, -1, 0, 0, 0, 1, 0, -1, 0, 0, 1},
};

// Per-lattice quantization parameters (FP32, deviation moments) and the code
// layout (2-byte elements in the lattice's layout).
struct MeCodes {
    uint32_t dir_bytes;
    uint64_t row_bytes;
    float lo[Q];     // lo_k - (M w)_k
    float step[Q];   // (hi_k - lo_k) / 65535
    float scale[Q];  // 65535 / (hi_k - lo_k)
};

// m' = M d (contraction off: the sums are exact in the order written).
__device__ __forceinline__ void me_encode_moments(const float d[Q], float m[Q]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        float s = 0.0f;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (me_mono(k, i) > 0) s += d[i];
            if (me_mono(k, i) < 0) s -= d[i];
        }
        m[k] = s;
    }
}

// d = M^-1 m'.
__device__ __forceinline__ void me_decode_moments(const float m[Q], float d[Q]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < Q; ++k)
            if (ME_MINV4[i][k] != 0) s += (float)ME_MINV4[i][k] * m[k];
        d[i] = s * 0.25f;
    }
}

// Nearest code, clamped to [0, 65535] (a NaN moment gives code 0).
__device__ __forceinline__ uint16_t me_quantize(float m, float lo, float scale) {
    const float t = fminf(fmaxf(rintf((m - lo) * scale), 0.0f), 65535.0f);
    return (uint16_t)t;
}

template <class P>
__device__ __forceinline__ const char* me_row(const void* base, const P& p, const MeCodes& mc, uint32_t x,
                                              uint32_t y) {
    return (const char*)base + (uint64_t)(x * p.ny + y) * mc.row_bytes;
}

// Decoded deviations of a cell from its codes.
__device__ __forceinline__ void me_load(const uint16_t* codes, const Params<float>& p, const MeCodes& mc,
                                        uint32_t x, uint32_t y, uint32_t z, float d[Q]) {
    const char* row = me_row(codes, p, mc, x, y);
    float m[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const uint16_t c = *(const uint16_t*)(row + (uint32_t)k * mc.dir_bytes + z * 2u);
        m[k] = fmaf((float)c, mc.step[k], mc.lo[k]);
    }
    me_decode_moments(m, d);
}

__device__ __forceinline__ void me_store(uint16_t* codes, const Params<float>& p, const MeCodes& mc, uint32_t x,
                                         uint32_t y, uint32_t z, const float d[Q]) {
    char* row = (char*)me_row(codes, p, mc, x, y);
    float m[Q];
    me_encode_moments(d, m);
#pragma unroll
    for (int k = 0; k < Q; ++k)
        __builtin_nontemporal_store(me_quantize(m[k], mc.lo[k], mc.scale[k]),
                                    (uint16_t*)(row + (uint32_t)k * mc.dir_bytes + z * 2u));
}

// Pass 1 of a step: decode f^n, collide, store the post-collision deviations
// in the FP32 work lattice; optionally rho/u of f^n.
template <bool FORCED, bool ROWS>
static __global__ void __launch_bounds__(256)
    k_me_collide(const uint16_t* __restrict__ codes, float* __restrict__ work, Params<float> p, MeCodes mc,
                 float* rho_out, float* u_out) {
    uint32_t x, y, z;
    if (!cell_of<ROWS>(p.nx, p.ny, p.nz, p.ncell, x, y, z)) return;
    const uint32_t cell = (x * p.ny + y) * p.nz + z;
    float d[Q];
    float rho = 1.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f;
    if (p.solid && p.solid[cell]) {
#pragma unroll
        for (int i = 0; i < Q; ++i) d[i] = 0.0f;
    } else {
        me_load(codes, p, mc, x, y, z, d);
        collide<float, FORCED>(d, p, cell, rho, ux, uy, uz);
    }
    store_local<TF32>(work, p, x, y, z, d);
    if (rho_out) rho_out[cell] = rho;
    if (u_out) {
        u_out[cell] = ux;
        u_out[(size_t)p.ncell + cell] = uy;
        u_out[2 * (size_t)p.ncell + cell] = uz;
    }
}

// Pass 2: pull the post-collision work lattice, apply the boundary, encode
// f^{n+1} into the codes.
template <int BC, bool ROWS>
static __global__ void __launch_bounds__(256)
    k_me_stream_encode(const float* __restrict__ work, uint16_t* __restrict__ codes, Params<float> p,
                       MeCodes mc) {
    uint32_t x, y, z;
    if (!cell_of<ROWS>(p.nx, p.ny, p.nz, p.ncell, x, y, z)) return;
    const uint32_t cell = (x * p.ny + y) * p.nz + z;
    float d[Q];
    if (BC == BC_HALFWAY && p.solid && p.solid[cell]) {
#pragma unroll
        for (int i = 0; i < Q; ++i) d[i] = 0.0f;
    } else {
        pull<TF32, BC>(work, p, x, y, z, d);
    }
    me_store(codes, p, mc, x, y, z, d);
}

// Encode in-place deviations of the work lattice (a loaded state).
template <bool ROWS>
static __global__ void __launch_bounds__(256)
    k_me_encode_local(const float* __restrict__ work, uint16_t* __restrict__ codes, Params<float> p, MeCodes mc) {
    uint32_t x, y, z;
    if (!cell_of<ROWS>(p.nx, p.ny, p.nz, p.ncell, x, y, z)) return;
    float d[Q];
    load_local<TF32>(work, p, x, y, z, d);
    me_store(codes, p, mc, x, y, z, d);
}

// Decode the codes into in-place deviations of the work lattice (before a
// read-back, which then reads the work lattice as a loaded FP32 state).
template <bool ROWS>
static __global__ void __launch_bounds__(256)
    k_me_decode_local(const uint16_t* __restrict__ codes, float* __restrict__ work, Params<float> p, MeCodes mc) {
    uint32_t x, y, z;
    if (!cell_of<ROWS>(p.nx, p.ny, p.nz, p.ncell, x, y, z)) return;
    const uint32_t cell = (x * p.ny + y) * p.nz + z;
    float d[Q];
    if (p.solid && p.solid[cell]) {
#pragma unroll
        for (int i = 0; i < Q; ++i) d[i] = 0.0f;
    } else {
        me_load(codes, p, mc, x, y, z, d);
    }
    store_local<TF32>(work, p, x, y, z, d);
}

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
// [block][0..18] minima, [block][19..37] maxima.
template <class Tag, int BC, bool ROWS>
static __global__ void __launch_bounds__(step_block<Tag>())
    k_moment_range(const typename Codec<Tag>::raw* __restrict__ in, Params<typename Codec<Tag>::C> p,
                   double* __restrict__ partials) {
    typedef typename Codec<Tag>::C C;
    constexpr int NW = step_block<Tag>() / 64;
    __shared__ double red[NW][2 * Q];
    double lo[Q], hi[Q];
    uint32_t x, y, z;
    const bool live = cell_of<ROWS>(p.nx, p.ny, p.nz, p.ncell, x, y, z);
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
            lo[k] = hi[k] = s;
        }
    } else {
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            lo[k] = 1e300;
            hi[k] = -1e300;
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
        partials[(size_t)block_linear_id() * 2 * Q + t] = v;
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
        const double p = partials[(size_t)b * 2 * Q + t];
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
