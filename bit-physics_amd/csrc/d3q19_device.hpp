// d3q19_device.hpp -- device-side building blocks of the D3Q19 engine (gfx950).
//
// Lattice: lexicographic D3Q19 velocity set of the reference
// (packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/reference/constants.py:39-69),
// opposite map of reference/bgk.py:164-187.
//
// Populations live on the GPU as deviations d_i = f_i - w_i ("shifted"
// storage), in one of four storage formats; all collision arithmetic is in
// the compute type C (float, or double for FP64 storage).  Storing the
// deviation keeps the O(1e-3) signal instead of the O(0.1) rest weight in the
// mantissa, which is what makes 16-bit storage usable and FP32 storage
// accurate to ~1e-10 absolute per step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace d3q19 {

constexpr int Q = 19;
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

template <class T>
__host__ __device__ constexpr T wgt(int i) {
    return i == 0 ? T(1) / T(3) : (i <= 6 ? T(1) / T(18) : T(1) / T(36));
}

// c_i . (x, y, z) with the zero components left out (every component of c_i
// is 0 or +-1), so no product with zero is formed.
template <class C>
__host__ __device__ __forceinline__ C cdot(int i, C x, C y, C z) {
    C r = C(0);
    bool any = false;
    if (CX[i] != 0) {
        r = CX[i] > 0 ? x : -x;
        any = true;
    }
    if (CY[i] != 0) {
        r = any ? (CY[i] > 0 ? r + y : r - y) : (CY[i] > 0 ? y : -y);
        any = true;
    }
    if (CZ[i] != 0) r = any ? (CZ[i] > 0 ? r + z : r - z) : (CZ[i] > 0 ? z : -z);
    return r;
}

enum { BC_PERIODIC = 0, BC_YWALLS = 1, BC_HALFWAY = 2 };
enum { FORCE_NONE = 0, FORCE_UNIFORM = 1, FORCE_FIELD = 2 };

// ---------------------------------------------------------------------------
// Storage formats.  raw = element type in HBM, C = compute type.
// ---------------------------------------------------------------------------
struct TF64 {};
struct TF32 {};
struct TF16 {};   // IEEE half of (f_i - w_i) * (2048 / w_i)
struct TBF16 {};  // bfloat16 of (f_i - w_i), rounded to nearest even
struct TBF16S {}; // bfloat16 of (f_i - w_i), stochastically rounded
struct TME16 {};  // 16-bit codes of the 19 moments of (f - w) (moment-encoded storage)

template <class Tag> struct Codec;

template <> struct Codec<TF64> {
    typedef double raw;
    typedef double C;
    __device__ static __forceinline__ double dec(raw v, int) { return v; }
    __device__ static __forceinline__ raw enc(double d, int) { return d; }
};

template <> struct Codec<TF32> {
    typedef float raw;
    typedef float C;
    __device__ static __forceinline__ float dec(raw v, int) { return v; }
    __device__ static __forceinline__ raw enc(float d, int) { return d; }
};

// Half: the deviation relative to w_i, times 2^11, so that relative
// deviations from 3e-8 to 32 stay in the normal FP16 range.
template <> struct Codec<TF16> {
    typedef uint16_t raw;
    typedef float C;
    __host__ __device__ static constexpr float scale(int i) {
        return i == 0 ? 3.0f * 2048.0f : (i <= 6 ? 18.0f * 2048.0f : 36.0f * 2048.0f);
    }
    __device__ static __forceinline__ float dec(raw v, int i) {
        _Float16 h = __builtin_bit_cast(_Float16, v);
        return (float)h * (1.0f / scale(i));
    }
    __device__ static __forceinline__ raw enc(float d, int i) {
        _Float16 h = (_Float16)(d * scale(i));  // round to nearest even
        return __builtin_bit_cast(uint16_t, h);
    }
};

template <> struct Codec<TBF16> {
    typedef uint16_t raw;
    typedef float C;
    __device__ static __forceinline__ float dec(raw v, int) {
        return __builtin_bit_cast(float, (uint32_t)v << 16);
    }
    __device__ static __forceinline__ raw enc(float d, int) {
        uint32_t b = __builtin_bit_cast(uint32_t, d);
        if ((b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu)) return (raw)0x7fc0u;  // NaN
        b += 0x7fffu + ((b >> 16) & 1u);  // round to nearest even
        return (raw)(b >> 16);
    }
    // Stochastic rounding with a uniform 16-bit r: rounds the magnitude up
    // with probability (dropped bits) / 2^16, so the stored value is unbiased.
    // Branch-free: Inf has zero low bits and stays Inf; NaN becomes the quiet NaN.
    __device__ static __forceinline__ raw enc_sr(float d, uint32_t r) {
        const uint32_t b = __builtin_bit_cast(uint32_t, d);
        const uint32_t v = (b + (r & 0xffffu)) >> 16;
        return (raw)(d != d ? 0x7fc0u : v);
    }
};

// Same format as TBF16; stores round stochastically (store_local).
template <> struct Codec<TBF16S> : Codec<TBF16> {};

// Moment codes: element k of a cell's 19 slots holds the 16-bit code of
// moment k of its deviations; populations are decoded and encoded per cell
// (load_local / store_local / pull below), never one element at a time.
template <> struct Codec<TME16> {
    typedef uint16_t raw;
    typedef float C;
};

// Quantization constants of a moment-encoded lattice (device memory, FP32),
// for deviation moments m'_k = sum_i M[k][i] (f_i - w_i) on the grid
// m'_k = lo'_k + code_k step_k, lo'_k = lo_k - (M w)_k rounded to FP32,
// step_k = the power of two at or above (hi_k - lo_k) / 65535:
//   decode  d_i = mag_i * sum_k sgn(Minv[i][k]) step_k code_k + L_i + Ll_i,
//           L_i + Ll_i = sum_k Minv[i][k] lo'_k as a two-float sum (every
//           non-zero of row i of M^-1 has magnitude mag_i: 1 for the rest
//           direction, 1/2 face, 1/4 edge);
//   encode  code_k = floor((m'_k - lo'_k) sc_k + r_k), clamped to [0, 65535],
//           sc_k = 1 / step_k, r_k in [0, 1) from a hash of cell, step and k
//           (stochastic rounding: the expected code is the exact position).
// Encode and decode therefore share one grid to FP32 rounding: with a
// rounded 1 / step or rounded offsets every re-encode shifted the moments by
// the same fraction of a code, and with nearest rounding a moment that stays
// near one value rounds the same way each step; either kept adding to the
// kinetic energy of the 384^3 decaying turbulence (1.1e-3 after 4000 steps,
// against 1.6e-4 with this grid).
struct MeQuant {
    float step[19], L[19], Ll[19], lo[19], sc[19];
};

// Threads per block of the population kernels: one row segment of 256 cells
// (4 or 8 bytes) or 512 cells (2 bytes) per direction and block.
template <class Tag>
constexpr int step_block() {
    return sizeof(typename Codec<Tag>::raw) == 2 ? 512 : 256;
}

// Threads per block of the fused step (k_step).  FP32 steps two rows of 256
// cells per block: 512 threads, 2 blocks per CU (below), ran 40.2-40.6
// against 39.97 GLUPS for 256 threads at 4 blocks per CU, the same 16
// wavefronts per CU (256^3 Taylor-Green; reference y-walls 40.8-40.9 against
// 40.0-40.2, lid-driven cavity 40.2 against 40.3).  512 threads uncapped
// ran 39.4, at 3 blocks 39.2.
template <class Tag>
constexpr int fused_block() {
    return std::is_same<Tag, TF32>::value ? 512 : step_block<Tag>();
}

// Wavefronts per SIMD the fused step is compiled for at least (1: no bound).
// The moment-coded step (60-80 VGPRs) is held to four: one of its
// instantiations was once allocated 142-166 (three wavefronts per SIMD).
template <class Tag>
constexpr int fused_min_waves() {
    return std::is_same<Tag, TME16>::value ? 4 : 1;
}

// Blocks of the fused step (fused_block threads) resident per CU, 0 = as
// many as registers allow.  The step streams 19 row segments in and 19 out
// per row; for some storages fewer concurrent blocks per CU keep HBM busier
// (256^3 on MI355X, GLUPS uncapped -> capped): FP64 18.65 -> 19.63 at 2
// blocks of 256 (uncapped: 5, by registers); FP32 lid-driven cavity 39.24 ->
// 40.31 at 4 blocks of 256, periodic unchanged (39.9), then 2 blocks of 512
// (fused_block); FP16 80.5 -> 82.9 and BF16 82.4 -> 83.5 at 3 blocks of 512
// (uncapped: 4).  Stochastically rounded BF16, with more ALU work per cell,
// runs best uncapped (82.9 against 81.1 at 3 blocks).
template <class Tag>
constexpr uint32_t step_blocks_per_cu() {
    if constexpr (std::is_same<Tag, TBF16S>::value) {
        return 0;
    } else {
        constexpr size_t es = sizeof(typename Codec<Tag>::raw);
        return es == 8 ? 2 : (es == 4 ? 2 : 3);
    }
}

// Plane chunks of the fused step (step_chunks in d3q19_launch.hpp): the rate
// a storage steps at (GLUPS, to estimate a step's duration), the duration
// from which a step is chunked, and the work per chunk (us).  Measured on
// MI355X (GLUPS, one launch -> chunks of about 210 us): FP32 256^3 40.7 ->
// 41.3 (2 chunks), 512^3 39.2 -> 41.1 (16); FP64 256^3 19.6 -> 20.4 (4);
// FP16 512^3 75.8 -> 78.5 (8).  Shorter launches lose more to their ramp-up
// and tail than they gain: FP16 256^3 80.3 -> 77.9 in 2 chunks, 384^3
// 85.9 -> 85.3 in 2, FP32 128^3 40.0 -> 37.6 in 2; so 16-bit storage
// chunks only steps of a millisecond or more.  The moment-coded step, bound
// by issue rather than by HBM, runs in one launch (256^3: 47.0 -> 46.5 in 2).
template <class Tag>
constexpr double step_chunk_glups() {
    if constexpr (std::is_same<Tag, TME16>::value) {
        return 0.0;
    } else {
        constexpr size_t es = sizeof(typename Codec<Tag>::raw);
        return es == 8 ? 20.0 : (es == 4 ? 41.0 : 82.0);
    }
}
template <class Tag>
constexpr double step_chunk_min_us() {
    return sizeof(typename Codec<Tag>::raw) == 2 ? 1000.0 : 300.0;
}
constexpr double STEP_CHUNK_US = 210.0;

// ---------------------------------------------------------------------------
// Population layout ("row-blocked"): the lattice is a sequence of z-rows
// r = x * ny + y; row r holds the 19 direction segments of nzp elements each
// (nzp = nz rounded up to 64), one after the other:
//     byte offset of (x, y, z, i) = r * row_bytes + i * dir_bytes + z * es,
//     dir_bytes = nzp * es, row_bytes = 19 * dir_bytes.
// The 19 read and 19 write streams of a block then fall in a few contiguous
// 19 * nzp * es regions instead of 38 regions a plane apart (measured on
// MI355X: 6.26 TB/s vs 5.77 TB/s for the same 19-in/19-out stream pattern).
// Kernels run on a (z, y, x) grid whose blocks span whole wavefronts of one
// row, so row addresses are wavefront-uniform 64-bit scalars and per-lane
// offsets are 32-bit.
// ---------------------------------------------------------------------------
template <class C> struct Params {
    uint32_t nx, ny, nz, ncell;
    uint32_t nzp, dir_bytes;  // population layout
    uint64_t row_bytes;
    C omega;  // 1 / tau
    C pref;   // 1 - 1 / (2 tau)   (Guo prefactor, reference/bgk.py:90)
    int32_t force_mode;
    uint32_t wall_mask;
    C fu[3];                // uniform force
    const C* force;         // (3, ncell) field or null
    const C* brink;         // (ncell) Brinkman solid fraction or null
    const uint8_t* solid;   // (ncell) solid flag or null
    C face_u[6][3];         // half-way wall velocities
    C wtop[3], wbot[3];     // reference y-wall velocities
    int32_t local;          // state holds pre-collision f^n in place (not yet streamed)
    uint32_t seed;          // per-step key of the BF16 stochastic rounding
    // collision coefficients per weight class k (rest, face, edge), each the
    // correctly rounded value of its exact product: om w_k, 3 om w_k,
    // 4.5 om w_k, 3 pref w_k, 9 pref w_k
    C k_base[3], k_odd[3], k_quad[3], k_fodd[3], k_fquad[3];
    // half-way z walls: the moving-wall term 6 w_i (c_i . u_face) of each
    // direction bounced at the z = 0 face (c_z > 0, [0]) and the z = nz - 1
    // face (c_z < 0, [1]), formed as the general path forms it
    C zterm[2][19];
    const MeQuant* me;  // moment-encoded storage: its quantization constants, else null
    uint32_t x0;        // row grid: the x plane of blockIdx.z = 0 (a fused step's plane chunk)
    // Moment codes in one ring of planes (me_row_in / me_row_out, me_ring_sync):
    uint32_t ring_np;        // planes in the ring, 0: the plain layout
    uint32_t ring_in;        // ring position of plane 0 of the state read
    uint32_t ring_out;       // ring position of plane 0 of the state written
    uint32_t ring_gap;       // step: the state written lies ring_gap planes before the one read
    uint32_t ring_target;    // step: the count ring_done[x] reaches once plane x is read
    const char* ring_saved;  // step: plane 0 of the state read, copied before the step
    uint32_t* ring_done;     // step: per plane, wavefronts whose pulls are consumed (monotonic)
    uint32_t* ring_err;      // step: set to 1 when a wait gave up
    uint32_t ring_spins;     // step: polls of a short count before a wait gives up
    // The adjoint's recomputation (k_step / k_collide_local with MOM): rho - 1
    // and the momentum sum m of the pulled f^n, (ncell, 4) in the compute type.
    C* mom;
};

// Fill the z-wall terms of p from its face velocities (host).
template <class C>
inline void set_zwall_terms(Params<C>& p) {
    for (int i = 0; i < 19; ++i) {
        p.zterm[0][i] = CZ[i] > 0 ? C(6) * wgt<C>(i) * cdot<C>(i, p.face_u[4][0], p.face_u[4][1], p.face_u[4][2]) : C(0);
        p.zterm[1][i] = CZ[i] < 0 ? C(6) * wgt<C>(i) * cdot<C>(i, p.face_u[5][0], p.face_u[5][1], p.face_u[5][2]) : C(0);
    }
}

// Fill the collision coefficients of p for relaxation time tau (host).
template <class C>
inline void set_collision_constants(Params<C>& p, double tau) {
    const double om = 1.0 / tau, pr = 1.0 - 0.5 / tau;
    p.omega = (C)om;
    p.pref = (C)pr;
    const double w[3] = {1.0 / 3.0, 1.0 / 18.0, 1.0 / 36.0};
    for (int k = 0; k < 3; ++k) {
        p.k_base[k] = (C)(om * w[k]);
        p.k_odd[k] = (C)(3.0 * om * w[k]);
        p.k_quad[k] = (C)(4.5 * om * w[k]);
        p.k_fodd[k] = (C)(3.0 * pr * w[k]);
        p.k_fquad[k] = (C)(9.0 * pr * w[k]);
    }
}

// Threads of a block along a row: the widest of 256, 128, 64 that leaves the
// fewest idle lanes on a row of nz cells (384 -> 128 x 3, 320 -> 64 x 5).
__host__ __device__ inline uint32_t row_block(uint32_t nz) {
    uint32_t best = 64, waste = ~0u;
    for (uint32_t b = 256; b >= 64; b /= 2) {
        const uint32_t w = (nz + b - 1) / b * b - nz;
        if (w < waste) {
            waste = w;
            best = b;
        }
    }
    return best;
}

// Padded row length for nz.
__host__ __device__ inline uint32_t padded_nz(uint32_t nz) { return (nz + 63u) / 64u * 64u; }

// Cell coordinates of the calling thread on the (z, y, x) grid; false for
// padding threads.  Every wavefront lies in one row (blockDim.x is a multiple
// of 64), so y is made an SGPR value.  x0: the plane of blockIdx.z = 0 (a
// launch over a chunk of planes).
__device__ __forceinline__ bool grid_cell(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t& x,
                                          uint32_t& y, uint32_t& z, uint32_t x0 = 0) {
    z = blockIdx.x * blockDim.x + threadIdx.x;
    y = __builtin_amdgcn_readfirstlane(blockIdx.y * blockDim.y + threadIdx.y);
    x = blockIdx.z + x0;
    return z < nz && y < ny && x < nx;
}

// Cell of the calling thread for a population kernel, false for padding
// threads.  ROWS: the (z, y, x) row grid above.  Otherwise one thread per cell
// in the reference's C order on a linear grid -- the layout of lattices whose
// rows are shorter than a wavefront (nz < 64, e.g. the reference's depth-3
// slabs), stored direction-major as the reference stores f:
// row_bytes = nz * es and dir_bytes = ncell * es in the offset formula above.
template <bool ROWS>
__device__ __forceinline__ bool cell_of(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t ncell, uint32_t& x,
                                        uint32_t& y, uint32_t& z, uint32_t x0 = 0) {
    if constexpr (ROWS) {
        return grid_cell(nx, ny, nz, x, y, z, x0);
    } else {
        const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= ncell) return false;
        z = c % nz;
        const uint32_t t = c / nz;
        y = t % ny;
        x = t / ny;
        return true;
    }
}

__device__ __forceinline__ uint32_t block_linear_id() {
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}

// Blocks of the grid.  Per-block partials of a reduction are stored
// component-major, [component][block], so that the final reduction of a
// component reads them contiguously.
__device__ __forceinline__ uint32_t block_count() { return gridDim.x * gridDim.y * gridDim.z; }

template <class C>
__device__ __forceinline__ const char* row_of(const void* base, const Params<C>& p, uint32_t x, uint32_t y) {
    return (const char*)base + (uint64_t)(x * p.ny + y) * p.row_bytes;
}

// Moment codes in a ring of ring_np planes (one buffer).  A step reads the
// state at ring position ring_in and writes the next one ring_gap planes
// before it, in place: plane x of the new state overwrites plane x - gap of
// the old one once the wavefronts that pull from that plane have consumed
// their codes (me_ring_sync), and plane 0 of the old state, which the last
// planes pull from after its slot is reused, is read from a copy made
// before the step (ring_saved).  Outside a step ring_in = ring_out.
// Plane x of the state read (its row y at + y * row_bytes).
template <class C>
__device__ __forceinline__ const char* me_plane_in(const void* base, const Params<C>& p, uint32_t x) {
    if (p.ring_np == 0) return (const char*)base + (uint64_t)(x * p.ny) * p.row_bytes;
    if (p.ring_saved && x == 0) return p.ring_saved;
    uint32_t pos = x + p.ring_in;
    if (pos >= p.ring_np) pos -= p.ring_np;
    return (const char*)base + (uint64_t)(pos * p.ny) * p.row_bytes;
}
template <class C>
__device__ __forceinline__ const char* me_row_in(const void* base, const Params<C>& p, uint32_t x, uint32_t y) {
    return me_plane_in(base, p, x) + (uint64_t)y * p.row_bytes;
}
template <class C>
__device__ __forceinline__ char* me_row_out(void* base, const Params<C>& p, uint32_t x, uint32_t y) {
    if (p.ring_np == 0) return (char*)row_of(base, p, x, y);
    uint32_t pos = x + p.ring_out;
    if (pos >= p.ring_np) pos -= p.ring_np;
    return (char*)base + (uint64_t)(pos * p.ny + y) * p.row_bytes;
}

// U: the address of a direction segment of a row, wavefront-uniform (the row
// grid: every wavefront lies in one row), passed through readfirstlane.  The
// compiler cannot reassociate the lane's 32-bit offset into the opaque
// uniform base, so each access takes the SGPR-base + 32-bit VGPR-offset form.
// Without it LLVM shares (row + z s) over the 19 stores of a cell and forms
// one 64-bit add per store (v_lshl_add_u64: 18 VALU per cell), and the
// half-way kernel does the same for nine of its loads.  Never for the linear
// grid of short-row lattices, where every lane has its own row; the kernels
// choose it per storage (k_step).
template <bool U>
__device__ __forceinline__ const char* seg_base(const char* p) {
    if constexpr (U) {
        const uint64_t v = (uint64_t)p;
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
        // rebuilt as a global-memory pointer, so the accesses stay global_*
        typedef __attribute__((address_space(1))) const char gchar;
        return (const char*)(gchar*)(((uint64_t)hi << 32) | lo);
    } else {
        return p;
    }
}

// Population i of a row at byte offset zoff; NT: with the nontemporal hint.
// In the fused step the hint depends on the
// direction.  A z-shifted pull (c_z != 0) straddles a line that the
// neighbouring wavefront of the same row also reads: loaded plainly, it stays
// in L2 (with the hint it was re-fetched, 12% more FP32 reads measured with
// FETCH_SIZE at 256^3).  A pull with c_z = 0 reads whole lines that no other
// wavefront touches: the hint keeps them from displacing the shared ones
// (256^3 step: FP32 39.9 against 38.9 GLUPS, FP16 81.1 against 80.0, FP64
// 18.4 against 17.9).
// The adjoint hints all its 4-byte loads: its backward pass measured 4%
// faster with them.
template <class Tag, bool NT = false>
__device__ __forceinline__ typename Codec<Tag>::C ld_at(const char* addr, int i) {
    typedef typename Codec<Tag>::raw R;
    const R* a = (const R*)addr;
    if constexpr (NT)
        return Codec<Tag>::dec(__builtin_nontemporal_load(a), i);
    else
        return Codec<Tag>::dec(*a, i);
}

template <class Tag, bool NT = false, bool U = false>
__device__ __forceinline__ typename Codec<Tag>::C ld(const char* row, int i, uint32_t dir_bytes, uint32_t zoff) {
    return ld_at<Tag, NT>(seg_base<U>(row + (uint32_t)i * dir_bytes) + zoff, i);
}

// Stores are nontemporal: each population is written once per step and read
// in the next (plain FP32 stores measured 12% slower at 256^3).
template <class Tag, bool U = false>
__device__ __forceinline__ void st(char* row, int i, uint32_t dir_bytes, uint32_t zoff, typename Codec<Tag>::C v) {
    typedef typename Codec<Tag>::raw R;
    __builtin_nontemporal_store(Codec<Tag>::enc(v, i), (R*)(seg_base<U>(row + (uint32_t)i * dir_bytes) + zoff));
}

// ---------------------------------------------------------------------------
// Moment codes (storage TME16).  The moments are the reference variant's
// (packages/lattice-boltzmann-d3q19-frontier-moment-encoded/src/lbm_me.cpp:
// 83-90): monomials of c over the 19 velocities.
// ---------------------------------------------------------------------------

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

// |Minv[i][k]| of the non-zeros of row i.
__host__ __device__ constexpr float me_row_mag(int i) { return i == 0 ? 1.0f : (i <= 6 ? 0.5f : 0.25f); }

// The quantization constants read through the constant address space:
// invariant for the kernel, so the loads are scalar (one per wavefront).
// Through the generic pointer the compiler could not rule out that the
// step's stores alias them and loaded them per lane (20 dwordx4 vector loads
// per thread; 256^3 step 31 against 47 GLUPS).
typedef __attribute__((address_space(4))) const MeQuant MeQuantC;
__device__ __forceinline__ MeQuantC* me_const(const MeQuant* q) { return (MeQuantC*)q; }

// Deviation d_i of the cell whose slots start at `row` + zoff, from its codes
// (the 3-5 codes of row i of M^-1; 91 over all 19 directions).  A pull
// decodes each direction from the codes of its own source cell.
template <bool U>
__device__ __forceinline__ float me_pop(const char* row, int i, uint32_t DB, uint32_t zoff, const MeQuant* qg) {
#pragma clang fp contract(off)
    MeQuantC* q = me_const(qg);
    float s = 0.0f;
    bool first = true;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int a = ME_MINV4[i][k];
        if (a != 0) {
            const float c = (float)*(const uint16_t*)(seg_base<U>(row + (uint32_t)k * DB) + zoff);
            const float st = a > 0 ? q->step[k] : -q->step[k];
            if (first) {
                s = c * st;
                first = false;
            } else {
                s = fmaf(c, st, s);
            }
        }
    }
    return fmaf(s, me_row_mag(i), q->L[i]) + q->Ll[i];
}

// Even-degree monomials: the same value for c and -c.
__host__ __device__ constexpr bool me_even(int k) { return k == 0 || (k >= 4 && k <= 9) || k >= 16; }

// Encode the deviations d of a cell and store its 19 codes.  The moments are
// summed over the opposite pairs (i, i + 1 = opp(i), i odd): the pair
// contributes mono_k(c_i) (d_i + d_opp) to an even moment and
// mono_k(c_i) (d_i - d_opp) to an odd one.  key: the cell's rounding hash
// (store_local); r_k is the top 16 bits of key * (odd constant of k).
template <bool U>
__device__ __forceinline__ void me_store_cell(char* row, uint32_t DB, uint32_t zoff, const float d[Q],
                                              const MeQuant* qg, uint32_t key) {
#pragma clang fp contract(off)
    MeQuantC* q = me_const(qg);
    float S[9], A[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        S[j] = d[2 * j + 1] + d[2 * j + 2];
        A[j] = d[2 * j + 1] - d[2 * j + 2];
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        float m = k == 0 ? d[0] : 0.0f;
        bool first = k != 0;
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const int c = me_mono(k, 2 * j + 1);
            if (c != 0) {
                const float t = me_even(k) ? S[j] : A[j];
                if (first) {
                    m = c > 0 ? t : -t;
                    first = false;
                } else {
                    m = c > 0 ? m + t : m - t;
                }
            }
        }
        // stochastic rounding; NaN and values below the range give 0
        const float r = (float)((key * (0x9e3779b1u * (uint32_t)(2 * k + 1))) >> 16) * (1.0f / 65536.0f);
        const float v = fminf(fmaxf(fmaf(m - q->lo[k], q->sc[k], r), 0.0f), 65535.0f);
        __builtin_nontemporal_store((uint16_t)(uint32_t)v, (uint16_t*)(seg_base<U>(row + (uint32_t)k * DB) + zoff));
    }
}

// In-place (unstreamed) populations of cell (x, y, z).
template <class Tag, bool U = false>
__device__ __forceinline__ void load_local(const void* in, const Params<typename Codec<Tag>::C>& p, uint32_t x,
                                           uint32_t y, uint32_t z, typename Codec<Tag>::C d[Q]) {
    constexpr uint32_t ES = sizeof(typename Codec<Tag>::raw);
    if constexpr (std::is_same<Tag, TME16>::value) {
        const char* row = me_row_in(in, p, x, y);
#pragma unroll
        for (int i = 0; i < Q; ++i) d[i] = me_pop<U>(row, i, p.dir_bytes, z * ES, p.me);
    } else {
        const char* row = row_of(in, p, x, y);
#pragma unroll
        for (int i = 0; i < Q; ++i) d[i] = ld<Tag, false, U>(row, i, p.dir_bytes, z * ES);
    }
}

// 32-bit integer hash (two xorshift-multiply rounds).
__device__ __forceinline__ uint32_t mix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

// The raw value store_local writes into slot i of cell (x, y, z).  For
// stochastically rounded BF16 the rounding bits come from a hash of (cell,
// step) and one multiply per population, so the value is a function of the
// cell, the step and d alone.
template <class Tag>
__device__ __forceinline__ typename Codec<Tag>::raw raw_of(const Params<typename Codec<Tag>::C>& p, uint32_t x,
                                                          uint32_t y, uint32_t z, typename Codec<Tag>::C v, int i) {
    if constexpr (std::is_same<Tag, TBF16S>::value) {
        // BF16 keeps 8 mantissa bits: nearest rounding of the small deviations
        // biases them the same way every step (kinetic energy of decaying
        // turbulence off by 1.4e-2 at 384^3); stochastic rounding removes the
        // bias.  key * odd constant (mod 2^32) is a bijection of the hashed
        // key, so its top 16 bits are uniform.
        const uint32_t key = mix32(((x * p.ny + y) * p.nz + z) ^ p.seed);
        const uint32_t r = (key * (0x9e3779b1u * (uint32_t)(2 * i + 1))) >> 16;
        return Codec<TBF16>::enc_sr(v, r);
    } else {
        return Codec<Tag>::enc(v, i);
    }
}

template <class Tag, bool U = false>
__device__ __forceinline__ void store_local(void* out, const Params<typename Codec<Tag>::C>& p, uint32_t x,
                                            uint32_t y, uint32_t z, const typename Codec<Tag>::C d[Q]) {
    typedef typename Codec<Tag>::raw R;
    constexpr uint32_t ES = sizeof(R);
    char* row = std::is_same<Tag, TME16>::value ? me_row_out(out, p, x, y) : (char*)row_of(out, p, x, y);
    uint32_t zoff = z * ES;
    // an opaque 32-bit lane offset: with the forced collision in the same
    // kernel LLVM otherwise widened it to 64 bits and added it per store
    if constexpr (U) asm volatile("" : "+v"(zoff));
    if constexpr (std::is_same<Tag, TME16>::value) {
        me_store_cell<U>(row, p.dir_bytes, zoff, d, p.me, mix32(((x * p.ny + y) * p.nz + z) ^ p.seed));
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i)
            __builtin_nontemporal_store(raw_of<Tag>(p, x, y, z, d[i], i),
                                        (R*)(seg_base<U>(row + (uint32_t)i * p.dir_bytes) + zoff));
    }
}

// Reference y-wall treatment (reference/bgk.py:190-215) applied in registers
// to the periodically streamed deviations of a cell in row y: at y = 0 the
// populations with c_y > 0 (at y = ny - 1 those with c_y < 0) are replaced by
// their opposites minus 6 w_i rho_wall (c_i . u_wall), rho_wall being the
// density of all 19 streamed values before the swap.
template <class C>
__device__ __forceinline__ void ywall_swap(const Params<C>& p, uint32_t y, C d[Q]) {
#pragma clang fp contract(off)
    const bool bot = (y == 0), top = (y + 1 == p.ny);
    if (bot || top) {
        C rw = C(0);
#pragma unroll
        for (int i = 0; i < Q; ++i) rw += d[i];
        rw += C(1);
        C nd[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) nd[i] = d[i];
        if (bot) {
#pragma unroll
            for (int i = 0; i < Q; ++i)
                if (CY[i] > 0) {
                    const C cu = cdot<C>(i, p.wbot[0], p.wbot[1], p.wbot[2]);
                    nd[i] = d[OPP[i]] - C(6) * wgt<C>(i) * rw * cu;
                }
        }
        if (top) {
#pragma unroll
            for (int i = 0; i < Q; ++i)
                if (CY[i] < 0) {
                    const C cu = cdot<C>(i, p.wtop[0], p.wtop[1], p.wtop[2]);
                    nd[i] = d[OPP[i]] - C(6) * wgt<C>(i) * rw * cu;
                }
        }
#pragma unroll
        for (int i = 0; i < Q; ++i) d[i] = nd[i];
    }
}

// Half-way mode: the wall face (D3Q19_FACE_* index) that the pull source
// x - c_i of a cell lies across, or -1.  An edge or corner direction at an
// edge or corner of the box crosses two or three wall faces; the last one in
// face order counts (its velocity enters the moving-wall term).
template <class C>
__device__ __forceinline__ int halfway_face(int i, uint32_t wm, uint32_t x, uint32_t y, uint32_t z,
                                            const Params<C>& p) {
    int face = -1;
    if (CX[i] > 0 && x == 0 && (wm & 1u)) face = 0;
    if (CX[i] < 0 && x + 1 == p.nx && (wm & 2u)) face = 1;
    if (CY[i] > 0 && y == 0 && (wm & 4u)) face = 2;
    if (CY[i] < 0 && y + 1 == p.ny && (wm & 8u)) face = 3;
    if (CZ[i] > 0 && z == 0 && (wm & 16u)) face = 4;
    if (CZ[i] < 0 && z + 1 == p.nz && (wm & 32u)) face = 5;
    return face;
}

// Periodic pull of all 19 directions (the load hints as described at ld).
template <class Tag, bool NT, bool U = false>
__device__ __forceinline__ void pull_periodic(const void* in, const Params<typename Codec<Tag>::C>& p, uint32_t z,
                                              const uint32_t xs[3], const uint32_t ys[3], const uint32_t zs[3],
                                              typename Codec<Tag>::C d[Q]) {
    constexpr uint32_t ES = sizeof(typename Codec<Tag>::raw);
    const uint32_t DB = p.dir_bytes;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const char* row = row_of(in, p, xs[1 - CX[i]], ys[1 - CY[i]]);
        if (CZ[i] == 0)
            d[i] = ld<Tag, true, U>(row, i, DB, z * ES);
        else
            d[i] = ld<Tag, NT, U>(row, i, DB, zs[1 - CZ[i]] * ES);
    }
}

// ---------------------------------------------------------------------------
// Pull streaming with boundary handling.  Produces the pre-collision
// deviations d_i(x) = f_i(x) - w_i of the reference state after streaming and
// bounce-back, from the stored post-collision deviations.  Direction i pulls
// from row (x - cx, y - cy) at z - cz.
// ---------------------------------------------------------------------------

// Source rows and z positions of a cell's pull: index 0/1/2 along an axis for
// velocity component +1/0/-1, wrapped periodically.
#define D3Q19_PULL_SOURCES(p, x, y, z)                                                   \
    const uint32_t xs[3] = {x == 0 ? p.nx - 1 : x - 1, x, x + 1 == p.nx ? 0 : x + 1};     \
    const uint32_t ys[3] = {y == 0 ? p.ny - 1 : y - 1, y, y + 1 == p.ny ? 0 : y + 1};     \
    const uint32_t zs[3] = {z == 0 ? p.nz - 1 : z - 1, z, z + 1 == p.nz ? 0 : z + 1}

// Half-way mode.  A lattice without solid cells splits its rows in two:
// rows on an x or y wall face (edge rows) and the others (interior rows).
// Whether a direction crosses an x or y face is the same for every lane of a
// row, and a z face is met only by the lanes at z = 0 and z = nz - 1, so
// neither kind needs a per-lane choice of address for the x and y faces.
// Lattices with solids take the general path (per-direction selects).
template <class C>
__device__ __forceinline__ bool halfway_edge_row(const Params<C>& p, uint32_t x, uint32_t y) {
    const uint32_t wm = p.wall_mask;
    return ((wm & 1u) && x == 0) || ((wm & 2u) && x + 1 == p.nx) || ((wm & 4u) && y == 0) ||
           ((wm & 8u) && y + 1 == p.ny);
}

// z-wall bounce values of 16-bit storage, loaded raw.  A lane at z = 0
// bounces, for each i with c_z > 0, its own opp(i); a lane at z = nz - 1, for
// each i with c_z < 0, its own opp(i).  The address is the same in every wall
// lane of a wavefront (one row, fixed z), so on the row grid it is a scalar
// base.  Each value is held zero-extended in a 32-bit register of its own:
// held as 16-bit values, the compiler packed two per register right after
// the loads, and the packing waited for them before the pull was issued (a
// full memory round trip more for every wavefront holding a wall lane).
template <class Tag, bool U>
__device__ __forceinline__ void zwall_fetch(const char* self, const Params<typename Codec<Tag>::C>& p, bool zlo,
                                            bool zhi, uint32_t rlo[Q], uint32_t rhi[Q]) {
    typedef typename Codec<Tag>::raw R;
    const uint32_t DB = p.dir_bytes;
    if (zlo) {
#pragma unroll
        for (int i = 0; i < Q; ++i)
            if (CZ[i] > 0) rlo[i] = *(const R*)seg_base<U>(self + (uint32_t)OPP[i] * DB);
    }
    if (zhi) {
        const uint32_t zo = (p.nz - 1) * (uint32_t)sizeof(R);
#pragma unroll
        for (int i = 0; i < Q; ++i)
            if (CZ[i] < 0) rhi[i] = *(const R*)(seg_base<U>(self + (uint32_t)OPP[i] * DB) + zo);
    }
}

// Pin the fetched z-wall values in their registers once the pull is issued:
// the decode of a value loaded in a branch was otherwise folded into that
// branch (next to the load), where it waited for the load before the pull.
__device__ __forceinline__ void zwall_pin(uint32_t rlo[Q], uint32_t rhi[Q]) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        if (CZ[i] > 0) asm volatile("" : "+v"(rlo[i]));
        if (CZ[i] < 0) asm volatile("" : "+v"(rhi[i]));
    }
}

// The z-wall lanes' bounced directions: opp(i) decodes with index i (same
// weight class), plus the face's precomputed moving-wall term.  Applied
// after any x/y bounce: at an edge the z face is the one that counts.
template <class Tag>
__device__ __forceinline__ void zwall_apply(const Params<typename Codec<Tag>::C>& p, bool zlo, bool zhi,
                                            const uint32_t rlo[Q], const uint32_t rhi[Q],
                                            typename Codec<Tag>::C d[Q]) {
#pragma clang fp contract(off)
    typedef typename Codec<Tag>::raw R;
    if (zlo) {
#pragma unroll
        for (int i = 0; i < Q; ++i)
            if (CZ[i] > 0) d[i] = Codec<Tag>::dec((R)rlo[i], i) + p.zterm[0][i];
    }
    if (zhi) {
#pragma unroll
        for (int i = 0; i < Q; ++i)
            if (CZ[i] < 0) d[i] = Codec<Tag>::dec((R)rhi[i], i) + p.zterm[1][i];
    }
}

// Interior rows: the periodic pull plus the z-wall bounce.  For 16-bit
// storage the wall lanes load their bounce values before the pull (loads
// complete in issue order, so the pull's wait covers them); for 4- and
// 8-byte storage after it, in one block for both faces (the earlier order
// ran FP32 1.7% slower at 256^3; two blocks, as for 16-bit storage, made
// the latency-bound 64^3 cavity 11% slower).
template <class Tag, bool NT = false, bool U = false>
__device__ __forceinline__ void pull_halfway_fast(const void* in, const Params<typename Codec<Tag>::C>& p,
                                                  uint32_t x, uint32_t y, uint32_t z,
                                                  typename Codec<Tag>::C d[Q]) {
#pragma clang fp contract(off)
    typedef typename Codec<Tag>::raw R;
    constexpr uint32_t ES = sizeof(R);
    D3Q19_PULL_SOURCES(p, x, y, z);
    const uint32_t wm = p.wall_mask;
    const char* self = row_of(in, p, x, y);
    const bool zlo = z == 0 && (wm & 16u), zhi = z + 1 == p.nz && (wm & 32u);
    if constexpr (sizeof(R) == 2) {
        uint32_t rlo[Q], rhi[Q];
        zwall_fetch<Tag, U>(self, p, zlo, zhi, rlo, rhi);
        pull_periodic<Tag, NT, U>(in, p, z, xs, ys, zs, d);
        zwall_pin(rlo, rhi);
        zwall_apply<Tag>(p, zlo, zhi, rlo, rhi, d);
    } else {
        pull_periodic<Tag, NT, U>(in, p, z, xs, ys, zs, d);
        if (zlo || zhi) {
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                if (CZ[i] == 0 || !(CZ[i] > 0 ? zlo : zhi)) continue;
                const typename Codec<Tag>::C v = ld_at<Tag, NT>(self + (uint32_t)OPP[i] * p.dir_bytes + z * ES, i);
                d[i] = v + p.zterm[CZ[i] > 0 ? 0 : 1][i];
            }
        }
    }
}

// Edge rows of 16-bit lattices: per direction, a wavefront-uniform choice
// between the pull source and the cell's own opposite population (bounce
// from an x or y face, moving-wall term of the face that counts: y over x),
// then the z-wall bounce as for interior rows.  (4- and 8-byte storage
// takes the general path on edge rows: the same at 256^3, and shorter at
// 64^3, where every row is one wavefront.)
template <class Tag, bool NT = false, bool U = false>
__device__ __forceinline__ void pull_halfway_edge(const void* in, const Params<typename Codec<Tag>::C>& p,
                                                  uint32_t x, uint32_t y, uint32_t z,
                                                  typename Codec<Tag>::C d[Q]) {
#pragma clang fp contract(off)
    typedef typename Codec<Tag>::C C;
    typedef typename Codec<Tag>::raw R;
    constexpr uint32_t ES = sizeof(R);
    D3Q19_PULL_SOURCES(p, x, y, z);
    const uint32_t DB = p.dir_bytes;
    const uint32_t wm = p.wall_mask;
    const char* self = row_of(in, p, x, y);
    const bool zlo = z == 0 && (wm & 16u), zhi = z + 1 == p.nz && (wm & 32u);
    uint32_t rlo[Q], rhi[Q];
    zwall_fetch<Tag, U>(self, p, zlo, zhi, rlo, rhi);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const bool wx = (CX[i] > 0 && x == 0 && (wm & 1u)) || (CX[i] < 0 && x + 1 == p.nx && (wm & 2u));
        const bool wy = (CY[i] > 0 && y == 0 && (wm & 4u)) || (CY[i] < 0 && y + 1 == p.ny && (wm & 8u));
        const bool cross = wx || wy;
        const char* base = cross ? self + (uint32_t)OPP[i] * DB
                                 : row_of(in, p, xs[1 - CX[i]], ys[1 - CY[i]]) + (uint32_t)i * DB;
        const uint32_t off = (CZ[i] == 0 || cross) ? z * ES : zs[1 - CZ[i]] * ES;
        const C v = CZ[i] == 0 ? ld_at<Tag, true>(seg_base<U>(base) + off, i)
                               : ld_at<Tag, NT>(seg_base<U>(base) + off, i);
        if (cross) {
            const int f = wy ? (CY[i] > 0 ? 2 : 3) : (CX[i] > 0 ? 0 : 1);
            d[i] = v + C(6) * wgt<C>(i) * cdot<C>(i, p.face_u[f][0], p.face_u[f][1], p.face_u[f][2]);
        } else {
            d[i] = v;
        }
    }
    zwall_pin(rlo, rhi);
    zwall_apply<Tag>(p, zlo, zhi, rlo, rhi, d);
}

// Half-way general pull (lattices with solid cells): per-direction selects
// between the pull source and the cell's own opposite population.
template <class Tag, bool NT = false>
__device__ __forceinline__ void pull_halfway_general(const void* in, const Params<typename Codec<Tag>::C>& p,
                                                     uint32_t x, uint32_t y, uint32_t z,
                                                     typename Codec<Tag>::C d[Q]) {
#pragma clang fp contract(off)
    typedef typename Codec<Tag>::C C;
    constexpr uint32_t ES = sizeof(typename Codec<Tag>::raw);
    D3Q19_PULL_SOURCES(p, x, y, z);
    const uint32_t DB = p.dir_bytes;
    const uint32_t wm = p.wall_mask;
    const char* self = row_of(in, p, x, y);
    // Solid source cells: all flags are loaded before any population, so
    // that no population load waits on a flag (none when there are no
    // solids: the branch is uniform).
    bool solid_src[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) solid_src[i] = false;
    if (p.solid) {
#pragma unroll
        for (int i = 1; i < Q; ++i)
            solid_src[i] = p.solid[(xs[1 - CX[i]] * p.ny + ys[1 - CY[i]]) * p.nz + zs[1 - CZ[i]]] != 0;
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        // The pull crosses a wall face; the face that counts for the
        // moving-wall term is the last in face order (z over y over x,
        // as halfway_face), with compile-time face indices.
        const bool wx = (CX[i] > 0 && x == 0 && (wm & 1u)) || (CX[i] < 0 && x + 1 == p.nx && (wm & 2u));
        const bool wy = (CY[i] > 0 && y == 0 && (wm & 4u)) || (CY[i] < 0 && y + 1 == p.ny && (wm & 8u));
        const bool wz = (CZ[i] > 0 && z == 0 && (wm & 16u)) || (CZ[i] < 0 && z + 1 == p.nz && (wm & 32u));
        const bool wall = wx || wy || wz;
        C u0 = C(0), u1 = C(0), u2 = C(0);
        if (CX[i] != 0) {
            const int f = CX[i] > 0 ? 0 : 1;
            u0 = wx ? p.face_u[f][0] : u0;
            u1 = wx ? p.face_u[f][1] : u1;
            u2 = wx ? p.face_u[f][2] : u2;
        }
        if (CY[i] != 0) {
            const int f = CY[i] > 0 ? 2 : 3;
            u0 = wy ? p.face_u[f][0] : u0;
            u1 = wy ? p.face_u[f][1] : u1;
            u2 = wy ? p.face_u[f][2] : u2;
        }
        if (CZ[i] != 0) {
            const int f = CZ[i] > 0 ? 4 : 5;
            u0 = wz ? p.face_u[f][0] : u0;
            u1 = wz ? p.face_u[f][1] : u1;
            u2 = wz ? p.face_u[f][2] : u2;
        }
        // Bounce-back (a wall face, or a solid source cell) reads the
        // opposite population of the cell itself, else the pull source.
        // One load per direction from a selected address: a load in one
        // arm of a branch would have to complete (and, for 16-bit
        // storage, be decoded) before the arms join, serialising the 19
        // loads.  opp(i) has the weight class of i, so the value decodes
        // with index i either way.
        const bool back = wall || solid_src[i];
        const char* src = back ? self + (uint32_t)OPP[i] * DB + z * ES
                               : row_of(in, p, xs[1 - CX[i]], ys[1 - CY[i]]) + (uint32_t)i * DB +
                                     zs[1 - CZ[i]] * ES;
        const C v = CZ[i] == 0 ? ld_at<Tag, true>(src, i) : ld_at<Tag, NT>(src, i);
        d[i] = wall ? v + C(6) * wgt<C>(i) * cdot<C>(i, u0, u1, u2) : v;
    }
}

// Pull of a moment-encoded lattice: each direction decoded from the codes of
// its source cell (me_pop); a bounce-back (half-way wall face or solid
// source) decodes the cell's own opposite direction instead, in a branch
// that is wavefront-uniform away from walls and solids.  The reference
// y-walls act on the pulled deviations (ywall_swap).
template <int BC, bool U>
__device__ __forceinline__ void me_pull(const void* in, const Params<float>& p, uint32_t x, uint32_t y, uint32_t z,
                                        float d[Q]) {
#pragma clang fp contract(off)
    D3Q19_PULL_SOURCES(p, x, y, z);
    const uint32_t DB = p.dir_bytes;
    uint32_t zo[3] = {zs[0] * 2u, zs[1] * 2u, zs[2] * 2u};
    if constexpr (U) asm volatile("" : "+v"(zo[0]), "+v"(zo[1]), "+v"(zo[2]));
    if constexpr (BC == BC_HALFWAY) {
        const uint32_t wm = p.wall_mask;
        const char* pl[3] = {me_plane_in(in, p, xs[0]), me_plane_in(in, p, xs[1]), me_plane_in(in, p, xs[2])};
        const uint64_t yo[3] = {(uint64_t)ys[0] * p.row_bytes, (uint64_t)ys[1] * p.row_bytes,
                                (uint64_t)ys[2] * p.row_bytes};
        const char* self = pl[1] + yo[1];
        bool solid_src[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) solid_src[i] = false;
        if (p.solid) {
#pragma unroll
            for (int i = 1; i < Q; ++i)
                solid_src[i] = p.solid[(xs[1 - CX[i]] * p.ny + ys[1 - CY[i]]) * p.nz + zs[1 - CZ[i]]] != 0;
        }
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int f = halfway_face<float>(i, wm, x, y, z, p);
            float v;
            if (f >= 0 || solid_src[i])
                v = me_pop<U>(self, OPP[i], DB, zo[1], p.me);
            else
                v = me_pop<U>(pl[1 - CX[i]] + yo[1 - CY[i]], i, DB, zo[1 - CZ[i]], p.me);
            if (f >= 0) v = v + 6.0f * wgt<float>(i) * cdot<float>(i, p.face_u[f][0], p.face_u[f][1], p.face_u[f][2]);
            d[i] = v;
        }
    } else {
        const char* pl[3] = {me_plane_in(in, p, xs[0]), me_plane_in(in, p, xs[1]), me_plane_in(in, p, xs[2])};
        const uint64_t yo[3] = {(uint64_t)ys[0] * p.row_bytes, (uint64_t)ys[1] * p.row_bytes,
                                (uint64_t)ys[2] * p.row_bytes};
#pragma unroll
        for (int i = 0; i < Q; ++i)
            d[i] = me_pop<U>(pl[1 - CX[i]] + yo[1 - CY[i]], i, DB, zo[1 - CZ[i]], p.me);
        // Every decoded value complete here: otherwise the forced periodic
        // step kept loads in flight across the force-mode branches of the
        // collision and spilled 23 VGPRs to scratch at 128.
#pragma unroll
        for (int i = 0; i < Q; ++i) asm volatile("" : "+v"(d[i]));
        if constexpr (BC == BC_YWALLS) ywall_swap<float>(p, y, d);
    }
}

template <class Tag, int BC, bool NT = false, bool U = false>
__device__ __forceinline__ void pull(const void* in, const Params<typename Codec<Tag>::C>& p, uint32_t x,
                                     uint32_t y, uint32_t z, typename Codec<Tag>::C d[Q]) {
    if constexpr (std::is_same<Tag, TME16>::value) {
        me_pull<BC, U>(in, p, x, y, z, d);
    } else if constexpr (BC == BC_HALFWAY) {
        constexpr bool short_raw = sizeof(typename Codec<Tag>::raw) == 2;
        if (p.solid || (!short_raw && halfway_edge_row(p, x, y)))
            pull_halfway_general<Tag, NT>(in, p, x, y, z, d);
        else if (halfway_edge_row(p, x, y))
            pull_halfway_edge<Tag, NT, U>(in, p, x, y, z, d);
        else
            pull_halfway_fast<Tag, NT, U>(in, p, x, y, z, d);
    } else {
        D3Q19_PULL_SOURCES(p, x, y, z);
        pull_periodic<Tag, NT, U>(in, p, z, xs, ys, zs, d);
        if constexpr (BC == BC_YWALLS) ywall_swap<typename Codec<Tag>::C>(p, y, d);
    }
}

// Moments of the deviations: drho = rho - 1, m = sum c_i f_i.
template <class C>
__device__ __forceinline__ void moments(const C d[Q], C& drho, C& mx, C& my, C& mz) {
#pragma clang fp contract(off)
    C r = C(0), a = C(0), b = C(0), c = C(0);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        r += d[i];
        if (CX[i] > 0) a += d[i];
        if (CX[i] < 0) a -= d[i];
        if (CY[i] > 0) b += d[i];
        if (CY[i] < 0) b -= d[i];
        if (CZ[i] > 0) c += d[i];
        if (CZ[i] < 0) c -= d[i];
    }
    drho = r;
    mx = a;
    my = b;
    mz = c;
}

// Body force at a cell (Guo), including the Brinkman drag -s * m.
template <class C>
__device__ __forceinline__ void cell_force(const Params<C>& p, uint32_t cell, C mx, C my, C mz,
                                           C& fx, C& fy, C& fz) {
#pragma clang fp contract(off)
    fx = C(0);
    fy = C(0);
    fz = C(0);
    if (p.force_mode == FORCE_UNIFORM) {
        fx = p.fu[0];
        fy = p.fu[1];
        fz = p.fu[2];
    } else if (p.force_mode == FORCE_FIELD) {
        fx = p.force[cell];
        fy = p.force[(size_t)p.ncell + cell];
        fz = p.force[2 * (size_t)p.ncell + cell];
    }
    if (p.brink) {
        const C s = p.brink[cell];
        fx -= s * mx;
        fy -= s * my;
        fz -= s * mz;
    }
}

// BGK collision with Guo forcing on deviations (reference/bgk.py:72-110):
//   u = (m + F/2)/rho,  f*_i = f_i - (f_i - feq_i(rho,u))/tau + G_i
//   feq_i - w_i = w_i (drho + rho (3 cu + 4.5 cu^2 - 1.5 u^2))
//   G_i = (1 - 1/(2 tau)) w_i [3 (c_i - u).F + 9 (c_i.u)(c_i.F)]
// evaluated per opposite pair (i, opp(i)): c_opp = -c_i flips the sign of
// the odd terms (3 cu, 3 c.F) and keeps the even ones.  The collision must
// leave the mass unchanged and add exactly F to the momentum; in float the
// rounded coefficients would miss both by a fixed relative amount every step
// (a systematic drift), so the rest population takes the mass residual and
// the three axis pairs take the momentum residual.  FORCED = false compiles
// the force terms out.  Returns rho and the Guo-corrected velocity
// (macroscopic_velocity, bgk.py:129-142).
//
// Contraction is off and every fused multiply-add is written out, so the
// result of a cell is fixed by this source alone: the scalar and vector-row
// steps, the read-back and the adjoint's recomputation agree bit for bit.
template <class C, bool FORCED = true>
__device__ __forceinline__ void collide(C d[Q], const Params<C>& p, uint32_t cell, C& rho_o,
                                        C& ux_o, C& uy_o, C& uz_o) {
#pragma clang fp contract(off)
    C drho, mx, my, mz;
    moments(d, drho, mx, my, mz);
    const C rho = C(1) + drho;
    C fx = C(0), fy = C(0), fz = C(0);
    if constexpr (FORCED) cell_force(p, cell, mx, my, mz, fx, fy, fz);
    const C rs = rho > C(1e-30) ? rho : C(1e-30);
    const C irs = C(1) / rs;
    C ux, uy, uz;
    if constexpr (FORCED) {
        ux = fma(C(0.5), fx, mx) * irs;
        uy = fma(C(0.5), fy, my) * irs;
        uz = fma(C(0.5), fz, mz) * irs;
    } else {
        ux = mx * irs;
        uy = my * irs;
        uz = mz * irs;
    }
    const C usq = fma(ux, ux, fma(uy, uy, uz * uz));
    const C om = p.omega;
    // even part without the cu^2 term: om w (drho - 1.5 rho u^2) [- 3 pref w u.F]
    const C t = fma(C(-1.5) * rho, usq, drho);
    C uf = C(0);
    if constexpr (FORCED) uf = fma(ux, fx, fma(uy, fy, uz * fz));
    C base[3], odd[3], quad[3];
#pragma unroll
    for (int k = 1; k < 3; ++k) {
        base[k] = p.k_base[k] * t;
        if constexpr (FORCED) base[k] = fma(-p.k_fodd[k], uf, base[k]);
        odd[k] = p.k_odd[k] * rho;
        quad[k] = p.k_quad[k] * rho;
    }
    C even[Q], oddv[Q];
#pragma unroll
    for (int i = 1; i < Q; i += 2) {
        const int k = i <= 6 ? 1 : 2;
        const C cu = cdot<C>(i, ux, uy, uz);
        even[i] = fma(quad[k] * cu, cu, base[k]);
        oddv[i] = odd[k] * cu;
        if constexpr (FORCED) {
            const C cf = cdot<C>(i, fx, fy, fz);
            even[i] = fma(p.k_fquad[k] * cu, cf, even[i]);
            oddv[i] = fma(p.k_fodd[k], cf, oddv[i]);
        }
    }
    // momentum residual: sum_i c_i (om feq_i + G_i) = om m + F
    C hx, hy, hz;
    if constexpr (FORCED) {
        hx = C(0.5) * fma(om, mx, fx);
        hy = C(0.5) * fma(om, my, fy);
        hz = C(0.5) * fma(om, mz, fz);
    } else {
        hx = C(0.5) * (om * mx);
        hy = C(0.5) * (om * my);
        hz = C(0.5) * (om * mz);
    }
    oddv[1] = hx - (oddv[7] + oddv[9] + oddv[11] + oddv[13]);
    oddv[3] = hy - (oddv[7] - oddv[9] + oddv[15] + oddv[17]);
    oddv[5] = hz - (oddv[11] - oddv[13] + oddv[15] - oddv[17]);
    // mass residual: sum_i (om feq_i + G_i) = om drho
    C sum_even = C(0);
#pragma unroll
    for (int i = 1; i < Q; i += 2) sum_even += even[i];
    const C rest = fma(C(-2), sum_even, om * drho);
    // relaxation as d + (om feq + G) - om d, one fma per population
    d[0] = fma(-om, d[0], d[0] + rest);
#pragma unroll
    for (int i = 1; i < Q; i += 2) {
        d[i] = fma(-om, d[i], d[i] + (even[i] + oddv[i]));
        d[i + 1] = fma(-om, d[i + 1], d[i + 1] + (even[i] - oddv[i]));
    }
    rho_o = rho;
    ux_o = ux;
    uy_o = uy;
    uz_o = uz;
}

// Read-back moments of pre-collision deviations (no collision), with the
// operations of collide(); a zero force gives the same value as the
// unforced branch there (fma(0.5, 0, m) == m).
template <class C>
__device__ __forceinline__ void readback_moments(const C d[Q], const Params<C>& p, uint32_t cell,
                                                 C& rho_o, C& ux, C& uy, C& uz) {
#pragma clang fp contract(off)
    C drho, mx, my, mz;
    moments(d, drho, mx, my, mz);
    const C rho = C(1) + drho;
    C fx, fy, fz;
    cell_force(p, cell, mx, my, mz, fx, fy, fz);
    const C rs = rho > C(1e-30) ? rho : C(1e-30);
    const C irs = C(1) / rs;
    ux = fma(C(0.5), fx, mx) * irs;
    uy = fma(C(0.5), fy, my) * irs;
    uz = fma(C(0.5), fz, mz) * irs;
    rho_o = rho;
}

__device__ __forceinline__ void cell_coords(uint32_t cell, uint32_t ny, uint32_t nz, uint32_t& x,
                                            uint32_t& y, uint32_t& z) {
    z = cell % nz;
    const uint32_t t = cell / nz;
    y = t % ny;
    x = t / ny;
}

// Wavefront (64-lane) sum of a double.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace d3q19
