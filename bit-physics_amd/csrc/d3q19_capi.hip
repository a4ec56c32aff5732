// d3q19_capi.hip -- C ABI of the D3Q19 engine (declared in include/d3q19.h).
//
// A lattice owns two population buffers (ping-pong, row-blocked layout) holding the
// post-collision deviations of the current step; the reference's
// pre-collision state f^n is produced on demand by the read-back kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/d3q19.h"
#include "d3q19_host.hpp"
#include "d3q19_kernels.hpp"
#include "d3q19_moment_codec.hpp"

using namespace d3q19;

namespace d3q19 {

thread_local std::string g_err;

int set_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace d3q19

namespace {

inline int fail(int code, const std::string& msg) { return set_error(code, msg); }

#define HIPCHK(expr) D3Q19_HIPCHK(expr)

inline uint32_t nblocks(uint64_t n) { return (uint32_t)((n + BLOCK - 1) / BLOCK); }

size_t storage_bytes(int s) {
    switch (s) {
        case D3Q19_F64: return 8;
        case D3Q19_F32: return 4;
        case D3Q19_F16:
        case D3Q19_BF16:
        case D3Q19_BF16SR:
        case D3Q19_ME16:
        case D3Q19_ME16_REF: return 2;
    }
    return 0;
}

// Device scratch allocation that frees itself.
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc(&p, n); }
};

}  // namespace

struct d3q19_lattice {
    d3q19_config cfg;
    hipStream_t stream = nullptr;  // nullptr: the default stream
    uint32_t ncell = 0;
    bool rows = true;       // row-blocked layout on the row grid; false: short rows
    uint32_t nzp = 0;       // padded row length (rows), nz (short rows)
    size_t row_bytes = 0;   // byte stride of a z-row: 19 * nzp * esize (rows), nz * esize
    size_t dir_bytes = 0;   // byte stride of a direction: nzp * esize (rows), ncell * esize
    size_t lat_bytes = 0;   // one population buffer
    size_t esize = 0;
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;
    void* force = nullptr;   // compute-type (3, ncell)
    void* brink = nullptr;   // compute-type (ncell)
    uint8_t* solid = nullptr;
    double* partials = nullptr;
    double* dout = nullptr;  // device results of reductions
    void* uscratch = nullptr;  // read-back velocity (3, ncell) in the compute type (enstrophy)
    bool has_state = false;
    bool local = false;  // buffer cur holds pre-collision f^n in place
    // D3Q19_ME16: buf[0] and buf[1] hold moment codes in the population layout
    // (2-byte elements, moment k in place of direction k); mq their
    // quantization constants (device).  D3Q19_ME16_REF: buf[0] the codes
    // (moment-major, k * ncell + cell), buf[1] and ref_b two FP64 population
    // arrays in the reference layout.
    bool me = false;
    MeQuant* mq = nullptr;
    bool me_ref = false;
    void* ref_b = nullptr;
    size_t codes_bytes = 0;
    bool me_ranges = false;
    double me_lo[Q] = {}, me_hi[Q] = {};
    double* mr_partials = nullptr;  // moment-range partials and result
    double* mr_out = nullptr;
    uint32_t chunk_override = 0;  // plane chunks per fused step, 0 = by step_chunks (D3Q19_STEP_CHUNKS)
    // D3Q19_ME16 on the row grid: the codes live in one ring of ring_np x
    // planes (buf[0]; buf[1] unused), the state's plane 0 at ring position
    // ring_base; a step writes the next state ring_gap planes before it, in
    // place (me_row_in, me_row_out, me_ring_sync).
    bool ring = false;
    uint32_t ring_np = 0, ring_base = 0, ring_gap = 0;
    uint32_t ring_blocks = 0;       // blocks of the fused step per plane
    uint32_t ring_epoch = 0;        // ring steps taken (ring_done reaches ring_blocks per step)
    size_t plane_bytes = 0;         // one x plane of the row-blocked layout
    void* ring_saved = nullptr;     // plane 0 of the state a step reads
    uint32_t* ring_done = nullptr;  // nx counters, ME_RING_STRIDE apart
    uint32_t* ring_err = nullptr;   // set by a wait that gave up
    bool ring_fault_test = false;   // D3Q19_RING_FAULT_TEST: every wait gives up at once
    int64_t steps = 0;
    size_t bytes = 0;
    uint32_t lds_per_cu = 0;  // LDS bytes per CU of the device, 0 if unknown
};

namespace {

// Polls of a short count before a ring step's wait gives up (about 10 ms).
constexpr uint32_t ME_RING_MAX_SPINS = 40000;

template <class C>
Params<C> make_params(const d3q19_lattice* L) {
    Params<C> p;
    std::memset(&p, 0, sizeof(p));
    p.nx = (uint32_t)L->cfg.nx;
    p.ny = (uint32_t)L->cfg.ny;
    p.nz = (uint32_t)L->cfg.nz;
    p.ncell = L->ncell;
    p.nzp = L->nzp;
    p.dir_bytes = (uint32_t)L->dir_bytes;
    p.row_bytes = L->row_bytes;
    set_collision_constants(p, L->cfg.tau);
    p.force_mode = L->cfg.force_mode;
    p.wall_mask = L->cfg.boundary == D3Q19_BC_HALFWAY ? L->cfg.wall_faces : 0u;
    for (int a = 0; a < 3; ++a) {
        p.fu[a] = (C)L->cfg.force[a];
        p.wtop[a] = (C)L->cfg.wall_velocity_top[a];
        p.wbot[a] = (C)L->cfg.wall_velocity_bottom[a];
        for (int k = 0; k < 6; ++k) p.face_u[k][a] = (C)L->cfg.face_velocity[k][a];
    }
    p.force = (const C*)L->force;
    p.brink = (const C*)L->brink;
    p.solid = L->cfg.boundary == D3Q19_BC_HALFWAY ? L->solid : nullptr;
    p.local = L->local ? 1 : 0;
    set_zwall_terms(p);
    p.me = L->mq;
    p.seed = (uint32_t)L->steps * 0x9e3779b9u + 0x7f4a7c15u;
    if (L->ring) {
        p.ring_np = L->ring_np;
        p.ring_in = p.ring_out = L->ring_base;
    }
    return p;
}

template <class Tag>
typename Codec<Tag>::raw* buf_of(const d3q19_lattice* L, int which) {
    return (typename Codec<Tag>::raw*)L->buf[L->ring ? 0 : which];
}

// Launch shape of every population kernel: (z, y, x) blocks whose x extent
// is a whole number of wavefronts of one row (64, 128 or 256 cells); for
// short rows a linear grid of one thread per cell.
struct StepGrid {
    dim3 grid, block;
    uint32_t nblocks;
    uint32_t step_lds = 0;  // dynamic LDS of a fused-step block (cap_lds)
};

// Dynamic LDS of a fused-step block: caps the resident blocks per CU at
// step_blocks_per_cu<Tag>().
template <class Tag>
uint32_t cap_lds(const d3q19_lattice* L) {
    return reserve_lds(L->lds_per_cu, step_blocks_per_cu<Tag>());
}

// block: threads per block (step_block, or fused_block for k_step).
template <class Tag>
StepGrid step_grid(const d3q19_lattice* L, uint32_t block = step_block<Tag>()) {
    StepGrid sg;
    const uint32_t nx = L->cfg.nx, ny = L->cfg.ny, nz = L->cfg.nz;
    sg.step_lds = cap_lds<Tag>(L);
    if (!L->rows) {
        sg.block = dim3(block, 1, 1);
        sg.grid = dim3((L->ncell + block - 1) / block, 1, 1);
        sg.nblocks = sg.grid.x;
        return sg;
    }
    const uint32_t bz = row_block(nz);
    const uint32_t by = block / bz;
    sg.block = dim3(bz, by, 1);
    sg.grid = dim3((nz + bz - 1) / bz, (ny + by - 1) / by, nx);
    sg.nblocks = sg.grid.x * sg.grid.y * sg.grid.z;
    return sg;
}

template <class Tag, bool FORCED, bool ROWS>
void launch_k_step_chunk(d3q19_lattice* L, const StepGrid& sg, const typename Codec<Tag>::raw* in,
                         typename Codec<Tag>::raw* out, const Params<typename Codec<Tag>::C>& p, float* r, float* v) {
    switch (L->cfg.boundary) {
        case D3Q19_BC_PERIODIC:
            k_step<Tag, BC_PERIODIC, FORCED, ROWS><<<sg.grid, sg.block, sg.step_lds, L->stream>>>(in, out, p, r, v);
            break;
        case D3Q19_BC_YWALLS:
            k_step<Tag, BC_YWALLS, FORCED, ROWS><<<sg.grid, sg.block, sg.step_lds, L->stream>>>(in, out, p, r, v);
            break;
        default:
            k_step<Tag, BC_HALFWAY, FORCED, ROWS><<<sg.grid, sg.block, sg.step_lds, L->stream>>>(in, out, p, r, v);
            break;
    }
}

// Plane chunks of the lattice's fused step (1 on the linear grid of short rows).
template <class Tag>
uint32_t fused_step_chunks(const d3q19_lattice* L) {
    return L->rows ? step_chunks(L->ncell, (uint32_t)L->cfg.nx, step_chunk_glups<Tag>(), step_chunk_min_us<Tag>(),
                                 STEP_CHUNK_US, L->chunk_override)
                   : 1u;
}

// One fused step, launched in plane chunks on the row grid (step_chunks):
// chunk k covers the x planes [k S, min((k + 1) S, nx)), S = chunk_planes.
template <class Tag, bool FORCED, bool ROWS>
void launch_k_step(d3q19_lattice* L, const StepGrid& sf, const typename Codec<Tag>::raw* in,
                   typename Codec<Tag>::raw* out, const Params<typename Codec<Tag>::C>& p0, float* r, float* v) {
    const uint32_t nx = (uint32_t)L->cfg.nx;
    const uint32_t n = ROWS ? fused_step_chunks<Tag>(L) : 1u;
    const uint32_t S = chunk_planes(nx, n);
    for (uint32_t x0 = 0; x0 < nx; x0 += S) {
        StepGrid sg = sf;
        Params<typename Codec<Tag>::C> p = p0;
        if (n > 1) {
            sg.grid.z = std::min(S, nx - x0);
            p.x0 = x0;
        }
        launch_k_step_chunk<Tag, FORCED, ROWS>(L, sg, in, out, p, r, v);
    }
}

template <class Tag, bool ROWS>
int launch_step_l(d3q19_lattice* L, int32_t n, float* rho, float* u) {
    typedef typename Codec<Tag>::C C;
    const StepGrid sg = step_grid<Tag>(L);                     // k_collide_local
    const StepGrid sf = step_grid<Tag>(L, fused_block<Tag>());  // k_step
    for (int32_t s = 0; s < n; ++s) {
        const Params<C> p = make_params<C>(L);
        const bool forced = p.force_mode != D3Q19_FORCE_NONE || p.brink != nullptr;
        const typename Codec<Tag>::raw* in = buf_of<Tag>(L, L->cur);
        typename Codec<Tag>::raw* out = buf_of<Tag>(L, L->cur ^ 1);
        float* r = (s == n - 1) ? rho : nullptr;
        float* v = (s == n - 1) ? u : nullptr;
        if (L->local) {
            // first step after a load: collide f^n in place, no streaming (a
            // ring collides each cell into its own slots: in == out)
            k_collide_local<Tag, ROWS><<<sg.grid, sg.block, 0, L->stream>>>(in, out, p, r, v);
            L->local = false;
        } else {
            Params<C> q = p;
            if (L->ring) {
                // plane 0 of the state read, for the last planes' pulls once its slot is reused
                HIPCHK(hipMemcpyAsync(L->ring_saved, (const char*)L->buf[0] + (size_t)L->ring_base * L->plane_bytes,
                                      L->plane_bytes, hipMemcpyDeviceToDevice, L->stream));
                q.ring_out = (L->ring_base + L->ring_np - L->ring_gap) % L->ring_np;
                q.ring_gap = L->ring_gap;
                q.ring_target = L->ring_blocks * (L->ring_epoch + 1u);
                q.ring_spins = ME_RING_MAX_SPINS;
                if (L->ring_fault_test) {  // a count no block reaches, one poll: the failure path
                    q.ring_target += 0x40000000u;
                    q.ring_spins = 1u;
                }
                q.ring_saved = (const char*)L->ring_saved;
                q.ring_done = L->ring_done;
                q.ring_err = L->ring_err;
            }
            if (forced)
                launch_k_step<Tag, true, ROWS>(L, sf, in, out, q, r, v);
            else
                launch_k_step<Tag, false, ROWS>(L, sf, in, out, q, r, v);
            if (L->ring) {
                L->ring_base = q.ring_out;
                L->ring_epoch++;
            }
        }
        HIPCHK(hipGetLastError());
        if (!L->ring) L->cur ^= 1;
        L->steps++;
    }
    return D3Q19_OK;
}

template <class Tag>
int launch_step(d3q19_lattice* L, int32_t n, float* rho, float* u) {
    return L->rows ? launch_step_l<Tag, true>(L, n, rho, u) : launch_step_l<Tag, false>(L, n, rho, u);
}

__global__ void k_clear_word(uint32_t* w) {
    if (threadIdx.x == 0) *w = 0u;
}

// A loaded state replaces whatever a failed ring step left: clear its flag
// (a kernel rather than a memset node, so that the load may be captured).
int ring_clear(d3q19_lattice* L) {
    if (!L->ring) return D3Q19_OK;
    k_clear_word<<<1, 64, 0, L->stream>>>(L->ring_err);
    HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

template <class Tag, class In>
int launch_init(d3q19_lattice* L, const In* f_dev) {
    typedef typename Codec<Tag>::C C;
    L->steps = 0;  // before make_params: the stochastic rounding of the load is keyed on it
    const Params<C> p = make_params<C>(L);
    const StepGrid sg = step_grid<Tag>(L);
    if (L->rows)
        k_load<Tag, In, true><<<sg.grid, sg.block, 0, L->stream>>>(f_dev, buf_of<Tag>(L, L->cur), p);
    else
        k_load<Tag, In, false><<<sg.grid, sg.block, 0, L->stream>>>(f_dev, buf_of<Tag>(L, L->cur), p);
    HIPCHK(hipGetLastError());
    if (int e = ring_clear(L)) return e;
    L->has_state = true;
    L->local = true;
    L->steps = 0;
    return D3Q19_OK;
}

// feq(rho, u) straight into the lattice storage.
template <class Tag, class In>
int launch_init_eq(d3q19_lattice* L, const In* rho_dev, const In* u_dev) {
    typedef typename Codec<Tag>::C C;
    L->steps = 0;  // before make_params: the stochastic rounding of the load is keyed on it
    const Params<C> p = make_params<C>(L);
    const StepGrid sg = step_grid<Tag>(L);
    if (L->rows)
        k_load_eq<Tag, In, true><<<sg.grid, sg.block, 0, L->stream>>>(rho_dev, u_dev, buf_of<Tag>(L, L->cur), p);
    else
        k_load_eq<Tag, In, false><<<sg.grid, sg.block, 0, L->stream>>>(rho_dev, u_dev, buf_of<Tag>(L, L->cur), p);
    HIPCHK(hipGetLastError());
    if (int e = ring_clear(L)) return e;
    L->has_state = true;
    L->local = true;
    L->steps = 0;
    return D3Q19_OK;
}

template <class Tag, class Out, bool ROWS>
void launch_readback_l(d3q19_lattice* L, const StepGrid& sg, const Params<typename Codec<Tag>::C>& p, Out* f,
                       Out* rho, Out* u, const typename Codec<Tag>::raw* in) {
    switch (L->cfg.boundary) {
        case D3Q19_BC_PERIODIC:
            k_readback<Tag, BC_PERIODIC, Out, ROWS><<<sg.grid, sg.block, 0, L->stream>>>(in, p, f, rho, u);
            break;
        case D3Q19_BC_YWALLS:
            k_readback<Tag, BC_YWALLS, Out, ROWS><<<sg.grid, sg.block, 0, L->stream>>>(in, p, f, rho, u);
            break;
        default:
            k_readback<Tag, BC_HALFWAY, Out, ROWS><<<sg.grid, sg.block, 0, L->stream>>>(in, p, f, rho, u);
            break;
    }
}

template <class Tag, class Out>
int launch_readback(d3q19_lattice* L, Out* f, Out* rho, Out* u) {
    typedef typename Codec<Tag>::C C;
    const Params<C> p = make_params<C>(L);
    const StepGrid sg = step_grid<Tag>(L);
    if (L->rows)
        launch_readback_l<Tag, Out, true>(L, sg, p, f, rho, u, buf_of<Tag>(L, L->cur));
    else
        launch_readback_l<Tag, Out, false>(L, sg, p, f, rho, u, buf_of<Tag>(L, L->cur));
    HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

template <class Tag, bool ROWS>
void launch_diag_l(d3q19_lattice* L, const StepGrid& sg, const Params<typename Codec<Tag>::C>& p,
                   const typename Codec<Tag>::raw* in) {
    switch (L->cfg.boundary) {
        case D3Q19_BC_PERIODIC:
            k_diag<Tag, BC_PERIODIC, ROWS><<<sg.grid, sg.block, 0, L->stream>>>(in, p, L->partials);
            break;
        case D3Q19_BC_YWALLS:
            k_diag<Tag, BC_YWALLS, ROWS><<<sg.grid, sg.block, 0, L->stream>>>(in, p, L->partials);
            break;
        default:
            k_diag<Tag, BC_HALFWAY, ROWS><<<sg.grid, sg.block, 0, L->stream>>>(in, p, L->partials);
            break;
    }
}

template <class Tag>
int launch_diag(d3q19_lattice* L) {
    typedef typename Codec<Tag>::C C;
    const Params<C> p = make_params<C>(L);
    const StepGrid sg = step_grid<Tag>(L);
    if (L->rows)
        launch_diag_l<Tag, true>(L, sg, p, buf_of<Tag>(L, L->cur));
    else
        launch_diag_l<Tag, false>(L, sg, p, buf_of<Tag>(L, L->cur));
    HIPCHK(hipGetLastError());
    k_sum_partials<<<NDIAG, BLOCK, 0, L->stream>>>(L->partials, sg.nblocks, NDIAG, L->dout);
    HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

// Dispatch on the lattice storage type.
#define D3Q19_DISPATCH(L, FN, ...)                                              \
    [&]() -> int {                                                              \
        switch ((L)->cfg.storage) {                                             \
            case D3Q19_F64: return FN<TF64>(__VA_ARGS__);                       \
            case D3Q19_F32: return FN<TF32>(__VA_ARGS__);                       \
            case D3Q19_F16: return FN<TF16>(__VA_ARGS__);                       \
            case D3Q19_BF16: return FN<TBF16>(__VA_ARGS__);                     \
            case D3Q19_BF16SR: return FN<TBF16S>(__VA_ARGS__);                  \
            case D3Q19_ME16: return FN<TME16>(__VA_ARGS__);                     \
        }                                                                       \
        return fail(D3Q19_EINVAL, "bad storage type");                          \
    }()

bool is_ct_double(const d3q19_lattice* L) {
    return L->cfg.storage == D3Q19_F64 || L->cfg.storage == D3Q19_ME16_REF;
}
size_t ct_size(const d3q19_lattice* L) { return is_ct_double(L) ? 8 : 4; }

int stage_to_compute(d3q19_lattice* L, const void* src, int32_t dt, int32_t on_device, size_t n,
                     void* dst) {
    return stage_in(src, dt, on_device, n, dst, is_ct_double(L), L->stream);
}

int check_lat(const d3q19_lattice* L) {
    if (!L) return fail(D3Q19_EINVAL, "null lattice handle");
    return D3Q19_OK;
}

int check_dtype(int32_t dt) {
    if (dt != D3Q19_F64 && dt != D3Q19_F32) return fail(D3Q19_EINVAL, "dtype must be D3Q19_F64 or D3Q19_F32");
    return D3Q19_OK;
}

}  // namespace

namespace d3q19 {

static __global__ void k_convert_f32_to_f64(const float* a, double* b, size_t n) {
    size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) b[i] = a[i];
}
static __global__ void k_convert_f64_to_f32(const double* a, float* b, size_t n) {
    size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

namespace {
struct Tmp {
    void* p = nullptr;
    ~Tmp() {
        if (p) (void)hipFree(p);
    }
};
}  // namespace

int stage_in(const void* src, int dt, int on_device, size_t n, void* dst, bool dst_double, hipStream_t s) {
    if (dt != D3Q19_F64 && dt != D3Q19_F32) return set_error(D3Q19_EINVAL, "dtype must be F64 or F32");
    if (!src || !dst) return set_error(D3Q19_EINVAL, "null array");
    const size_t es = dt == D3Q19_F64 ? 8 : 4;
    if ((dt == D3Q19_F64) == dst_double) {
        D3Q19_HIPCHK(hipMemcpyAsync(dst, src, n * es, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
        if (!on_device) D3Q19_HIPCHK(hipStreamSynchronize(s));
        return D3Q19_OK;
    }
    Tmp tmp;
    const void* d = src;
    if (!on_device) {
        if (hipMalloc(&tmp.p, n * es) != hipSuccess) return set_error(D3Q19_ENOMEM, "staging allocation failed");
        D3Q19_HIPCHK(hipMemcpyAsync(tmp.p, src, n * es, hipMemcpyHostToDevice, s));
        d = tmp.p;
    }
    if (dst_double)
        k_convert_f32_to_f64<<<grid_for(n, BLOCK), BLOCK, 0, s>>>((const float*)d, (double*)dst, n);
    else
        k_convert_f64_to_f32<<<grid_for(n, BLOCK), BLOCK, 0, s>>>((const double*)d, (float*)dst, n);
    D3Q19_HIPCHK(hipGetLastError());
    D3Q19_HIPCHK(hipStreamSynchronize(s));
    return D3Q19_OK;
}

int stage_out(const void* src, bool src_double, size_t n, void* dst, int dt, int on_device, hipStream_t s) {
    if (dt != D3Q19_F64 && dt != D3Q19_F32) return set_error(D3Q19_EINVAL, "dtype must be F64 or F32");
    if (!src || !dst) return set_error(D3Q19_EINVAL, "null array");
    const size_t es = dt == D3Q19_F64 ? 8 : 4;
    const bool same = (dt == D3Q19_F64) == src_double;
    Tmp tmp;
    const void* from = src;
    if (!same) {
        void* conv = dst;
        if (!on_device) {
            if (hipMalloc(&tmp.p, n * es) != hipSuccess) return set_error(D3Q19_ENOMEM, "staging allocation failed");
            conv = tmp.p;
        }
        if (src_double)
            k_convert_f64_to_f32<<<grid_for(n, BLOCK), BLOCK, 0, s>>>((const double*)src, (float*)conv, n);
        else
            k_convert_f32_to_f64<<<grid_for(n, BLOCK), BLOCK, 0, s>>>((const float*)src, (double*)conv, n);
        D3Q19_HIPCHK(hipGetLastError());
        from = conv;
        if (on_device) {
            D3Q19_HIPCHK(hipStreamSynchronize(s));
            return D3Q19_OK;
        }
    }
    D3Q19_HIPCHK(hipMemcpyAsync(dst, from, n * es, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    D3Q19_HIPCHK(hipStreamSynchronize(s));
    return D3Q19_OK;
}

}  // namespace d3q19

namespace {
template <class Tag, bool ROWS>
void launch_moment_range_l(d3q19_lattice* L, const StepGrid& sg, const Params<typename Codec<Tag>::C>& p,
                           const typename Codec<Tag>::raw* in, uint32_t per) {
    switch (L->cfg.boundary) {
        case D3Q19_BC_PERIODIC:
            k_moment_range<Tag, BC_PERIODIC, ROWS><<<sg.grid, sg.block, 0, L->stream>>>(in, p, L->mr_partials, per);
            break;
        case D3Q19_BC_YWALLS:
            k_moment_range<Tag, BC_YWALLS, ROWS><<<sg.grid, sg.block, 0, L->stream>>>(in, p, L->mr_partials, per);
            break;
        default:
            k_moment_range<Tag, BC_HALFWAY, ROWS><<<sg.grid, sg.block, 0, L->stream>>>(in, p, L->mr_partials, per);
            break;
    }
}

// Cells each thread of k_moment_range folds: up to 8, while the grid keeps
// 2048 blocks or more.
inline uint32_t moment_range_per(uint32_t nblocks) { return std::max(1u, std::min(8u, nblocks / 2048u)); }

// stored: the moments of the stored state as it lies in the buffer (the
// post-collision populations of the last step, or a loaded state) instead of
// the state f^n the read-back gives.
template <class Tag>
int launch_moment_range(d3q19_lattice* L, bool stored) {
    typedef typename Codec<Tag>::C C;
    Params<C> p = make_params<C>(L);
    if (stored) p.local = 1;
    StepGrid sg = step_grid<Tag>(L);
    const uint32_t per = moment_range_per(sg.nblocks);
    if (L->rows) {
        sg.grid.z = (sg.grid.z + per - 1) / per;
        sg.nblocks = sg.grid.x * sg.grid.y * sg.grid.z;
        launch_moment_range_l<Tag, true>(L, sg, p, buf_of<Tag>(L, L->cur), per);
    } else {
        sg.grid.x = (sg.grid.x + per - 1) / per;
        sg.nblocks = sg.grid.x;
        launch_moment_range_l<Tag, false>(L, sg, p, buf_of<Tag>(L, L->cur), per);
    }
    HIPCHK(hipGetLastError());
    k_moment_range_final<<<2 * Q, 256, 0, L->stream>>>(L->mr_partials, sg.nblocks, L->mr_out);
    HIPCHK(hipGetLastError());
    return D3Q19_OK;
}
}  // namespace

namespace {
template <class Tag, bool ROWS>
void launch_health_l(d3q19_lattice* L, const StepGrid& sg, const Params<typename Codec<Tag>::C>& p,
                     const typename Codec<Tag>::raw* in) {
    switch (L->cfg.boundary) {
        case D3Q19_BC_PERIODIC:
            k_health<Tag, BC_PERIODIC, ROWS><<<sg.grid, sg.block, 0, L->stream>>>(in, p, L->partials);
            break;
        case D3Q19_BC_YWALLS:
            k_health<Tag, BC_YWALLS, ROWS><<<sg.grid, sg.block, 0, L->stream>>>(in, p, L->partials);
            break;
        default:
            k_health<Tag, BC_HALFWAY, ROWS><<<sg.grid, sg.block, 0, L->stream>>>(in, p, L->partials);
            break;
    }
}

template <class Tag>
int launch_health(d3q19_lattice* L) {
    typedef typename Codec<Tag>::C C;
    const Params<C> p = make_params<C>(L);
    const StepGrid sg = step_grid<Tag>(L);
    if (L->rows)
        launch_health_l<Tag, true>(L, sg, p, buf_of<Tag>(L, L->cur));
    else
        launch_health_l<Tag, false>(L, sg, p, buf_of<Tag>(L, L->cur));
    HIPCHK(hipGetLastError());
    k_health_final<<<NHEALTH, 256, 0, L->stream>>>(L->partials, sg.nblocks, L->dout);
    HIPCHK(hipGetLastError());
    return D3Q19_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
// D3Q19_ME16_REF: the moment codes stepped in the moment-encoded variant's own
// FP64 arithmetic (packages/lattice-boltzmann-d3q19-frontier-moment-encoded,
// src/lbm_me.cpp:386-409: decode, collide, stream + bounce, encode).  A step
// is two kernels: decode + collide into the FP64 post-collision array ref_b,
// then pull-stream + y-walls + encode back into the codes.  buf[1] (ref_a)
// holds decoded populations for read-back and the loads' staging.
// Boundaries: periodic or the reference's y-walls.
// ---------------------------------------------------------------------------
extern "C" int d3q19_destroy(d3q19_lattice* L);

namespace {

// Quantization constants of a D3Q19_ME16 lattice from its ranges: the
// deviation-moment grid lo'_k = fp32(lo_k - (M w)_k), step_k = the power of
// two at or above (hi_k - lo_k) / 65535 (so the 65536 codes cover [lo, hi]
// and 1 / step_k is exact), and the decode offsets sum_k Minv[i][k] lo'_k
// formed in FP64 and kept as a sum of two floats.
MeQuant me_quant(const double lo[19], const double hi[19]) {
    MeQuant q;
    double lod[Q];
    for (int k = 0; k < Q; ++k) {
        double mw = 0.0;  // moment k of the rest state
        for (int i = 0; i < Q; ++i) mw += me_mono(k, i) * wgt<double>(i);
        const double st = std::exp2(std::ceil(std::log2((hi[k] - lo[k]) / 65535.0)));
        lod[k] = (double)(float)(lo[k] - mw);
        q.step[k] = (float)st;
        q.lo[k] = (float)lod[k];
        q.sc[k] = (float)(1.0 / st);
    }
    for (int i = 0; i < Q; ++i) {
        double l = 0.0;
        for (int k = 0; k < Q; ++k) l += 0.25 * ME_MINV4[i][k] * lod[k];
        q.L[i] = (float)l;
        q.Ll[i] = (float)(l - (double)q.L[i]);
    }
    return q;
}

MeRef make_me_ref(const d3q19_lattice* L) {
    MeRef r;
    for (int k = 0; k < Q; ++k) {
        r.lo[k] = L->me_lo[k];
        r.hi[k] = L->me_hi[k];
    }
    return r;
}

double* ref_a(const d3q19_lattice* L) { return (double*)L->buf[1]; }
double* ref_b(const d3q19_lattice* L) { return (double*)L->ref_b; }

int create_me_ref(const d3q19_config* cfg, void* stream, d3q19_lattice** out) {
    if (cfg->boundary != D3Q19_BC_PERIODIC && cfg->boundary != D3Q19_BC_YWALLS)
        return fail(D3Q19_EINVAL, "D3Q19_ME16_REF runs the reference's step: periodic or y-wall boundaries");
    d3q19_lattice* L = new d3q19_lattice();
    L->cfg = *cfg;
    L->ncell = (uint32_t)((uint64_t)cfg->nx * cfg->ny * cfg->nz);
    L->me_ref = true;
    L->rows = false;
    L->esize = 2;
    L->nzp = (uint32_t)cfg->nz;
    L->dir_bytes = (size_t)L->ncell * 2;
    L->row_bytes = (size_t)cfg->nz * 2;
    L->codes_bytes = (size_t)Q * L->ncell * 2;
    L->lat_bytes = (size_t)Q * L->ncell * 8;
    L->stream = (hipStream_t)stream;
    const size_t n3 = 3 * (size_t)L->ncell;
    if (hipMalloc(&L->buf[0], L->codes_bytes) != hipSuccess || hipMalloc(&L->buf[1], L->lat_bytes) != hipSuccess ||
        hipMalloc(&L->ref_b, L->lat_bytes) != hipSuccess || hipMalloc((void**)&L->dout, 8 * sizeof(double)) != hipSuccess) {
        d3q19_destroy(L);
        return fail(D3Q19_ENOMEM, "device allocation of the lattice failed");
    }
    L->bytes = L->codes_bytes + 2 * L->lat_bytes + 64;
    if (cfg->force_mode == D3Q19_FORCE_UNIFORM) {
        // the reference's bgk_step takes a per-cell force array
        std::vector<double> fh(n3);
        for (size_t a = 0; a < 3; ++a)
            for (size_t c = 0; c < L->ncell; ++c) fh[a * L->ncell + c] = cfg->force[a];
        if (hipMalloc(&L->force, n3 * sizeof(double)) != hipSuccess ||
            hipMemcpy(L->force, fh.data(), n3 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            d3q19_destroy(L);
            return fail(D3Q19_ENOMEM, "force array allocation failed");
        }
        L->bytes += n3 * sizeof(double);
    }
    *out = L;
    return D3Q19_OK;
}

int me_ref_decode(d3q19_lattice* L) {
    k_me_ref_decode<<<nblocks(L->ncell), BLOCK, 0, L->stream>>>((const uint16_t*)L->buf[0], ref_a(L), make_me_ref(L),
                                                               L->ncell);
    HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

int me_ref_encode(d3q19_lattice* L, const double* f) {
    k_me_ref_encode<<<nblocks(L->ncell), BLOCK, 0, L->stream>>>(f, (uint16_t*)L->buf[0], make_me_ref(L), L->ncell);
    HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

// f (19, ncell) of dtype dt, host or device -> FP64 array A -> codes.
int me_ref_set_populations(d3q19_lattice* L, const void* f, int32_t dt, int32_t on_device) {
    if (!L->me_ranges) return fail(D3Q19_ESTATE, "D3Q19_ME16_REF lattice needs d3q19_set_moment_ranges first");
    if (int e = stage_in(f, dt, on_device, (size_t)Q * L->ncell, ref_a(L), true, L->stream)) return e;
    if (int e = me_ref_encode(L, ref_a(L))) return e;
    HIPCHK(hipStreamSynchronize(L->stream));
    L->has_state = true;
    L->local = true;
    L->steps = 0;
    return D3Q19_OK;
}

// feq_field(rho, u) in FP64 (rho may be NULL: rho = 1), then its codes.
int me_ref_set_equilibrium(d3q19_lattice* L, const void* rho, const void* u, int32_t dt, int32_t on_device) {
    if (!L->me_ranges) return fail(D3Q19_ESTATE, "D3Q19_ME16_REF lattice needs d3q19_set_moment_ranges first");
    const size_t nc = L->ncell;
    double* ru = ref_b(L);  // rho (nc) then u (3 nc), FP64
    if (rho)
        if (int e = stage_in(rho, dt, on_device, nc, ru, true, L->stream)) return e;
    if (int e = stage_in(u, dt, on_device, 3 * nc, ru + nc, true, L->stream)) return e;
    k_fn_feq<double><<<nblocks(nc), BLOCK, 0, L->stream>>>(rho ? ru : nullptr, ru + nc, ref_a(L), (uint32_t)nc);
    HIPCHK(hipGetLastError());
    if (int e = me_ref_encode(L, ref_a(L))) return e;
    HIPCHK(hipStreamSynchronize(L->stream));
    L->has_state = true;
    L->local = true;
    L->steps = 0;
    return D3Q19_OK;
}

int me_ref_step(d3q19_lattice* L, int32_t n) {
    const uint32_t nc = L->ncell;
    const double* force = L->cfg.force_mode == D3Q19_FORCE_NONE ? nullptr : (const double*)L->force;
    const MeRef r = make_me_ref(L);
    MeRefWalls w;
    w.on = L->cfg.boundary == D3Q19_BC_YWALLS;
    for (int a = 0; a < 3; ++a) {
        w.top[a] = L->cfg.wall_velocity_top[a];
        w.bot[a] = L->cfg.wall_velocity_bottom[a];
    }
    uint16_t* codes = (uint16_t*)L->buf[0];
    for (int32_t s = 0; s < n; ++s) {
        k_me_ref_decode_collide<<<nblocks(nc), BLOCK, 0, L->stream>>>(codes, ref_b(L), force, r, nc, L->cfg.tau);
        HIPCHK(hipGetLastError());
        k_me_ref_stream_encode<<<nblocks(nc), BLOCK, 0, L->stream>>>(ref_b(L), codes, r, w, L->cfg.nx, L->cfg.ny,
                                                                     L->cfg.nz);
        HIPCHK(hipGetLastError());
        L->steps++;
    }
    return D3Q19_OK;
}

// Decoded f^n and/or density_field / macroscopic_velocity(f^n, force) of it.
int me_ref_read(d3q19_lattice* L, void* f, void* rho, void* u, int32_t dt, int32_t on_device) {
    const size_t nc = L->ncell;
    if (int e = me_ref_decode(L)) return e;
    if (f)
        if (int e = stage_out(ref_a(L), true, (size_t)Q * nc, f, dt, on_device, L->stream)) return e;
    if (rho || u) {
        const double* force = L->cfg.force_mode == D3Q19_FORCE_NONE ? nullptr : (const double*)L->force;
        double* ru = ref_b(L);
        k_fn_moments<double><<<nblocks(nc), BLOCK, 0, L->stream>>>(ref_a(L), force, ru, nullptr, ru + nc,
                                                                   (uint32_t)nc);
        HIPCHK(hipGetLastError());
        if (rho)
            if (int e = stage_out(ru, true, nc, rho, dt, on_device, L->stream)) return e;
        if (u)
            if (int e = stage_out(ru + nc, true, 3 * nc, u, dt, on_device, L->stream)) return e;
    }
    HIPCHK(hipStreamSynchronize(L->stream));
    return D3Q19_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Allocation-free read-back of the moments to host memory.  The stateful
// d3q19_read_moments stages host outputs through a device buffer it
// allocates per call; this variant stages them through `scratch`, a
// fixed-size device array of the library (one per GPU, alive until the
// process ends), so it makes no hipMalloc/hipFree and no allocating call.
// The array's address on a device is looked up once per device with
// hipGetSymbolAddress (in host code `scratch` names a host-side stand-in, not
// device memory) and kept in g_scratch_ptr; only that pointer is passed to
// kernels and copies.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)D3Q19_NOALLOC_SCRATCH_BYTES;
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

// Device address of `scratch` on device `dev` (looked up on first use, with
// that device current, then cached).
int scratch_on(int dev, void** out) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail(D3Q19_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int saved = 0;
        HIPCHK(hipGetDevice(&saved));
        HIPCHK(hipSetDevice(dev));
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        const hipError_t r = hipSetDevice(saved);
        if (e != hipSuccess || !ptr)
            return fail(D3Q19_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        HIPCHK(r);
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return D3Q19_OK;
}

// Copy the state of a moment-coded ring between host memory (planes 0 .. nx - 1
// in order, the layout of a two-lattice state) and the ring: at most two
// contiguous runs of planes.
int ring_copy(d3q19_lattice* L, void* host, hipMemcpyKind kind) {
    const uint32_t nx = (uint32_t)L->cfg.nx;
    const uint32_t first = std::min(nx, L->ring_np - L->ring_base);
    char* ring = (char*)L->buf[0];
    char* h = (char*)host;
    const size_t pb = L->plane_bytes;
    if (kind == hipMemcpyDeviceToHost) {
        HIPCHK(hipMemcpyAsync(h, ring + (size_t)L->ring_base * pb, (size_t)first * pb, kind, L->stream));
        if (first < nx) HIPCHK(hipMemcpyAsync(h + (size_t)first * pb, ring, (size_t)(nx - first) * pb, kind, L->stream));
    } else {
        HIPCHK(hipMemcpyAsync(ring + (size_t)L->ring_base * pb, h, (size_t)first * pb, kind, L->stream));
        if (first < nx) HIPCHK(hipMemcpyAsync(ring, h + (size_t)first * pb, (size_t)(nx - first) * pb, kind, L->stream));
    }
    return D3Q19_OK;
}

// After a synchronisation: whether a ring step's wait gave up (its result
// would be unreliable).  The flag stays set until a state is loaded
// (ring_clear).
int ring_check(d3q19_lattice* L) {
    if (!L->ring) return D3Q19_OK;
    uint32_t err = 0;
    HIPCHK(hipMemcpyAsync(&err, L->ring_err, sizeof(err), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    if (err)
        return fail(D3Q19_EHIP, "a moment-coded step waited too long for the pulls of the planes it overwrites; "
                                "the lattice state is unreliable");
    return D3Q19_OK;
}

}  // namespace

extern "C" {

const char* d3q19_last_error(void) { return g_err.c_str(); }
const char* d3q19_version(void) { return "d3q19-mi355x 0.1.0 (gfx950)"; }

void d3q19_config_default(d3q19_config* cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->tau = 0.6;
    cfg->storage = D3Q19_F32;
    cfg->boundary = D3Q19_BC_PERIODIC;
    cfg->force_mode = D3Q19_FORCE_NONE;
}

int d3q19_create(const d3q19_config* cfg, void* stream, d3q19_lattice** out) {
    if (!cfg || !out) return fail(D3Q19_EINVAL, "null argument");
    *out = nullptr;
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return fail(D3Q19_EINVAL, "extents must be >= 1");
    const uint64_t ncell = (uint64_t)cfg->nx * cfg->ny * cfg->nz;
    const bool me = cfg->storage == D3Q19_ME16;
    const size_t es = storage_bytes(cfg->storage);
    if (es == 0) return fail(D3Q19_EINVAL, "storage must be F64, F32, F16, BF16, BF16SR or ME16");
    if (ncell >= (1ull << 32)) return fail(D3Q19_EINVAL, "lattice too large (nx * ny * nz must be < 2^32)");
    if (cfg->nx > 65535 || cfg->ny > 65535) return fail(D3Q19_EINVAL, "nx and ny must be <= 65535");
    if (!(cfg->tau > 0.5)) return fail(D3Q19_EINVAL, "tau must be > 0.5");
    if (cfg->boundary < D3Q19_BC_PERIODIC || cfg->boundary > D3Q19_BC_HALFWAY)
        return fail(D3Q19_EINVAL, "bad boundary mode");
    if (cfg->force_mode < D3Q19_FORCE_NONE || cfg->force_mode > D3Q19_FORCE_FIELD)
        return fail(D3Q19_EINVAL, "bad force mode");
    if (cfg->boundary == D3Q19_BC_HALFWAY && (cfg->wall_faces & ~63u))
        return fail(D3Q19_EINVAL, "wall_faces has bits beyond the six faces");
    if (hipSetDevice(cfg->device) != hipSuccess) return fail(D3Q19_EHIP, "hipSetDevice failed");
    if (cfg->storage == D3Q19_ME16_REF) return create_me_ref(cfg, stream, out);

    d3q19_lattice* L = new d3q19_lattice();
    L->cfg = *cfg;
    L->ncell = (uint32_t)ncell;
    L->esize = es;
    L->me = me;
    // Rows shorter than a wavefront would leave most lanes idle and most of a
    // padded row empty: such lattices are stored direction-major (the
    // reference's layout) and stepped on a linear grid, while the 32-bit
    // direction offsets 18 * ncell * es fit.
    L->rows = cfg->nz >= 64 || (uint64_t)Q * ncell * es >= (1ull << 32);
    if (L->rows) {
        L->nzp = padded_nz((uint32_t)cfg->nz);
        L->dir_bytes = (size_t)L->nzp * es;
        L->row_bytes = (size_t)Q * L->nzp * es;
        L->lat_bytes = (size_t)cfg->nx * cfg->ny * L->row_bytes;
    } else {
        L->nzp = (uint32_t)cfg->nz;
        L->dir_bytes = (size_t)ncell * es;
        L->row_bytes = (size_t)cfg->nz * es;
        L->lat_bytes = (size_t)Q * ncell * es;
    }
    L->stream = (hipStream_t)stream;  // NULL: the default stream
    L->lds_per_cu = device_lds_per_cu(cfg->device);
    if (const char* chunks = std::getenv("D3Q19_STEP_CHUNKS")) {
        const long c = std::strtol(chunks, nullptr, 10);
        L->chunk_override = c > 0 && c <= 65535 ? (uint32_t)c : 0u;
    }
    const size_t lat_bytes = L->lat_bytes;
    // Moment codes on the row grid: one ring of nx + gap planes instead of
    // two lattices (the gap from the blocks a CU holds: at most 4 of the
    // moment-coded step's 512-thread blocks).
    if (me && L->rows) {
        const StepGrid sf = step_grid<TME16>(L, fused_block<TME16>());
        const uint64_t resident = (uint64_t)std::max(device_cu_count(cfg->device), 1u) * 4u;
        L->ring = true;
        L->ring_gap = ring_gap((uint32_t)cfg->nx, (uint64_t)sf.grid.x * sf.grid.y, resident);
        if (const char* gap = std::getenv("D3Q19_RING_GAP")) {  // a test knob: a tighter gap
            const long g = std::strtol(gap, nullptr, 10);
            if (g >= 2 && g < (long)L->ring_gap) L->ring_gap = (uint32_t)g;
        }
        if (const char* ft = std::getenv("D3Q19_RING_FAULT_TEST"))  // a test knob: the waits' failure path
            L->ring_fault_test = std::strtol(ft, nullptr, 10) == 1;
        L->ring_np = (uint32_t)cfg->nx + L->ring_gap;
        L->plane_bytes = (size_t)cfg->ny * L->row_bytes;
        L->ring_blocks = sf.grid.x * sf.grid.y;
    }
    const size_t buf0_bytes = L->ring ? (size_t)L->ring_np * L->plane_bytes : lat_bytes;
    // partial sums: one slot per block of the population grid and of the
    // linear 256-thread grids (ncell / 256)
    const uint32_t nb = (uint32_t)std::max<uint64_t>(
        nblocks(ncell), L->rows ? (uint64_t)cfg->nx * cfg->ny * ((cfg->nz + 63) / 64) : 0);
    if (hipMalloc(&L->buf[0], buf0_bytes) != hipSuccess ||
        (!L->ring && hipMalloc(&L->buf[1], lat_bytes) != hipSuccess) ||
        (L->ring && (hipMalloc(&L->ring_saved, L->plane_bytes) != hipSuccess ||
                     hipMalloc((void**)&L->ring_done, (size_t)cfg->nx * ME_RING_STRIDE * sizeof(uint32_t)) != hipSuccess ||
                     hipMalloc((void**)&L->ring_err, sizeof(uint32_t)) != hipSuccess)) ||
        hipMalloc((void**)&L->partials, (size_t)nb * NDIAG * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&L->dout, 8 * sizeof(double)) != hipSuccess ||
        (me && hipMalloc((void**)&L->mq, sizeof(MeQuant)) != hipSuccess)) {
        d3q19_destroy(L);
        return fail(D3Q19_ENOMEM, "device allocation of the lattice failed");
    }
    if (L->ring && (hipMemset(L->ring_done, 0, (size_t)cfg->nx * ME_RING_STRIDE * sizeof(uint32_t)) != hipSuccess ||
                    hipMemset(L->ring_err, 0, sizeof(uint32_t)) != hipSuccess)) {
        d3q19_destroy(L);
        return fail(D3Q19_EHIP, "clearing the ring counters failed");
    }
    L->bytes = (L->ring ? buf0_bytes + L->plane_bytes + ((size_t)cfg->nx * ME_RING_STRIDE + 1) * sizeof(uint32_t)
                        : 2 * lat_bytes) +
               (size_t)nb * NDIAG * sizeof(double) + 64 + (me ? sizeof(MeQuant) : 0);
    // Row padding (z >= nz of a padded row) is never written by a kernel:
    // zero it once, so that the raw state (read_state, checkpoints) is a
    // function of the lattice's content alone.
    if (L->rows && L->nzp != (uint32_t)cfg->nz) {
        if (hipMemset(L->buf[0], 0, buf0_bytes) != hipSuccess ||
            (L->buf[1] && hipMemset(L->buf[1], 0, lat_bytes) != hipSuccess)) {
            d3q19_destroy(L);
            return fail(D3Q19_EHIP, "clearing the lattice padding failed");
        }
    }
    *out = L;
    return D3Q19_OK;
}

int d3q19_destroy(d3q19_lattice* L) {
    if (!L) return D3Q19_OK;
    (void)hipStreamSynchronize(L->stream);
    void* ps[] = {L->buf[0],      L->buf[1],    L->force,       L->brink,     L->solid,
                  L->partials,    L->dout,      L->uscratch,    L->mr_partials, L->mr_out,
                  L->ref_b,       L->mq,        L->ring_saved,  L->ring_done, L->ring_err};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    delete L;
    return D3Q19_OK;
}

int d3q19_set_stream(d3q19_lattice* L, void* stream) {
    if (int e = check_lat(L)) return e;
    if ((hipStream_t)stream != L->stream) HIPCHK(hipStreamSynchronize(L->stream));
    L->stream = (hipStream_t)stream;
    return D3Q19_OK;
}

int d3q19_set_populations(d3q19_lattice* L, const void* f, int32_t dt, int32_t on_device) {
    if (int e = check_lat(L)) return e;
    if (int e = check_dtype(dt)) return e;
    if (!f) return fail(D3Q19_EINVAL, "null populations");
    if (L->me_ref) return me_ref_set_populations(L, f, dt, on_device);
    if (L->me && !L->me_ranges) return fail(D3Q19_ESTATE, "D3Q19_ME16 lattice needs d3q19_set_moment_ranges first");
    const size_t n = (size_t)Q * L->ncell;
    const size_t es = dt == D3Q19_F64 ? 8 : 4;
    DevBuf tmp;
    const void* src = f;
    if (!on_device) {
        if (tmp.alloc(n * es) != hipSuccess) return fail(D3Q19_ENOMEM, "staging allocation failed");
        HIPCHK(hipMemcpyAsync(tmp.p, f, n * es, hipMemcpyHostToDevice, L->stream));
        src = tmp.p;
    }
    int rc;
    if (dt == D3Q19_F64)
        rc = D3Q19_DISPATCH(L, launch_init, L, (const double*)src);
    else
        rc = D3Q19_DISPATCH(L, launch_init, L, (const float*)src);
    if (rc) return rc;
    if (!on_device) HIPCHK(hipStreamSynchronize(L->stream));
    return D3Q19_OK;
}

int d3q19_set_equilibrium(d3q19_lattice* L, const void* rho, const void* u, int32_t dt,
                          int32_t on_device) {
    if (int e = check_lat(L)) return e;
    if (int e = check_dtype(dt)) return e;
    if (!u) return fail(D3Q19_EINVAL, "null velocity");
    if (L->me_ref) return me_ref_set_equilibrium(L, rho, u, dt, on_device);
    if (L->me && !L->me_ranges) return fail(D3Q19_ESTATE, "D3Q19_ME16 lattice needs d3q19_set_moment_ranges first");
    const size_t nc = L->ncell;
    const size_t es = dt == D3Q19_F64 ? 8 : 4;
    DevBuf drho, du, df;
    const void* prho = rho;
    const void* pu = u;
    if (!on_device) {
        if (du.alloc(3 * nc * es) != hipSuccess) return fail(D3Q19_ENOMEM, "staging allocation failed");
        HIPCHK(hipMemcpyAsync(du.p, u, 3 * nc * es, hipMemcpyHostToDevice, L->stream));
        pu = du.p;
        if (rho) {
            if (drho.alloc(nc * es) != hipSuccess) return fail(D3Q19_ENOMEM, "staging allocation failed");
            HIPCHK(hipMemcpyAsync(drho.p, rho, nc * es, hipMemcpyHostToDevice, L->stream));
            prho = drho.p;
        }
    }
    if (!L->me) {
        const int rc = dt == D3Q19_F64
                           ? D3Q19_DISPATCH(L, launch_init_eq, L, (const double*)prho, (const double*)pu)
                           : D3Q19_DISPATCH(L, launch_init_eq, L, (const float*)prho, (const float*)pu);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(L->stream));
        return D3Q19_OK;
    }
    // ME16: populations first, then their codes
    if (df.alloc((size_t)Q * nc * es) != hipSuccess) return fail(D3Q19_ENOMEM, "staging allocation failed");
    if (dt == D3Q19_F64)
        k_fn_feq<double><<<nblocks(nc), BLOCK, 0, L->stream>>>((const double*)prho, (const double*)pu,
                                                               (double*)df.p, (uint32_t)nc);
    else
        k_fn_feq<float><<<nblocks(nc), BLOCK, 0, L->stream>>>((const float*)prho, (const float*)pu,
                                                              (float*)df.p, (uint32_t)nc);
    HIPCHK(hipGetLastError());
    int rc = d3q19_set_populations(L, df.p, dt, 1);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(L->stream));
    return D3Q19_OK;
}

int d3q19_set_force_field(d3q19_lattice* L, const void* force, int32_t dt, int32_t on_device) {
    if (int e = check_lat(L)) return e;
    if (!force) return fail(D3Q19_EINVAL, "null force field");
    const size_t n = 3 * (size_t)L->ncell;
    if (!L->force) {
        if (hipMalloc(&L->force, n * ct_size(L)) != hipSuccess)
            return fail(D3Q19_ENOMEM, "force field allocation failed");
        L->bytes += n * ct_size(L);
    }
    return stage_to_compute(L, force, dt, on_device, n, L->force);
}

int d3q19_set_brinkman(d3q19_lattice* L, const void* s, int32_t dt, int32_t on_device) {
    if (int e = check_lat(L)) return e;
    if (L->me_ref) return fail(D3Q19_EINVAL, "D3Q19_ME16_REF runs the reference's step, which has no Brinkman term");
    if (!s) {
        if (L->brink) {
            HIPCHK(hipStreamSynchronize(L->stream));
            (void)hipFree(L->brink);
            L->brink = nullptr;
        }
        return D3Q19_OK;
    }
    const size_t n = L->ncell;
    if (!L->brink) {
        if (hipMalloc(&L->brink, n * ct_size(L)) != hipSuccess)
            return fail(D3Q19_ENOMEM, "Brinkman field allocation failed");
        L->bytes += n * ct_size(L);
    }
    return stage_to_compute(L, s, dt, on_device, n, L->brink);
}

int d3q19_set_solid(d3q19_lattice* L, const uint8_t* solid, int32_t on_device) {
    if (int e = check_lat(L)) return e;
    if (L->cfg.boundary != D3Q19_BC_HALFWAY) return fail(D3Q19_EINVAL, "solid cells need D3Q19_BC_HALFWAY");
    if (!solid) {
        if (L->solid) {
            HIPCHK(hipStreamSynchronize(L->stream));
            (void)hipFree(L->solid);
            L->solid = nullptr;
        }
        return D3Q19_OK;
    }
    if (!L->solid) {
        if (hipMalloc((void**)&L->solid, L->ncell) != hipSuccess)
            return fail(D3Q19_ENOMEM, "solid mask allocation failed");
        L->bytes += L->ncell;
    }
    HIPCHK(hipMemcpyAsync(L->solid, solid, L->ncell,
                          on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    return D3Q19_OK;
}

int d3q19_step(d3q19_lattice* L, int32_t n) {
    if (int e = check_lat(L)) return e;
    if (!L->has_state) return fail(D3Q19_ESTATE, "lattice has no state; call set_populations/set_equilibrium");
    if (n < 0) return fail(D3Q19_EINVAL, "negative step count");
    if (L->cfg.force_mode == D3Q19_FORCE_FIELD && !L->force)
        return fail(D3Q19_ESTATE, "FORCE_FIELD lattice needs d3q19_set_force_field first");
    if (n == 0) return D3Q19_OK;
    if (L->me_ref) return me_ref_step(L, n);
    return D3Q19_DISPATCH(L, launch_step, L, n, nullptr, nullptr);
}

int d3q19_step_moments(d3q19_lattice* L, int32_t n, float* rho, float* u) {
    if (int e = check_lat(L)) return e;
    if (!L->has_state) return fail(D3Q19_ESTATE, "lattice has no state; call set_populations/set_equilibrium");
    if (n < 1) return fail(D3Q19_EINVAL, "step_moments needs n >= 1");
    if (L->cfg.force_mode == D3Q19_FORCE_FIELD && !L->force)
        return fail(D3Q19_ESTATE, "FORCE_FIELD lattice needs d3q19_set_force_field first");
    if (L->me_ref) return fail(D3Q19_EINVAL, "step_moments is not available for D3Q19_ME16_REF lattices");
    return D3Q19_DISPATCH(L, launch_step, L, n, rho, u);
}

int d3q19_read_populations(d3q19_lattice* L, void* f, int32_t dt, int32_t on_device) {
    if (int e = check_lat(L)) return e;
    if (int e = check_dtype(dt)) return e;
    if (!f) return fail(D3Q19_EINVAL, "null output");
    if (!L->has_state) return fail(D3Q19_ESTATE, "lattice has no state");
    if (L->me_ref) return me_ref_read(L, f, nullptr, nullptr, dt, on_device);
    const size_t n = (size_t)Q * L->ncell;
    const size_t es = dt == D3Q19_F64 ? 8 : 4;
    DevBuf tmp;
    void* dst = f;
    if (!on_device) {
        if (tmp.alloc(n * es) != hipSuccess) return fail(D3Q19_ENOMEM, "staging allocation failed");
        dst = tmp.p;
    }
    int rc = dt == D3Q19_F64
                 ? D3Q19_DISPATCH(L, launch_readback, L, (double*)dst, (double*)nullptr, (double*)nullptr)
                 : D3Q19_DISPATCH(L, launch_readback, L, (float*)dst, (float*)nullptr, (float*)nullptr);
    if (rc) return rc;
    if (!on_device) HIPCHK(hipMemcpyAsync(f, dst, n * es, hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    return ring_check(L);
}

int d3q19_read_moments(d3q19_lattice* L, void* rho, void* u, int32_t dt, int32_t on_device) {
    if (int e = check_lat(L)) return e;
    if (int e = check_dtype(dt)) return e;
    if (!L->has_state) return fail(D3Q19_ESTATE, "lattice has no state");
    if (!rho && !u) return D3Q19_OK;
    if (L->me_ref) return me_ref_read(L, nullptr, rho, u, dt, on_device);
    const size_t nc = L->ncell;
    const size_t es = dt == D3Q19_F64 ? 8 : 4;
    DevBuf trho, tu;
    void* drho = rho;
    void* du = u;
    if (!on_device) {
        if (rho) {
            if (trho.alloc(nc * es) != hipSuccess) return fail(D3Q19_ENOMEM, "staging allocation failed");
            drho = trho.p;
        }
        if (u) {
            if (tu.alloc(3 * nc * es) != hipSuccess) return fail(D3Q19_ENOMEM, "staging allocation failed");
            du = tu.p;
        }
    }
    int rc = dt == D3Q19_F64
                 ? D3Q19_DISPATCH(L, launch_readback, L, (double*)nullptr, (double*)drho, (double*)du)
                 : D3Q19_DISPATCH(L, launch_readback, L, (float*)nullptr, (float*)drho, (float*)du);
    if (rc) return rc;
    if (!on_device) {
        if (rho) HIPCHK(hipMemcpyAsync(rho, drho, nc * es, hipMemcpyDeviceToHost, L->stream));
        if (u) HIPCHK(hipMemcpyAsync(u, du, 3 * nc * es, hipMemcpyDeviceToHost, L->stream));
    }
    HIPCHK(hipStreamSynchronize(L->stream));
    return ring_check(L);
}

int d3q19_read_moments_noalloc(d3q19_lattice* L, void* rho, void* u, int32_t dt) {
    if (int e = check_lat(L)) return e;
    if (int e = check_dtype(dt)) return e;
    if (!L->has_state) return fail(D3Q19_ESTATE, "lattice has no state");
    if (!rho && !u) return D3Q19_OK;
    if (L->me_ref) return fail(D3Q19_EINVAL, "read_moments_noalloc is not available for D3Q19_ME16_REF lattices");
    const size_t nc = L->ncell;
    const size_t es = dt == D3Q19_F64 ? 8 : 4;
    const size_t need = 4 * nc * es;  // rho (ncell) then u (3, ncell)
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(D3Q19_EINVAL, "read_moments_noalloc stages rho and u (" + std::to_string(need) +
                                      " bytes) in a fixed " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                                      "-byte device buffer: lattice too large, use d3q19_read_moments");
    void* base = nullptr;
    if (int e = scratch_on(L->cfg.device, &base)) return e;
    char* drho = (char*)base;
    char* du = drho + nc * es;
    // every staged element is written by the read-back kernel: nothing to clear
    int rc = dt == D3Q19_F64
                 ? D3Q19_DISPATCH(L, launch_readback, L, (double*)nullptr, rho ? (double*)drho : (double*)nullptr,
                                  u ? (double*)du : (double*)nullptr)
                 : D3Q19_DISPATCH(L, launch_readback, L, (float*)nullptr, rho ? (float*)drho : (float*)nullptr,
                                  u ? (float*)du : (float*)nullptr);
    if (rc) return rc;
    if (rho) HIPCHK(hipMemcpyAsync(rho, drho, nc * es, hipMemcpyDeviceToHost, L->stream));
    if (u) HIPCHK(hipMemcpyAsync(u, du, 3 * nc * es, hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    return ring_check(L);
}

int d3q19_diagnostics(d3q19_lattice* L, double out[5]) {
    if (int e = check_lat(L)) return e;
    if (!out) return fail(D3Q19_EINVAL, "null output");
    if (!L->has_state) return fail(D3Q19_ESTATE, "lattice has no state");
    if (L->me_ref) return fail(D3Q19_EINVAL, "diagnostics is not available for D3Q19_ME16_REF lattices");
    int rc = D3Q19_DISPATCH(L, launch_diag, L);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, L->dout, NDIAG * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    return ring_check(L);
}

int d3q19_enstrophy(d3q19_lattice* L, double* out) {
    if (int e = check_lat(L)) return e;
    if (!out) return fail(D3Q19_EINVAL, "null output");
    if (!L->has_state) return fail(D3Q19_ESTATE, "lattice has no state");
    if (L->me_ref) return fail(D3Q19_EINVAL, "enstrophy is not available for D3Q19_ME16_REF lattices");
    // Central differences wrap around every axis: only a fully periodic
    // lattice (no wall face, no solid cell) has a well-defined value here.
    const bool periodic = L->cfg.boundary == D3Q19_BC_PERIODIC ||
                          (L->cfg.boundary == D3Q19_BC_HALFWAY && L->cfg.wall_faces == 0 && !L->solid);
    if (!periodic)
        return fail(D3Q19_EINVAL, "enstrophy uses periodic central differences: the lattice has walls or solids");
    const size_t ubytes = 3 * (size_t)L->ncell * ct_size(L);
    if (!L->uscratch) {
        if (hipMalloc(&L->uscratch, ubytes) != hipSuccess) return fail(D3Q19_ENOMEM, "velocity scratch allocation failed");
        L->bytes += ubytes;
    }
    const uint32_t g = nblocks(L->ncell);
    if (is_ct_double(L)) {
        int rc = D3Q19_DISPATCH(L, launch_readback, L, (double*)nullptr, (double*)nullptr, (double*)L->uscratch);
        if (rc) return rc;
        k_enstrophy<double><<<g, BLOCK, 0, L->stream>>>((const double*)L->uscratch, L->cfg.nx, L->cfg.ny,
                                                        L->cfg.nz, L->partials);
    } else {
        int rc = D3Q19_DISPATCH(L, launch_readback, L, (float*)nullptr, (float*)nullptr, (float*)L->uscratch);
        if (rc) return rc;
        k_enstrophy<float><<<g, BLOCK, 0, L->stream>>>((const float*)L->uscratch, L->cfg.nx, L->cfg.ny,
                                                       L->cfg.nz, L->partials);
    }
    HIPCHK(hipGetLastError());
    k_sum_partials<<<1, BLOCK, 0, L->stream>>>(L->partials, g, 1, L->dout);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, L->dout, sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    return ring_check(L);
}

int d3q19_moment_basis(int32_t m[361], int32_t minv4[361]) {
    if (!m || !minv4) return fail(D3Q19_EINVAL, "null output");
    for (int k = 0; k < Q; ++k)
        for (int i = 0; i < Q; ++i) {
            m[k * Q + i] = me_mono(k, i);
            minv4[i * Q + k] = ME_MINV4[i][k];
        }
    return D3Q19_OK;
}

int d3q19_set_moment_ranges(d3q19_lattice* L, const double lo[19], const double hi[19]) {
    if (int e = check_lat(L)) return e;
    if (!L->me && !L->me_ref) return fail(D3Q19_EINVAL, "moment ranges apply to D3Q19_ME16 and D3Q19_ME16_REF lattices");
    if (!lo || !hi) return fail(D3Q19_EINVAL, "null ranges");
    if (L->has_state) return fail(D3Q19_ESTATE, "set the moment ranges before loading a state");
    for (int k = 0; k < Q; ++k)
        if (!(std::isfinite(lo[k]) && std::isfinite(hi[k]) && hi[k] > lo[k]))
            return fail(D3Q19_EINVAL, "each moment range needs finite lo < hi");
    for (int k = 0; k < Q; ++k) {
        L->me_lo[k] = lo[k];
        L->me_hi[k] = hi[k];
    }
    if (L->mq) {
        const MeQuant q = me_quant(lo, hi);
        HIPCHK(hipMemcpyAsync(L->mq, &q, sizeof(q), hipMemcpyHostToDevice, L->stream));
        HIPCHK(hipStreamSynchronize(L->stream));
    }
    L->me_ranges = true;
    return D3Q19_OK;
}


namespace {
int moment_range(d3q19_lattice* L, double lo[19], double hi[19], bool stored) {
    if (int e = check_lat(L)) return e;
    if (!lo || !hi) return fail(D3Q19_EINVAL, "null output");
    if (!L->has_state) return fail(D3Q19_ESTATE, "lattice has no state");
    if (L->me_ref) return fail(D3Q19_EINVAL, "moment_range is not available for D3Q19_ME16_REF lattices");
    if (!L->mr_partials || !L->mr_out) {
        // one slot pair per block of the population grid (<= the partials of the diagnostics)
        const size_t nb = L->rows ? (size_t)L->cfg.nx * L->cfg.ny * ((L->cfg.nz + 63) / 64) : nblocks(L->ncell);
        const size_t pbytes = std::max<size_t>(nb, nblocks(L->ncell)) * 2 * Q * sizeof(double);
        if (hipMalloc((void**)&L->mr_partials, pbytes) != hipSuccess ||
            hipMalloc((void**)&L->mr_out, 2 * Q * sizeof(double)) != hipSuccess) {
            // leave neither buffer behind, so that the next call allocates both
            if (L->mr_partials) (void)hipFree(L->mr_partials);
            if (L->mr_out) (void)hipFree(L->mr_out);
            L->mr_partials = nullptr;
            L->mr_out = nullptr;
            return fail(D3Q19_ENOMEM, "moment-range buffers");
        }
        L->bytes += pbytes + 2 * Q * sizeof(double);
    }
    if (int rc = D3Q19_DISPATCH(L, launch_moment_range, L, stored)) return rc;
    double out[2 * Q];
    HIPCHK(hipMemcpyAsync(out, L->mr_out, sizeof(out), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    for (int k = 0; k < Q; ++k) {
        lo[k] = out[k];
        hi[k] = out[Q + k];
    }
    return D3Q19_OK;
}
}  // namespace

int d3q19_moment_range(d3q19_lattice* L, double lo[19], double hi[19]) { return moment_range(L, lo, hi, false); }

int d3q19_moment_range_stored(d3q19_lattice* L, double lo[19], double hi[19]) {
    return moment_range(L, lo, hi, true);
}

size_t d3q19_state_bytes(const d3q19_lattice* L) {
    if (!L) return 0;
    return L->me_ref ? L->codes_bytes : L->lat_bytes;
}

int d3q19_read_state(d3q19_lattice* L, void* dst, int32_t* local, int64_t* steps) {
    if (int e = check_lat(L)) return e;
    if (!dst) return fail(D3Q19_EINVAL, "null output");
    if (!L->has_state) return fail(D3Q19_ESTATE, "lattice has no state");
    if (L->ring) {
        // planes 0 .. nx - 1 in order: the ring positions base .. base + nx - 1, wrapped
        if (int e = ring_copy(L, dst, hipMemcpyDeviceToHost)) return e;
    } else {
        const void* src = L->me_ref ? L->buf[0] : L->buf[L->cur];
        HIPCHK(hipMemcpyAsync(dst, src, d3q19_state_bytes(L), hipMemcpyDeviceToHost, L->stream));
    }
    HIPCHK(hipStreamSynchronize(L->stream));
    if (int e = ring_check(L)) return e;
    if (local) *local = L->local ? 1 : 0;
    if (steps) *steps = L->steps;
    return D3Q19_OK;
}

int d3q19_write_state(d3q19_lattice* L, const void* src, int32_t local, int64_t steps) {
    if (int e = check_lat(L)) return e;
    if (!src) return fail(D3Q19_EINVAL, "null state");
    if (steps < 0) return fail(D3Q19_EINVAL, "negative step count");
    if ((L->me || L->me_ref) && !L->me_ranges)
        return fail(D3Q19_ESTATE, "a moment-encoded lattice needs d3q19_set_moment_ranges first");
    if (L->ring) {
        if (int e = ring_copy(L, const_cast<void*>(src), hipMemcpyHostToDevice)) return e;
        if (int e = ring_clear(L)) return e;
    } else {
        void* dst = L->me_ref ? L->buf[0] : L->buf[L->cur];
        HIPCHK(hipMemcpyAsync(dst, src, d3q19_state_bytes(L), hipMemcpyHostToDevice, L->stream));
    }
    HIPCHK(hipStreamSynchronize(L->stream));
    L->has_state = true;
    L->local = L->me_ref ? true : local != 0;
    L->steps = steps;
    return D3Q19_OK;
}

int d3q19_set_step_count(d3q19_lattice* L, int64_t n) {
    if (int e = check_lat(L)) return e;
    if (n < 0) return fail(D3Q19_EINVAL, "negative step count");
    L->steps = n;
    return D3Q19_OK;
}

int d3q19_health(d3q19_lattice* L, double out[4]) {
    if (int e = check_lat(L)) return e;
    if (!out) return fail(D3Q19_EINVAL, "null output");
    if (!L->has_state) return fail(D3Q19_ESTATE, "lattice has no state");
    if (L->me_ref) return fail(D3Q19_EINVAL, "health is not available for D3Q19_ME16_REF lattices");
    if (int rc = D3Q19_DISPATCH(L, launch_health, L)) return rc;
    HIPCHK(hipMemcpyAsync(out, L->dout, NHEALTH * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    return ring_check(L);
}

int d3q19_synchronize(d3q19_lattice* L) {
    if (int e = check_lat(L)) return e;
    HIPCHK(hipStreamSynchronize(L->stream));
    return ring_check(L);
}

int64_t d3q19_step_count(const d3q19_lattice* L) { return L ? L->steps : -1; }

int32_t d3q19_step_launches(const d3q19_lattice* L) {
    if (!L) return -1;
    if (L->me_ref) return 2;
    switch (L->cfg.storage) {
        case D3Q19_F64: return (int32_t)fused_step_chunks<TF64>(L);
        case D3Q19_F32: return (int32_t)fused_step_chunks<TF32>(L);
        case D3Q19_F16: return (int32_t)fused_step_chunks<TF16>(L);
        case D3Q19_BF16: return (int32_t)fused_step_chunks<TBF16>(L);
        case D3Q19_BF16SR: return (int32_t)fused_step_chunks<TBF16S>(L);
        case D3Q19_ME16: return (int32_t)fused_step_chunks<TME16>(L);
    }
    return -1;
}
size_t d3q19_device_bytes(const d3q19_lattice* L) { return L ? L->bytes : 0; }

// ---------------------------------------------------------------------------
// Stateless drop-ins.
// ---------------------------------------------------------------------------
static int check_fn(const void* a, const void* b, int64_t ncell, int32_t dt) {
    if (!a || !b) return fail(D3Q19_EINVAL, "null array");
    if (ncell < 1 || (uint64_t)ncell >= (1ull << 32)) return fail(D3Q19_EINVAL, "bad cell count (1 .. 2^32 - 1)");
    return check_dtype(dt);
}

int d3q19_bgk_step(const void* f, void* out, const void* force, int32_t nx, int32_t ny, int32_t nz,
                   double tau, int32_t dt, void* stream) {
    if (nx < 1 || ny < 1 || nz < 1) return fail(D3Q19_EINVAL, "extents must be >= 1");
    const int64_t nc = (int64_t)nx * ny * nz;
    if (int e = check_fn(f, out, nc, dt)) return e;
    if (f == out) return fail(D3Q19_EINVAL, "bgk_step cannot run in place");
    hipStream_t s = (hipStream_t)stream;
    if (dt == D3Q19_F64)
        k_fn_bgk_step<double><<<nblocks(nc), BLOCK, 0, s>>>((const double*)f, (double*)out,
                                                             (const double*)force, nx, ny, nz, tau);
    else
        k_fn_bgk_step<float><<<nblocks(nc), BLOCK, 0, s>>>((const float*)f, (float*)out,
                                                            (const float*)force, nx, ny, nz, (float)tau);
    HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

int d3q19_stream(const void* f, void* out, int32_t nx, int32_t ny, int32_t nz, int32_t dt, void* stream) {
    if (nx < 1 || ny < 1 || nz < 1) return fail(D3Q19_EINVAL, "extents must be >= 1");
    const int64_t nc = (int64_t)nx * ny * nz;
    if (int e = check_fn(f, out, nc, dt)) return e;
    if (f == out) return fail(D3Q19_EINVAL, "stream cannot run in place");
    hipStream_t s = (hipStream_t)stream;
    if (dt == D3Q19_F64)
        k_fn_stream<double><<<nblocks(nc), BLOCK, 0, s>>>((const double*)f, (double*)out, nx, ny, nz);
    else
        k_fn_stream<float><<<nblocks(nc), BLOCK, 0, s>>>((const float*)f, (float*)out, nx, ny, nz);
    HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

int d3q19_bounce_back_y_walls(const void* f, void* out, const double top[3], const double bot[3],
                              int32_t nx, int32_t ny, int32_t nz, int32_t dt, void* stream) {
    if (nx < 1 || ny < 1 || nz < 1) return fail(D3Q19_EINVAL, "extents must be >= 1");
    const int64_t nc = (int64_t)nx * ny * nz;
    if (int e = check_fn(f, out, nc, dt)) return e;
    if (f == out) return fail(D3Q19_EINVAL, "bounce-back cannot run in place");
    const double zero[3] = {0, 0, 0};
    if (!top) top = zero;
    if (!bot) bot = zero;
    hipStream_t s = (hipStream_t)stream;
    if (dt == D3Q19_F64)
        k_fn_bb_ywalls<double><<<nblocks(nc), BLOCK, 0, s>>>((const double*)f, (double*)out, nx, ny, nz,
                                                              top[0], top[1], top[2], bot[0], bot[1], bot[2]);
    else
        k_fn_bb_ywalls<float><<<nblocks(nc), BLOCK, 0, s>>>(
            (const float*)f, (float*)out, nx, ny, nz, (float)top[0], (float)top[1], (float)top[2],
            (float)bot[0], (float)bot[1], (float)bot[2]);
    HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

int d3q19_feq_field(const void* rho, const void* u, void* out, int64_t nc, int32_t dt, void* stream) {
    if (int e = check_fn(u, out, nc, dt)) return e;
    hipStream_t s = (hipStream_t)stream;
    if (dt == D3Q19_F64)
        k_fn_feq<double><<<nblocks(nc), BLOCK, 0, s>>>((const double*)rho, (const double*)u, (double*)out,
                                                        (uint32_t)nc);
    else
        k_fn_feq<float><<<nblocks(nc), BLOCK, 0, s>>>((const float*)rho, (const float*)u, (float*)out,
                                                       (uint32_t)nc);
    HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

int d3q19_feq(const void* rho, const void* u, void* out, int64_t nc, int32_t dt, void* stream) {
    if (int e = check_fn(u, out, nc, dt)) return e;
    hipStream_t s = (hipStream_t)stream;
    if (dt == D3Q19_F64)
        k_fn_feq<double, true><<<nblocks(nc), BLOCK, 0, s>>>((const double*)rho, (const double*)u, (double*)out,
                                                        (uint32_t)nc);
    else
        k_fn_feq<float, true><<<nblocks(nc), BLOCK, 0, s>>>((const float*)rho, (const float*)u, (float*)out,
                                                       (uint32_t)nc);
    HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

static int fn_moments(const void* f, const void* force, void* rho, void* mom, void* u, int64_t nc,
                      int32_t dt, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (dt == D3Q19_F64)
        k_fn_moments<double><<<nblocks(nc), BLOCK, 0, s>>>((const double*)f, (const double*)force,
                                                            (double*)rho, (double*)mom, (double*)u,
                                                            (uint32_t)nc);
    else
        k_fn_moments<float><<<nblocks(nc), BLOCK, 0, s>>>((const float*)f, (const float*)force,
                                                           (float*)rho, (float*)mom, (float*)u, (uint32_t)nc);
    HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

int d3q19_density_field(const void* f, void* rho, int64_t nc, int32_t dt, void* stream) {
    if (int e = check_fn(f, rho, nc, dt)) return e;
    return fn_moments(f, nullptr, rho, nullptr, nullptr, nc, dt, stream);
}

int d3q19_momentum_field(const void* f, void* mom, int64_t nc, int32_t dt, void* stream) {
    if (int e = check_fn(f, mom, nc, dt)) return e;
    return fn_moments(f, nullptr, nullptr, mom, nullptr, nc, dt, stream);
}

int d3q19_macroscopic_velocity(const void* f, const void* force, void* u, int64_t nc, int32_t dt,
                               void* stream) {
    if (int e = check_fn(f, u, nc, dt)) return e;
    return fn_moments(f, force, nullptr, nullptr, u, nc, dt, stream);
}

}  // extern "C"
