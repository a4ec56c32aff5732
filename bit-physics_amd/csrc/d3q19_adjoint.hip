// d3q19_adjoint.hip -- discrete adjoint of the D3Q19 BGK + Guo (+ Brinkman) step.
//
// Forward map (per-cell force F, optional Brinkman solid fraction s):
//   f^0 = feq(1, u0),  f^{n+1} = S C(f^n; F - s m),  u^T = (m + (F - s m)/2) / rho  at f^T
// with C the BGK+Guo collision and S the streaming with its boundary:
// periodic, half-way bounce-back on wall faces (a permutation plus the
// constant moving-wall term 6 w_i c_i.u_face), or the reference y-walls
// S = B P (periodic streaming P, then the wall map B on rows 0 and ny - 1,
// affine in the streamed values: reference/bgk.py:190-215).  The engine
// stores post-collision states, so the forward runs on
//   H_0 = S^-1 f^0,  H_{j+1} = k_step(H_j),  f^j = S H_j.
// The y-wall map discards the populations it replaces, so S has no inverse
// there: H_0 is f^0 itself (local layout), H_1 = C(f^0) by a local collision,
// and f^0 is read back from H_0 directly.
// Backward (lambda^n = dL/df^n):
//   lambda^T   = J_u(f^T)^T dL/du^T
//   lambda^n   = J_C(f^n)^T S^T lambda^{n+1},
//                (S^T mu)_i(x) = mu_i(x + c_i), or mu_opp(i)(x) when x + c_i is across a wall;
//                y-walls: S^T = P^T B^T, B^T applied on the wall rows first
//   dL/dF, dL/ds accumulate the parameter parts of each step's Jacobian
//   dL/du0     = dfeq(1,u0)/du0^T lambda^0
// The states between checkpoints (every K steps) are recomputed by the
// forward kernel itself, which also stores rho - 1 and the momentum sum m of
// each pulled f^n (k_step with MOM): J_C(f^n) depends on f^n only through
// them, so the reverse step reads 4 values per cell instead of pulling 19,
// and a segment keeps K moment slots instead of K - 1 states.  Checkpoints
// may be stored as FP16 (f - w) and expanded to the working precision on use.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/d3q19.h"
#include "d3q19_host.hpp"
#include "d3q19_kernels.hpp"

using namespace d3q19;

namespace {

// True when the pull source x - c of a cell crosses a wall face (half-way mode).
template <class C>
__device__ __forceinline__ bool crosses_wall(int cx, int cy, int cz, uint32_t x, uint32_t y, uint32_t z,
                                             const Params<C>& p) {
    const uint32_t wm = p.wall_mask;
    return (cx > 0 && x == 0 && (wm & 1u)) || (cx < 0 && x + 1 == p.nx && (wm & 2u)) ||
           (cy > 0 && y == 0 && (wm & 4u)) || (cy < 0 && y + 1 == p.ny && (wm & 8u)) ||
           (cz > 0 && z == 0 && (wm & 16u)) || (cz < 0 && z + 1 == p.nz && (wm & 32u));
}

template <class C>
__device__ __forceinline__ C feq_dev(int i, const C* u0, uint32_t N, uint32_t c) {
    const C ux = u0[c], uy = u0[(size_t)N + c], uz = u0[2 * (size_t)N + c];
    const C cu = C(CX[i]) * ux + C(CY[i]) * uy + C(CZ[i]) * uz;
    return wgt<C>(i) * (C(3) * cu + C(4.5) * cu * cu - C(1.5) * (ux * ux + uy * uy + uz * uz));
}

// Initial post-collision pre-image H_0 = S^-1 feq(1, u0):
//   H_0,i(x) = feq_i(1, u0(x + c_i)) - w_i, or, when x + c_i is across a wall,
//   feq_opp(i)(1, u0(x)) - w_i + 6 w_i (c_i . u_face), u_face the velocity of the
//   face that the pull of direction opp(i) at x reflects from (the pull adds
//   6 w_opp (c_opp . u_face), which cancels it).
// Reference y-walls: H_0 = f^0 - w in local layout (S has no inverse).
template <class Tag, int BC>
__global__ void __launch_bounds__(step_block<Tag>())
    k_adj_unstream(const typename Codec<Tag>::C* __restrict__ u0, typename Codec<Tag>::raw* __restrict__ out,
                   Params<typename Codec<Tag>::C> p) {
    typedef typename Codec<Tag>::C C;
    uint32_t x, y, z;
    if (!grid_cell(p.nx, p.ny, p.nz, x, y, z)) return;
    const uint32_t N = p.ncell;
    const uint32_t self = (x * p.ny + y) * p.nz + z;
    C d[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int face = BC == BC_HALFWAY ? halfway_face(OPP[i], p.wall_mask, x, y, z, p) : -1;
        if (BC == BC_YWALLS) {
            d[i] = feq_dev<C>(i, u0, N, self);
        } else if (face >= 0) {
            const C cu = cdot<C>(i, p.face_u[face][0], p.face_u[face][1], p.face_u[face][2]);
            d[i] = feq_dev<C>(OPP[i], u0, N, self) + C(6) * wgt<C>(i) * cu;
        } else {
            const uint32_t sx = (x + p.nx + CX[i]) % p.nx, sy = (y + p.ny + CY[i]) % p.ny,
                           sz = (z + p.nz + CZ[i]) % p.nz;
            d[i] = feq_dev<C>(i, u0, N, (sx * p.ny + sy) * p.nz + sz);
        }
    }
    store_local<Tag>(out, p, x, y, z, d);
}

// Seed at f^T: u = ((1 - s/2) m + F/2)/rho
//   lambda^T_j = (g.c_j (1 - s/2) - g.u) / rho ;  gF = g / (2 rho) ;  gs = -(g.m) / (2 rho)
template <class Tag, int BC, bool BRINK>
__global__ void __launch_bounds__(step_block<Tag>())
    k_adj_seed(const typename Codec<Tag>::raw* __restrict__ H, Params<typename Codec<Tag>::C> p,
               const typename Codec<Tag>::C* __restrict__ g, typename Codec<Tag>::raw* __restrict__ lam,
               typename Codec<Tag>::C* __restrict__ gF, typename Codec<Tag>::C* __restrict__ gs) {
    typedef typename Codec<Tag>::C C;
    uint32_t x, y, z;
    if (!grid_cell(p.nx, p.ny, p.nz, x, y, z)) return;
    const uint32_t cell = (x * p.ny + y) * p.nz + z;
    C d[Q];
    pull<Tag, BC, true>(H, p, x, y, z, d);
    C drho, mx, my, mz;
    moments(d, drho, mx, my, mz);
    const C rho = C(1) + drho;
    const C ir = C(1) / rho;
    const uint32_t N = p.ncell;
    const C s = BRINK ? p.brink[cell] : C(0);
    const C fx = p.force[cell] - s * mx, fy = p.force[(size_t)N + cell] - s * my, fz = p.force[2 * (size_t)N + cell] - s * mz;
    const C ux = (mx + C(0.5) * fx) * ir, uy = (my + C(0.5) * fy) * ir, uz = (mz + C(0.5) * fz) * ir;
    const C gx = g[cell], gy = g[(size_t)N + cell], gz = g[2 * (size_t)N + cell];
    const C gu = gx * ux + gy * uy + gz * uz;
    const C keep = C(1) - C(0.5) * s;
    C l[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) l[j] = ((C(CX[j]) * gx + C(CY[j]) * gy + C(CZ[j]) * gz) * keep - gu) * ir;
    store_local<Tag>(lam, p, x, y, z, l);
    gF[cell] = C(0.5) * gx * ir;
    gF[(size_t)N + cell] = C(0.5) * gy * ir;
    gF[2 * (size_t)N + cell] = C(0.5) * gz * ir;
    if (BRINK) gs[cell] = -C(0.5) * (gx * mx + gy * my + gz * mz) * ir;
}

// One reverse step: lambda_out = J_C(f^n)^T S^T lambda_in; the parameter
// parts go to gF (force) and gs (Brinkman fraction).  mom: rho - 1 and m of
// f^n, (ncell, 4), stored by the recomputation (store_adjoint_moments).
template <class Tag, int BC, bool BRINK>
__global__ void __launch_bounds__(step_block<Tag>())
    k_adj_step(const typename Codec<Tag>::C* __restrict__ mom, Params<typename Codec<Tag>::C> p,
               const typename Codec<Tag>::raw* __restrict__ lam_in, typename Codec<Tag>::raw* __restrict__ lam_out,
               typename Codec<Tag>::C* __restrict__ gF, typename Codec<Tag>::C* __restrict__ gs) {
    typedef typename Codec<Tag>::C C;
    constexpr uint32_t ES = sizeof(typename Codec<Tag>::raw);
    uint32_t x, y, z;
    if (!grid_cell(p.nx, p.ny, p.nz, x, y, z)) return;
    const uint32_t cell = (x * p.ny + y) * p.nz + z;
    const uint32_t N = p.ncell;
    // S^T: mu_i(x) = lambda_i(x + c_i)  (the pull run backwards; bounce links reversed)
    const uint32_t xs[3] = {x == 0 ? p.nx - 1 : x - 1, x, x + 1 == p.nx ? 0 : x + 1};
    const uint32_t ys[3] = {y == 0 ? p.ny - 1 : y - 1, y, y + 1 == p.ny ? 0 : y + 1};
    const uint32_t zs[3] = {z == 0 ? p.nz - 1 : z - 1, z, z + 1 == p.nz ? 0 : z + 1};
    const char* self_row = row_of(lam_in, p, x, y);
    C mu[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        // one load per direction from a selected address (no load inside a
        // branch arm, whose join would wait for every load in flight); the
        // nontemporal hint on the c_z = 0 directions only, as the step's pull
        const bool back = BC == BC_HALFWAY && crosses_wall(-CX[i], -CY[i], -CZ[i], x, y, z, p);
        const char* src = back ? self_row + (uint32_t)OPP[i] * p.dir_bytes + z * ES
                               : row_of(lam_in, p, xs[1 + CX[i]], ys[1 + CY[i]]) + (uint32_t)i * p.dir_bytes +
                                     zs[1 + CZ[i]] * ES;
        mu[i] = CZ[i] == 0 ? ld_at<Tag, true>(src, i) : ld_at<Tag, false>(src, i);
    }
    const Vec4<C> m4 = __builtin_nontemporal_load((const Vec4<C>*)mom + cell);
    const C drho = m4.x, mx = m4.y, my = m4.z, mz = m4.w;
    const C rho = C(1) + drho;
    const C ir = C(1) / rho;
    const C s = BRINK ? p.brink[cell] : C(0);
    const C m[3] = {mx, my, mz};
    const C F[3] = {p.force[cell] - s * mx, p.force[(size_t)N + cell] - s * my, p.force[2 * (size_t)N + cell] - s * mz};
    const C u[3] = {(mx + C(0.5) * F[0]) * ir, (my + C(0.5) * F[1]) * ir, (mz + C(0.5) * F[2]) * ir};
    const C usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const C om = p.omega, pref = p.pref;
    C a_rho = C(0);
    C b[3] = {C(0), C(0), C(0)};
    C ge[3] = {C(0), C(0), C(0)};
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const C c[3] = {C(CX[i]), C(CY[i]), C(CZ[i])};
        const C sw = mu[i] * wgt<C>(i);
        const C cu = c[0] * u[0] + c[1] * u[1] + c[2] * u[2];
        const C cF = c[0] * F[0] + c[1] * F[1] + c[2] * F[2];
        a_rho += sw * (C(1) + C(3) * cu + C(4.5) * cu * cu - C(1.5) * usq);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            // d(feq_i/tau)/du_a = w rho (3 c_a + 9 cu c_a - 3 u_a)/tau ; dG/du_a = pref w (-3 F_a + 9 c_a cF)
            b[a] += sw * (om * rho * (C(3) * c[a] + C(9) * cu * c[a] - C(3) * u[a]) +
                          pref * (C(9) * c[a] * cF - C(3) * F[a]));
            // explicit dG/dF_a = pref w (3 (c_a - u_a) + 9 cu c_a)
            ge[a] += sw * pref * (C(3) * (c[a] - u[a]) + C(9) * cu * c[a]);
        }
    }
    a_rho *= om;
    // total derivative through the effective force (explicit + via u = (m + F/2)/rho)
    const C gt[3] = {ge[0] + C(0.5) * b[0] * ir, ge[1] + C(0.5) * b[1] * ir, ge[2] + C(0.5) * b[2] * ir};
    const C keep = C(1) - om;
    C l[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) {
        C v = keep * mu[j] + a_rho +
              (b[0] * (C(CX[j]) - u[0]) + b[1] * (C(CY[j]) - u[1]) + b[2] * (C(CZ[j]) - u[2])) * ir;
        if (BRINK) v -= s * (C(CX[j]) * gt[0] + C(CY[j]) * gt[1] + C(CZ[j]) * gt[2]);  // F = F0 - s m
        l[j] = v;
    }
    store_local<Tag>(lam_out, p, x, y, z, l);
#pragma unroll
    for (int a = 0; a < 3; ++a) gF[(size_t)a * N + cell] += gt[a];
    if (BRINK) gs[cell] -= gt[0] * m[0] + gt[1] * m[1] + gt[2] * m[2];
}

// Reference y-walls, S = B P: on row 0 (c_y > 0) and row ny - 1 (c_y < 0) the
// streamed values d of a cell become
//   (B d)_i = d_opp(i) - a_i (1 + sum_j d_j)   for the replaced directions i,
//   (B d)_i = d_i                              for the others,
// a_i = 6 w_i (c_i . u_wall), the density taken over all 19 streamed values
// (ywall_swap).  S^T = P^T B^T; this applies B^T in place to an adjoint state
// on the wall rows:
//   (B^T l)_j = [j kept] l_j + [opp(j) replaced] l_opp(j) - sum_{i replaced} a_i l_i.
// One thread per wall cell: nx * rows * nz threads, rows = 2 (1 when ny = 1,
// where both walls act on the one row).
template <class Tag>
__global__ void __launch_bounds__(256)
    k_adj_ywalls_T(typename Codec<Tag>::raw* __restrict__ lam, Params<typename Codec<Tag>::C> p) {
    typedef typename Codec<Tag>::C C;
    const uint32_t rows = p.ny > 1 ? 2u : 1u;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)p.nx * rows * p.nz) return;
    const uint32_t z = (uint32_t)(t % p.nz);
    const uint32_t r = (uint32_t)((t / p.nz) % rows);
    const uint32_t x = (uint32_t)(t / ((uint64_t)p.nz * rows));
    const uint32_t y = r == 0 ? 0u : p.ny - 1;
    const bool bot = (y == 0), top = (y + 1 == p.ny);
    C l[Q];
    load_local<Tag>(lam, p, x, y, z, l);
    bool rep[Q];
    C acc = C(0);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        rep[i] = (bot && CY[i] > 0) || (top && CY[i] < 0);
        if (rep[i]) {
            const C cu = CY[i] > 0 ? cdot<C>(i, p.wbot[0], p.wbot[1], p.wbot[2])
                                   : cdot<C>(i, p.wtop[0], p.wtop[1], p.wtop[2]);
            acc += C(6) * wgt<C>(i) * cu * l[i];
        }
    }
    C o[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) o[j] = (rep[j] ? C(0) : l[j]) + (rep[OPP[j]] ? l[OPP[j]] : C(0)) - acc;
    store_local<Tag>(lam, p, x, y, z, o);
}

// dL/du0_a = sum_i lambda^0_i w_i (3 c_a + 9 cu c_a - 3 u_a)   (rho = 1)
template <class Tag>
__global__ void __launch_bounds__(step_block<Tag>())
    k_adj_u0(const typename Codec<Tag>::raw* __restrict__ lam, Params<typename Codec<Tag>::C> p,
             const typename Codec<Tag>::C* __restrict__ u0, typename Codec<Tag>::C* __restrict__ gu0) {
    typedef typename Codec<Tag>::C C;
    uint32_t x, y, z;
    if (!grid_cell(p.nx, p.ny, p.nz, x, y, z)) return;
    const uint32_t cell = (x * p.ny + y) * p.nz + z;
    const uint32_t N = p.ncell;
    C l[Q];
    load_local<Tag>(lam, p, x, y, z, l);
    const C u[3] = {u0[cell], u0[(size_t)N + cell], u0[2 * (size_t)N + cell]};
    C g[3] = {C(0), C(0), C(0)};
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const C c[3] = {C(CX[i]), C(CY[i]), C(CZ[i])};
        const C sw = l[i] * wgt<C>(i);
        const C cu = c[0] * u[0] + c[1] * u[1] + c[2] * u[2];
#pragma unroll
        for (int a = 0; a < 3; ++a) g[a] += sw * (C(3) * c[a] + C(9) * cu * c[a] - C(3) * u[a]);
    }
    gu0[cell] = g[0];
    gu0[(size_t)N + cell] = g[1];
    gu0[2 * (size_t)N + cell] = g[2];
}

// Re-encode a state between storage formats (checkpoint compression).
template <class TIn, class TOut>
__global__ void __launch_bounds__(256)
    k_convert_state(const typename Codec<TIn>::raw* __restrict__ in, Params<typename Codec<TIn>::C> pin,
                    typename Codec<TOut>::raw* __restrict__ out, Params<typename Codec<TOut>::C> pout) {
    uint32_t x, y, z;
    if (!grid_cell(pin.nx, pin.ny, pin.nz, x, y, z)) return;
    typename Codec<TIn>::C a[Q];
    typename Codec<TOut>::C b[Q];
    load_local<TIn>(in, pin, x, y, z, a);
#pragma unroll
    for (int i = 0; i < Q; ++i) b[i] = (typename Codec<TOut>::C)a[i];
    store_local<TOut>(out, pout, x, y, z, b);
}

}  // namespace

struct d3q19_adjoint {
    d3q19_config cfg;
    d3q19_adjoint_options opt;
    int32_t T = 0, K = 0, nck = 0;
    bool ckpt_same = true;  // checkpoints in the working storage
    bool brink = false;
    hipStream_t stream = nullptr;  // nullptr: the default stream
    uint32_t ncell = 0, nzp = 0;
    size_t es = 0, state_bytes = 0, ckpt_bytes = 0;
    std::vector<void*> ckpt;  // H_{kK}
    std::vector<void*> seg;   // working states: the forward's and the recomputation's ping-pong
    void* moms = nullptr;     // K moment slots of a segment, mom_bytes each (store_adjoint_moments)
    size_t mom_bytes = 0;
    void* lam[2] = {nullptr, nullptr};
    void* u0 = nullptr;
    void* force = nullptr;
    void* s = nullptr;    // Brinkman fraction
    void* gF = nullptr;
    void* gu = nullptr;   // staged dL/du^T, reused for dL/du0
    void* gs = nullptr;   // dL/ds
    void* Hfinal = nullptr;
    bool have_forward = false;
    size_t bytes = 0;
    uint32_t lds_per_cu = 0;  // LDS bytes per CU of the device, 0 if unknown
};

namespace {

bool ct_double(const d3q19_adjoint* A) { return A->cfg.storage == D3Q19_F64; }

template <class C>
Params<C> adj_params(const d3q19_adjoint* A, size_t es) {
    Params<C> p;
    std::memset(&p, 0, sizeof(p));
    p.nx = A->cfg.nx;
    p.ny = A->cfg.ny;
    p.nz = A->cfg.nz;
    p.ncell = A->ncell;
    p.nzp = A->nzp;
    p.dir_bytes = (uint32_t)(A->nzp * es);
    p.row_bytes = (uint64_t)Q * A->nzp * es;
    set_collision_constants(p, A->cfg.tau);
    p.force_mode = FORCE_FIELD;
    p.force = (const C*)A->force;
    p.brink = A->brink ? (const C*)A->s : nullptr;
    p.wall_mask = A->cfg.boundary == D3Q19_BC_HALFWAY ? A->cfg.wall_faces : 0u;
    for (int a = 0; a < 3; ++a) {
        p.wtop[a] = (C)A->cfg.wall_velocity_top[a];
        p.wbot[a] = (C)A->cfg.wall_velocity_bottom[a];
        for (int k = 0; k < 6; ++k) p.face_u[k][a] = (C)A->cfg.face_velocity[k][a];
    }
    set_zwall_terms(p);
    return p;
}

// (z, y, x) launch shape shared by every kernel here.
template <class Tag>
void adj_grid(const d3q19_adjoint* A, dim3& grid, dim3& block, uint32_t threads = step_block<Tag>()) {
    const uint32_t nz = A->cfg.nz;
    const uint32_t bz = row_block(nz);
    const uint32_t by = threads / bz;
    block = dim3(bz, by, 1);
    grid = dim3((nz + bz - 1) / bz, (A->cfg.ny + by - 1) / by, A->cfg.nx);
}

template <class Tag, bool MOM>
int step_into_t(d3q19_adjoint* A, void* in, void* out, bool first, void* mom) {
    typedef typename Codec<Tag>::C C;
    typedef typename Codec<Tag>::raw R;
    Params<C> p = adj_params<C>(A, A->es);
    p.mom = (C*)mom;
    dim3 g, b, gf, bf;
    adj_grid<Tag>(A, g, b);                       // k_collide_local
    adj_grid<Tag>(A, gf, bf, fused_block<Tag>());  // k_step
    const R* src = (const R*)in;
    // the forward step takes the lattice's resident-block cap (the backward
    // step's registers already hold it to 6 (FP32) and 3 (FP64) blocks per CU,
    // and lower caps measured slower)
    const uint32_t lds = reserve_lds(A->lds_per_cu, step_blocks_per_cu<Tag>());
    switch (A->cfg.boundary) {
        case D3Q19_BC_HALFWAY:
            k_step<Tag, BC_HALFWAY, true, true, MOM><<<gf, bf, lds, A->stream>>>(src, (R*)out, p, nullptr, nullptr);
            break;
        case D3Q19_BC_YWALLS:
            if (first)
                k_collide_local<Tag, true, MOM><<<g, b, 0, A->stream>>>(src, (R*)out, p, nullptr, nullptr);
            else
                k_step<Tag, BC_YWALLS, true, true, MOM><<<gf, bf, lds, A->stream>>>(src, (R*)out, p, nullptr, nullptr);
            break;
        default:
            k_step<Tag, BC_PERIODIC, true, true, MOM><<<gf, bf, lds, A->stream>>>(src, (R*)out, p, nullptr, nullptr);
    }
    D3Q19_HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

// One forward step H_{j+1} = C(S H_j).  first: j = 0 of a y-wall chain,
// whose H_0 is f^0 itself, so the step is the local collision H_1 = C(f^0).
// mom: where the step also stores the moments of f^j (the recomputation), or null.
template <class Tag>
int step_into(d3q19_adjoint* A, void* in, void* out, bool first, void* mom = nullptr) {
    if (mom) return step_into_t<Tag, true>(A, in, out, first, mom);
    return step_into_t<Tag, false>(A, in, out, first, nullptr);
}


// Working state <-> FP16 checkpoint.
template <class Tag>
int convert(d3q19_adjoint* A, const void* in, void* out, bool to_ckpt) {
    typedef typename Codec<Tag>::C C;
    dim3 g, b;
    adj_grid<Tag>(A, g, b, 256);
    const Params<C> pw = adj_params<C>(A, A->es);
    const Params<float> ph = adj_params<float>(A, 2);
    if (to_ckpt)
        k_convert_state<Tag, TF16><<<g, b, 0, A->stream>>>((const typename Codec<Tag>::raw*)in, pw,
                                                           (uint16_t*)out, ph);
    else
        k_convert_state<TF16, Tag><<<g, b, 0, A->stream>>>((const uint16_t*)in, ph,
                                                           (typename Codec<Tag>::raw*)out, pw);
    D3Q19_HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

template <class Tag, int BC>
int run_forward_t(d3q19_adjoint* A, void* u_out_dev, int out_dt) {
    typedef typename Codec<Tag>::C C;
    typedef typename Codec<Tag>::raw R;
    const Params<C> p = adj_params<C>(A, A->es);
    dim3 g, b;
    adj_grid<Tag>(A, g, b);
    void* h0 = A->ckpt_same ? A->ckpt[0] : A->seg[0];
    k_adj_unstream<Tag, BC><<<g, b, 0, A->stream>>>((const C*)A->u0, (R*)h0, p);
    D3Q19_HIPCHK(hipGetLastError());
    if (!A->ckpt_same)
        if (int e = convert<Tag>(A, h0, A->ckpt[0], true)) return e;
    void* cur = h0;
    for (int32_t j = 0; j < A->T; ++j) {
        const int32_t nxt = j + 1;
        const bool is_ck = (nxt % A->K == 0) && (nxt / A->K < A->nck);
        void* out;
        if (is_ck && A->ckpt_same)
            out = A->ckpt[nxt / A->K];
        else
            out = (cur == A->seg[0]) ? A->seg[1] : A->seg[0];
        if (int e = step_into<Tag>(A, cur, out, j == 0)) return e;
        if (is_ck && !A->ckpt_same)
            if (int e = convert<Tag>(A, out, A->ckpt[nxt / A->K], true)) return e;
        cur = out;
    }
    A->Hfinal = cur;
    if (out_dt == D3Q19_F64)
        k_readback<Tag, BC, double><<<g, b, 0, A->stream>>>((const R*)cur, p, nullptr, nullptr, (double*)u_out_dev);
    else
        k_readback<Tag, BC, float><<<g, b, 0, A->stream>>>((const R*)cur, p, nullptr, nullptr, (float*)u_out_dev);
    D3Q19_HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

template <class Tag>
int run_forward(d3q19_adjoint* A, void* u_out_dev, int out_dt) {
    switch (A->cfg.boundary) {
        case D3Q19_BC_HALFWAY: return run_forward_t<Tag, BC_HALFWAY>(A, u_out_dev, out_dt);
        case D3Q19_BC_YWALLS: return run_forward_t<Tag, BC_YWALLS>(A, u_out_dev, out_dt);
        default: return run_forward_t<Tag, BC_PERIODIC>(A, u_out_dev, out_dt);
    }
}

// Moment slot j of a segment (rho - 1 and m of f^{n0 + j}, (ncell, 4)).
void* mom_slot(const d3q19_adjoint* A, int32_t j) { return (char*)A->moms + (size_t)j * A->mom_bytes; }

template <class Tag, int BC, bool BRINK>
int run_backward_t(d3q19_adjoint* A) {
    typedef typename Codec<Tag>::C C;
    typedef typename Codec<Tag>::raw R;
    const Params<C> p = adj_params<C>(A, A->es);
    dim3 g, b;
    adj_grid<Tag>(A, g, b);
    k_adj_seed<Tag, BC, BRINK><<<g, b, 0, A->stream>>>((const R*)A->Hfinal, p, (const C*)A->gu, (R*)A->lam[0],
                                                        (C*)A->gF, (C*)A->gs);
    D3Q19_HIPCHK(hipGetLastError());
    int cur = 0;
    for (int32_t k = A->nck - 1; k >= 0; --k) {
        const int32_t n0 = k * A->K;
        const int32_t len = std::min(A->K, A->T - n0);
        // recompute H_{n0 + 1} .. H_{n0 + len} from the checkpoint through the
        // two working states, storing the moments of f^{n0 + j} in slot j
        const void* h = A->ckpt[k];
        if (!A->ckpt_same) {
            if (int e = convert<Tag>(A, A->ckpt[k], A->seg[0], false)) return e;
            h = A->seg[0];
        }
        for (int32_t j = 0; j < len; ++j) {
            void* out = (h == A->seg[0]) ? A->seg[1] : A->seg[0];
            if (int e = step_into<Tag>(A, const_cast<void*>(h), out, k == 0 && j == 0, mom_slot(A, j))) return e;
            h = out;
        }
        for (int32_t j = len - 1; j >= 0; --j) {
            if constexpr (BC == BC_YWALLS) {
                const uint32_t n = p.nx * (p.ny > 1 ? 2u : 1u) * p.nz;
                k_adj_ywalls_T<Tag><<<(n + 255) / 256, 256, 0, A->stream>>>((R*)A->lam[cur], p);
                D3Q19_HIPCHK(hipGetLastError());
            }
            k_adj_step<Tag, BC, BRINK><<<g, b, 0, A->stream>>>((const C*)mom_slot(A, j), p, (const R*)A->lam[cur],
                                                                (R*)A->lam[cur ^ 1], (C*)A->gF, (C*)A->gs);
            D3Q19_HIPCHK(hipGetLastError());
            cur ^= 1;
        }
    }
    k_adj_u0<Tag><<<g, b, 0, A->stream>>>((const R*)A->lam[cur], p, (const C*)A->u0, (C*)A->gu);
    D3Q19_HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

template <class Tag>
int run_backward(d3q19_adjoint* A) {
    switch (A->cfg.boundary) {
        case D3Q19_BC_HALFWAY:
            return A->brink ? run_backward_t<Tag, BC_HALFWAY, true>(A) : run_backward_t<Tag, BC_HALFWAY, false>(A);
        case D3Q19_BC_YWALLS:
            return A->brink ? run_backward_t<Tag, BC_YWALLS, true>(A) : run_backward_t<Tag, BC_YWALLS, false>(A);
        default:
            return A->brink ? run_backward_t<Tag, BC_PERIODIC, true>(A) : run_backward_t<Tag, BC_PERIODIC, false>(A);
    }
}

void free_all(d3q19_adjoint* A) {
    for (void* p : A->ckpt)
        if (p) (void)hipFree(p);
    for (void* p : A->seg)
        if (p) (void)hipFree(p);
    void* ps[] = {A->moms, A->lam[0], A->lam[1], A->u0, A->force, A->s, A->gF, A->gu, A->gs};
    for (void* p : ps)
        if (p) (void)hipFree(p);
}

// Per-cell fields of the adjoint: u0, force, dL/dF, dL/du (3 each), and the
// Brinkman fraction with its gradient.
size_t field_bytes(const d3q19_adjoint* A) {
    const size_t vec = 3 * (size_t)A->ncell * A->es;
    const size_t scal = (size_t)A->ncell * A->es;
    return 4 * vec + (A->brink ? 2 * scal : 0);
}

// Device bytes of the adjoint with checkpoint interval K (adjoint_plan_bytes).
size_t plan_bytes(const d3q19_adjoint* A, int32_t K) {
    return adjoint_plan_bytes(A->T, K, A->ckpt_bytes, A->mom_bytes, A->state_bytes, field_bytes(A));
}

}  // namespace

extern "C" {

void d3q19_adjoint_options_default(d3q19_adjoint_options* opt) {
    if (!opt) return;
    std::memset(opt, 0, sizeof(*opt));
    opt->steps = 1;
    opt->checkpoint_every = 0;
    opt->checkpoint_storage = -1;  // the lattice storage
    opt->brinkman = 0;
    opt->memory_budget = 0;
}

int d3q19_adjoint_create(const d3q19_config* cfg, const d3q19_adjoint_options* opt, void* stream,
                         d3q19_adjoint** out) {
    if (!cfg || !opt || !out) return set_error(D3Q19_EINVAL, "null argument");
    *out = nullptr;
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return set_error(D3Q19_EINVAL, "extents must be >= 1");
    if (cfg->boundary != D3Q19_BC_PERIODIC && cfg->boundary != D3Q19_BC_HALFWAY &&
        cfg->boundary != D3Q19_BC_YWALLS)
        return set_error(D3Q19_EINVAL, "unknown boundary");
    if (cfg->boundary == D3Q19_BC_HALFWAY && (cfg->wall_faces & ~63u))
        return set_error(D3Q19_EINVAL, "wall_faces has bits beyond the six faces");
    if (cfg->force_mode != D3Q19_FORCE_FIELD) return set_error(D3Q19_EINVAL, "adjoint needs force_mode FIELD");
    if (cfg->storage != D3Q19_F32 && cfg->storage != D3Q19_F64)
        return set_error(D3Q19_EINVAL, "adjoint storage must be F32 or F64");
    if (!(cfg->tau > 0.5)) return set_error(D3Q19_EINVAL, "tau must be > 0.5");
    if (opt->steps < 1) return set_error(D3Q19_EINVAL, "steps must be >= 1");
    const int32_t cks = opt->checkpoint_storage < 0 ? cfg->storage : opt->checkpoint_storage;
    if (cks != cfg->storage && cks != D3Q19_F16)
        return set_error(D3Q19_EINVAL, "checkpoint storage must be the lattice storage or F16");
    const uint64_t ncell = (uint64_t)cfg->nx * cfg->ny * cfg->nz;
    // four or more full states must fit in memory, so no adjoint that can be
    // allocated reaches this bound; the (3, ncell) field offsets are 64-bit anyway
    if (3 * ncell >= (1ull << 32)) return set_error(D3Q19_EINVAL, "lattice too large (3 * ncell must be < 2^32)");
    if (cfg->nx > 65535 || cfg->ny > 65535) return set_error(D3Q19_EINVAL, "nx and ny must be <= 65535");
    if (hipSetDevice(cfg->device) != hipSuccess) return set_error(D3Q19_EHIP, "hipSetDevice failed");

    d3q19_adjoint* A = new d3q19_adjoint();
    A->cfg = *cfg;
    A->opt = *opt;
    A->T = opt->steps;
    A->ncell = (uint32_t)ncell;
    A->nzp = padded_nz((uint32_t)cfg->nz);
    A->es = cfg->storage == D3Q19_F64 ? 8 : 4;
    A->ckpt_same = cks == cfg->storage;
    A->brink = opt->brinkman != 0;
    const size_t rows = (size_t)cfg->nx * cfg->ny;
    A->state_bytes = rows * Q * A->nzp * A->es;
    A->ckpt_bytes = rows * Q * A->nzp * (A->ckpt_same ? A->es : 2);
    A->mom_bytes = 4 * (size_t)ncell * A->es;
    int32_t K = opt->checkpoint_every;
    if (K <= 0) K = adjoint_least_memory_K(A->T, A->ckpt_bytes, A->mom_bytes, A->state_bytes, field_bytes(A));
    K = std::min(K, A->T);
    if (opt->memory_budget && plan_bytes(A, K) > opt->memory_budget) {
        const size_t need = plan_bytes(A, K);
        delete A;
        return set_error(D3Q19_ENOMEM, "adjoint needs " + std::to_string(need) + " bytes with checkpoint interval " +
                                           std::to_string(K) + ", over the memory budget");
    }
    A->K = K;
    A->nck = (A->T + K - 1) / K;
    A->stream = (hipStream_t)stream;
    A->lds_per_cu = device_lds_per_cu(cfg->device);
    const size_t vec = 3 * (size_t)ncell * A->es;
    const size_t scal = (size_t)ncell * A->es;
    bool ok = true;
    A->ckpt.assign(A->nck, nullptr);
    A->seg.assign(2, nullptr);
    for (auto& p : A->ckpt) ok = ok && hipMalloc(&p, A->ckpt_bytes) == hipSuccess;
    for (auto& p : A->seg) ok = ok && hipMalloc(&p, A->state_bytes) == hipSuccess;
    ok = ok && hipMalloc(&A->moms, (size_t)K * A->mom_bytes) == hipSuccess;
    ok = ok && hipMalloc(&A->lam[0], A->state_bytes) == hipSuccess &&
         hipMalloc(&A->lam[1], A->state_bytes) == hipSuccess;
    ok = ok && hipMalloc(&A->u0, vec) == hipSuccess && hipMalloc(&A->force, vec) == hipSuccess;
    ok = ok && hipMalloc(&A->gF, vec) == hipSuccess && hipMalloc(&A->gu, vec) == hipSuccess;
    if (A->brink) ok = ok && hipMalloc(&A->s, scal) == hipSuccess && hipMalloc(&A->gs, scal) == hipSuccess;
    if (!ok) {
        free_all(A);
        delete A;
        return set_error(D3Q19_ENOMEM, "adjoint checkpoint allocation failed");
    }
    A->bytes = plan_bytes(A, K);
    *out = A;
    return D3Q19_OK;
}

int d3q19_adjoint_destroy(d3q19_adjoint* A) {
    if (!A) return D3Q19_OK;
    (void)hipStreamSynchronize(A->stream);
    free_all(A);
    delete A;
    return D3Q19_OK;
}

int d3q19_adjoint_set_stream(d3q19_adjoint* A, void* stream) {
    if (!A) return set_error(D3Q19_EINVAL, "null adjoint handle");
    // work queued on the previous stream completes before any on the new one
    if ((hipStream_t)stream != A->stream) D3Q19_HIPCHK(hipStreamSynchronize(A->stream));
    A->stream = (hipStream_t)stream;
    return D3Q19_OK;
}

int d3q19_adjoint_forward(d3q19_adjoint* A, const void* u0, const void* force, const void* brinkman,
                          void* u_final, int32_t dt, int32_t on_device) {
    if (!A) return set_error(D3Q19_EINVAL, "null adjoint handle");
    // a forward pass that fails (part-way, the checkpoints then belong to no
    // single trajectory) leaves backward refused until one completes
    A->have_forward = false;
    if (!u0 || !force || !u_final) return set_error(D3Q19_EINVAL, "null array");
    if (A->brink && !brinkman) return set_error(D3Q19_EINVAL, "this adjoint needs a Brinkman field");
    if (dt != D3Q19_F64 && dt != D3Q19_F32) return set_error(D3Q19_EINVAL, "dtype must be F64 or F32");
    const size_t n3 = 3 * (size_t)A->ncell;
    if (int e = stage_in(u0, dt, on_device, n3, A->u0, ct_double(A), A->stream)) return e;
    if (int e = stage_in(force, dt, on_device, n3, A->force, ct_double(A), A->stream)) return e;
    if (A->brink)
        if (int e = stage_in(brinkman, dt, on_device, A->ncell, A->s, ct_double(A), A->stream)) return e;
    // u^T in the caller's dtype: written directly when on device, else via gu scratch
    void* dst = u_final;
    const bool direct = on_device != 0;
    if (!direct) dst = A->gu;
    const int out_dt = direct ? dt : (ct_double(A) ? D3Q19_F64 : D3Q19_F32);
    int rc = ct_double(A) ? run_forward<TF64>(A, dst, out_dt) : run_forward<TF32>(A, dst, out_dt);
    if (rc) return rc;
    if (!direct) {
        if (int e = stage_out(A->gu, ct_double(A), n3, u_final, dt, 0, A->stream)) return e;
    } else {
        D3Q19_HIPCHK(hipStreamSynchronize(A->stream));
    }
    A->have_forward = true;
    return D3Q19_OK;
}

int d3q19_adjoint_backward(d3q19_adjoint* A, const void* grad_u_final, void* grad_u0, void* grad_force,
                           void* grad_brinkman, int32_t dt, int32_t on_device) {
    if (!A) return set_error(D3Q19_EINVAL, "null adjoint handle");
    if (!A->have_forward)
        return set_error(D3Q19_ESTATE, "adjoint_backward needs an adjoint_forward since the last backward "
                                       "(one backward per forward)");
    if (!grad_u_final) return set_error(D3Q19_EINVAL, "null dL/du_final");
    if (grad_brinkman && !A->brink) return set_error(D3Q19_EINVAL, "this adjoint has no Brinkman parameter");
    // The recomputation overwrites the segment buffers, one of which holds
    // H_T (the seed's input): a second backward would start from the wrong
    // state, so it is refused until the next forward.
    A->have_forward = false;
    const size_t n3 = 3 * (size_t)A->ncell;
    if (int e = stage_in(grad_u_final, dt, on_device, n3, A->gu, ct_double(A), A->stream)) return e;
    int rc = ct_double(A) ? run_backward<TF64>(A) : run_backward<TF32>(A);
    if (rc) return rc;
    if (grad_u0)
        if (int e = stage_out(A->gu, ct_double(A), n3, grad_u0, dt, on_device, A->stream)) return e;
    if (grad_force)
        if (int e = stage_out(A->gF, ct_double(A), n3, grad_force, dt, on_device, A->stream)) return e;
    if (grad_brinkman)
        if (int e = stage_out(A->gs, ct_double(A), A->ncell, grad_brinkman, dt, on_device, A->stream)) return e;
    D3Q19_HIPCHK(hipStreamSynchronize(A->stream));
    return D3Q19_OK;
}

int32_t d3q19_adjoint_checkpoint_every(const d3q19_adjoint* A) { return A ? A->K : -1; }
size_t d3q19_adjoint_memory_bytes(const d3q19_adjoint* A) { return A ? A->bytes : 0; }

}  // extern "C"
