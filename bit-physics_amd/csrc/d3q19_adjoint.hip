// d3q19_adjoint.hip -- discrete adjoint of the periodic Guo-forced D3Q19 step.
//
// Forward map (periodic, per-cell force F):
//   f^0 = feq(1, u0),  f^{n+1} = S C(f^n; F),  u^T = (m(f^T) + F/2) / rho(f^T)
// with C the BGK+Guo collision and S the streaming permutation.  The engine
// stores post-collision states, so the forward runs on
//   H_0 = S^-1 f^0,  H_{j+1} = k_step(H_j),  f^j = S H_j.
// Backward (lambda^n = dL/df^n):
//   lambda^T   = J_u(f^T)^T dL/du^T
//   lambda^n   = J_C(f^n)^T S^T lambda^{n+1},   (S^T mu)_i(x) = mu_i(x + c_i)
//   dL/dF     += dC/dF(f^n)^T S^T lambda^{n+1}  +  dL/du^T * 1/(2 rho^T)
//   dL/du0     = dfeq(1,u0)/du0^T lambda^0
// States between checkpoints (every K steps) are recomputed by the forward
// kernel itself, writing straight into the segment slots.
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

// Initial post-collision "pre-image": H_0 = S^-1 feq(1, u0), i.e.
// H_0,i(y) = feq_i(1, u0(y + c_i)) - w_i.
template <class Tag>
__global__ void __launch_bounds__(step_block<Tag>())
    k_adj_unstream(const typename Codec<Tag>::C* __restrict__ u0, typename Codec<Tag>::raw* __restrict__ out,
                   Params<typename Codec<Tag>::C> p) {
    typedef typename Codec<Tag>::C C;
    uint32_t x, y, z;
    if (!grid_cell(p.nx, p.ny, p.nz, x, y, z)) return;
    const uint32_t N = p.ncell;
    C d[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const uint32_t sx = (x + p.nx + CX[i]) % p.nx, sy = (y + p.ny + CY[i]) % p.ny, sz = (z + p.nz + CZ[i]) % p.nz;
        const uint32_t src = (sx * p.ny + sy) * p.nz + sz;
        const C ux = u0[src], uy = u0[N + src], uz = u0[2 * N + src];
        const C cu = C(CX[i]) * ux + C(CY[i]) * uy + C(CZ[i]) * uz;
        d[i] = wgt<C>(i) * (C(3) * cu + C(4.5) * cu * cu - C(1.5) * (ux * ux + uy * uy + uz * uz));
    }
    store_local<Tag>(out, p, x, y, z, d);
}

// Seed: lambda^T_j = sum_a g_a (c_ja - u_a) / rho ;  gF_a = g_a / (2 rho).
template <class Tag>
__global__ void __launch_bounds__(step_block<Tag>())
    k_adj_seed(const typename Codec<Tag>::raw* __restrict__ H, Params<typename Codec<Tag>::C> p,
               const typename Codec<Tag>::C* __restrict__ g, typename Codec<Tag>::raw* __restrict__ lam,
               typename Codec<Tag>::C* __restrict__ gF) {
    typedef typename Codec<Tag>::C C;
    uint32_t x, y, z;
    if (!grid_cell(p.nx, p.ny, p.nz, x, y, z)) return;
    const uint32_t cell = (x * p.ny + y) * p.nz + z;
    C d[Q];
    pull<Tag, BC_PERIODIC>(H, p, x, y, z, d);
    C drho, mx, my, mz;
    moments(d, drho, mx, my, mz);
    const C rho = C(1) + drho;
    const C ir = C(1) / rho;
    const uint32_t N = p.ncell;
    const C fx = p.force[cell], fy = p.force[N + cell], fz = p.force[2 * N + cell];
    const C ux = (mx + C(0.5) * fx) * ir, uy = (my + C(0.5) * fy) * ir, uz = (mz + C(0.5) * fz) * ir;
    const C gx = g[cell], gy = g[N + cell], gz = g[2 * N + cell];
    const C gu = gx * ux + gy * uy + gz * uz;
    C l[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) l[j] = (C(CX[j]) * gx + C(CY[j]) * gy + C(CZ[j]) * gz - gu) * ir;
    store_local<Tag>(lam, p, x, y, z, l);
    gF[cell] = C(0.5) * gx * ir;
    gF[N + cell] = C(0.5) * gy * ir;
    gF[2 * N + cell] = C(0.5) * gz * ir;
}

// One reverse step: lambda_out = J_C(f^n)^T S^T lambda_in, gF += dC/dF^T S^T lambda_in.
template <class Tag>
__global__ void __launch_bounds__(step_block<Tag>())
    k_adj_step(const typename Codec<Tag>::raw* __restrict__ H, Params<typename Codec<Tag>::C> p,
               const typename Codec<Tag>::raw* __restrict__ lam_in, typename Codec<Tag>::raw* __restrict__ lam_out,
               typename Codec<Tag>::C* __restrict__ gF) {
    typedef typename Codec<Tag>::C C;
    constexpr uint32_t ES = sizeof(typename Codec<Tag>::raw);
    uint32_t x, y, z;
    if (!grid_cell(p.nx, p.ny, p.nz, x, y, z)) return;
    const uint32_t cell = (x * p.ny + y) * p.nz + z;
    const uint32_t N = p.ncell;
    // S^T: mu_i(x) = lambda_i(x + c_i)  (the pull run backwards)
    const uint32_t xs[3] = {x == 0 ? p.nx - 1 : x - 1, x, x + 1 == p.nx ? 0 : x + 1};
    const uint32_t ys[3] = {y == 0 ? p.ny - 1 : y - 1, y, y + 1 == p.ny ? 0 : y + 1};
    const uint32_t zs[3] = {z == 0 ? p.nz - 1 : z - 1, z, z + 1 == p.nz ? 0 : z + 1};
    C mu[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i)
        mu[i] = ld<Tag>(row_of(lam_in, p, xs[1 + CX[i]], ys[1 + CY[i]]), i, p.dir_bytes, zs[1 + CZ[i]] * ES);
    C d[Q];
    pull<Tag, BC_PERIODIC>(H, p, x, y, z, d);
    C drho, mx, my, mz;
    moments(d, drho, mx, my, mz);
    const C rho = C(1) + drho;
    const C ir = C(1) / rho;
    const C F[3] = {p.force[cell], p.force[N + cell], p.force[2 * N + cell]};
    const C u[3] = {(mx + C(0.5) * F[0]) * ir, (my + C(0.5) * F[1]) * ir, (mz + C(0.5) * F[2]) * ir};
    const C usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const C om = p.omega, pref = p.pref;
    C a_rho = C(0);
    C b[3] = {C(0), C(0), C(0)};
    C ge[3] = {C(0), C(0), C(0)};
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const C c[3] = {C(CX[i]), C(CY[i]), C(CZ[i])};
        const C s = mu[i] * wgt<C>(i);
        const C cu = c[0] * u[0] + c[1] * u[1] + c[2] * u[2];
        const C cF = c[0] * F[0] + c[1] * F[1] + c[2] * F[2];
        a_rho += s * (C(1) + C(3) * cu + C(4.5) * cu * cu - C(1.5) * usq);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            // d(feq_i/tau)/du_a = w rho (3 c_a + 9 cu c_a - 3 u_a)/tau ; dG/du_a = pref w (-3 F_a + 9 c_a cF)
            b[a] += s * (om * rho * (C(3) * c[a] + C(9) * cu * c[a] - C(3) * u[a]) +
                         pref * (C(9) * c[a] * cF - C(3) * F[a]));
            // explicit dG/dF_a = pref w (3 (c_a - u_a) + 9 cu c_a)
            ge[a] += s * pref * (C(3) * (c[a] - u[a]) + C(9) * cu * c[a]);
        }
    }
    a_rho *= om;
    const C keep = C(1) - om;
    C l[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j)
        l[j] = keep * mu[j] + a_rho +
               (b[0] * (C(CX[j]) - u[0]) + b[1] * (C(CY[j]) - u[1]) + b[2] * (C(CZ[j]) - u[2])) * ir;
    store_local<Tag>(lam_out, p, x, y, z, l);
#pragma unroll
    for (int a = 0; a < 3; ++a) gF[(size_t)a * N + cell] += ge[a] + C(0.5) * b[a] * ir;
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
    const C u[3] = {u0[cell], u0[N + cell], u0[2 * N + cell]};
    C g[3] = {C(0), C(0), C(0)};
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const C c[3] = {C(CX[i]), C(CY[i]), C(CZ[i])};
        const C s = l[i] * wgt<C>(i);
        const C cu = c[0] * u[0] + c[1] * u[1] + c[2] * u[2];
#pragma unroll
        for (int a = 0; a < 3; ++a) g[a] += s * (C(3) * c[a] + C(9) * cu * c[a] - C(3) * u[a]);
    }
    gu0[cell] = g[0];
    gu0[N + cell] = g[1];
    gu0[2 * N + cell] = g[2];
}

}  // namespace

struct d3q19_adjoint {
    d3q19_config cfg;
    int32_t T = 0, K = 0, nck = 0;
    hipStream_t stream = nullptr;  // nullptr: the default stream
    uint32_t ncell = 0;
    size_t es = 0, state_bytes = 0, row_bytes = 0;
    uint32_t nzp = 0;
    std::vector<void*> ckpt;  // H_{kK}
    std::vector<void*> seg;   // recompute slots (>= 2, also the forward ping-pong)
    void* lam[2] = {nullptr, nullptr};
    void* u0 = nullptr;
    void* force = nullptr;
    void* gF = nullptr;
    void* gu = nullptr;    // staged dL/du^T, reused for dL/du0
    void* Hfinal = nullptr;
    bool have_forward = false;
    size_t bytes = 0;
};

namespace {

bool ct_double(const d3q19_adjoint* A) { return A->cfg.storage == D3Q19_F64; }

template <class C>
Params<C> adj_params(const d3q19_adjoint* A) {
    Params<C> p;
    std::memset(&p, 0, sizeof(p));
    p.nx = A->cfg.nx;
    p.ny = A->cfg.ny;
    p.nz = A->cfg.nz;
    p.ncell = A->ncell;
    p.nzp = A->nzp;
    p.dir_bytes = (uint32_t)(A->nzp * A->es);
    p.row_bytes = A->row_bytes;
    set_collision_constants(p, A->cfg.tau);
    p.force_mode = FORCE_FIELD;
    p.force = (const C*)A->force;
    return p;
}

// (z, y, x) launch shape shared by every kernel here.
template <class Tag>
void adj_grid(const d3q19_adjoint* A, dim3& grid, dim3& block) {
    const uint32_t nz = A->cfg.nz;
    const uint32_t bz = nz >= 256 ? 256 : (nz >= 128 ? 128 : 64);
    const uint32_t by = step_block<Tag>() / bz;
    block = dim3(bz, by, 1);
    grid = dim3((nz + bz - 1) / bz, (A->cfg.ny + by - 1) / by, A->cfg.nx);
}

template <class Tag>
int step_into(d3q19_adjoint* A, void* in, void* out) {
    typedef typename Codec<Tag>::C C;
    typedef typename Codec<Tag>::raw R;
    const Params<C> p = adj_params<C>(A);
    dim3 g, b;
    adj_grid<Tag>(A, g, b);
    k_step<Tag, BC_PERIODIC, true><<<g, b, 0, A->stream>>>((const R*)in, (R*)out, p, nullptr, nullptr);
    D3Q19_HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

template <class Tag>
int run_forward(d3q19_adjoint* A, void* u_out_dev, int out_dt) {
    typedef typename Codec<Tag>::C C;
    typedef typename Codec<Tag>::raw R;
    const Params<C> p = adj_params<C>(A);
    dim3 g, b;
    adj_grid<Tag>(A, g, b);
    k_adj_unstream<Tag><<<g, b, 0, A->stream>>>((const C*)A->u0, (R*)A->ckpt[0], p);
    D3Q19_HIPCHK(hipGetLastError());
    void* cur = A->ckpt[0];
    int pp = 0;  // next ping-pong slot
    for (int32_t j = 0; j < A->T; ++j) {
        const int32_t nxt = j + 1;
        void* out;
        if (nxt % A->K == 0 && nxt / A->K < A->nck) {
            out = A->ckpt[nxt / A->K];
        } else {
            out = A->seg[pp];
            if (out == cur) out = A->seg[pp ^ 1];
            pp = (out == A->seg[0]) ? 1 : 0;
        }
        if (int e = step_into<Tag>(A, cur, out)) return e;
        cur = out;
    }
    A->Hfinal = cur;
    if (out_dt == D3Q19_F64)
        k_readback<Tag, BC_PERIODIC, double><<<g, b, 0, A->stream>>>((const R*)cur, p, nullptr, nullptr,
                                                                     (double*)u_out_dev);
    else
        k_readback<Tag, BC_PERIODIC, float><<<g, b, 0, A->stream>>>((const R*)cur, p, nullptr, nullptr,
                                                                    (float*)u_out_dev);
    D3Q19_HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

template <class Tag>
int run_backward(d3q19_adjoint* A) {
    typedef typename Codec<Tag>::C C;
    typedef typename Codec<Tag>::raw R;
    const Params<C> p = adj_params<C>(A);
    dim3 g, b;
    adj_grid<Tag>(A, g, b);
    k_adj_seed<Tag><<<g, b, 0, A->stream>>>((const R*)A->Hfinal, p, (const C*)A->gu, (R*)A->lam[0], (C*)A->gF);
    D3Q19_HIPCHK(hipGetLastError());
    int cur = 0;
    for (int32_t k = A->nck - 1; k >= 0; --k) {
        const int32_t n0 = k * A->K;
        const int32_t len = std::min(A->K, A->T - n0);
        // slot 0 = checkpoint, slots 1..len-1 recomputed into seg[0..len-2]
        auto slot = [&](int32_t j) -> void* { return j == 0 ? A->ckpt[k] : A->seg[j - 1]; };
        for (int32_t j = 1; j < len; ++j)
            if (int e = step_into<Tag>(A, slot(j - 1), slot(j))) return e;
        for (int32_t j = len - 1; j >= 0; --j) {
            k_adj_step<Tag><<<g, b, 0, A->stream>>>((const R*)slot(j), p, (const R*)A->lam[cur],
                                                     (R*)A->lam[cur ^ 1], (C*)A->gF);
            D3Q19_HIPCHK(hipGetLastError());
            cur ^= 1;
        }
    }
    k_adj_u0<Tag><<<g, b, 0, A->stream>>>((const R*)A->lam[cur], p, (const C*)A->u0, (C*)A->gu);
    D3Q19_HIPCHK(hipGetLastError());
    return D3Q19_OK;
}

void free_all(d3q19_adjoint* A) {
    for (void* p : A->ckpt)
        if (p) (void)hipFree(p);
    for (void* p : A->seg)
        if (p) (void)hipFree(p);
    void* ps[] = {A->lam[0], A->lam[1], A->u0, A->force, A->gF, A->gu};
    for (void* p : ps)
        if (p) (void)hipFree(p);
}

}  // namespace

extern "C" {

int d3q19_adjoint_create(const d3q19_config* cfg, int32_t steps, int32_t K, void* stream,
                         d3q19_adjoint** out) {
    if (!cfg || !out) return set_error(D3Q19_EINVAL, "null argument");
    *out = nullptr;
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return set_error(D3Q19_EINVAL, "extents must be >= 1");
    if (cfg->boundary != D3Q19_BC_PERIODIC) return set_error(D3Q19_EINVAL, "adjoint needs a periodic lattice");
    if (cfg->force_mode != D3Q19_FORCE_FIELD) return set_error(D3Q19_EINVAL, "adjoint needs force_mode FIELD");
    if (cfg->storage != D3Q19_F32 && cfg->storage != D3Q19_F64)
        return set_error(D3Q19_EINVAL, "adjoint storage must be F32 or F64");
    if (!(cfg->tau > 0.5)) return set_error(D3Q19_EINVAL, "tau must be > 0.5");
    if (steps < 1) return set_error(D3Q19_EINVAL, "steps must be >= 1");
    const uint64_t ncell = (uint64_t)cfg->nx * cfg->ny * cfg->nz;
    if (ncell * 8 >= (1ull << 32)) return set_error(D3Q19_EINVAL, "lattice too large");
    if (cfg->nx > 65535 || cfg->ny > 65535) return set_error(D3Q19_EINVAL, "nx and ny must be <= 65535");
    if (K <= 0) K = std::max(2, (int32_t)std::lround(std::sqrt((double)steps)));
    K = std::min(K, steps);
    if (hipSetDevice(cfg->device) != hipSuccess) return set_error(D3Q19_EHIP, "hipSetDevice failed");

    d3q19_adjoint* A = new d3q19_adjoint();
    A->cfg = *cfg;
    A->T = steps;
    A->K = K;
    A->nck = (steps + K - 1) / K;
    A->ncell = (uint32_t)ncell;
    A->nzp = padded_nz((uint32_t)cfg->nz);
    A->es = cfg->storage == D3Q19_F64 ? 8 : 4;
    A->row_bytes = (size_t)Q * A->nzp * A->es;
    A->state_bytes = (size_t)cfg->nx * cfg->ny * A->row_bytes;
    A->stream = (hipStream_t)stream;  // NULL: the default stream
    const size_t state = A->state_bytes;
    const size_t lam = A->state_bytes;  // same layout and element size
    const size_t vec = 3 * (size_t)ncell * A->es;
    bool ok = true;
    A->ckpt.assign(A->nck, nullptr);
    A->seg.assign(std::max(2, K - 1), nullptr);
    for (auto& p : A->ckpt) ok = ok && hipMalloc(&p, state) == hipSuccess;
    for (auto& p : A->seg) ok = ok && hipMalloc(&p, state) == hipSuccess;
    ok = ok && hipMalloc(&A->lam[0], lam) == hipSuccess && hipMalloc(&A->lam[1], lam) == hipSuccess;
    ok = ok && hipMalloc(&A->u0, vec) == hipSuccess && hipMalloc(&A->force, vec) == hipSuccess;
    ok = ok && hipMalloc(&A->gF, vec) == hipSuccess && hipMalloc(&A->gu, vec) == hipSuccess;
    if (!ok) {
        free_all(A);
        delete A;
        return set_error(D3Q19_ENOMEM, "adjoint checkpoint allocation failed");
    }
    A->bytes = (A->ckpt.size() + A->seg.size()) * state + 2 * lam + 4 * vec;
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
    A->stream = (hipStream_t)stream;
    return D3Q19_OK;
}

int d3q19_adjoint_forward(d3q19_adjoint* A, const void* u0, const void* force, void* u_final, int32_t dt,
                          int32_t on_device) {
    if (!A) return set_error(D3Q19_EINVAL, "null adjoint handle");
    if (!u0 || !force || !u_final) return set_error(D3Q19_EINVAL, "null array");
    if (dt != D3Q19_F64 && dt != D3Q19_F32) return set_error(D3Q19_EINVAL, "dtype must be F64 or F32");
    const size_t n3 = 3 * (size_t)A->ncell;
    if (int e = stage_in(u0, dt, on_device, n3, A->u0, ct_double(A), A->stream)) return e;
    if (int e = stage_in(force, dt, on_device, n3, A->force, ct_double(A), A->stream)) return e;
    // u^T in the caller's dtype: written directly when on device, else via gu scratch
    void* dst = u_final;
    const bool direct = on_device != 0;
    if (!direct) dst = A->gu;  // scratch in compute type
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
                           int32_t dt, int32_t on_device) {
    if (!A) return set_error(D3Q19_EINVAL, "null adjoint handle");
    if (!A->have_forward) return set_error(D3Q19_ESTATE, "adjoint_backward needs adjoint_forward first");
    if (!grad_u_final) return set_error(D3Q19_EINVAL, "null dL/du_final");
    const size_t n3 = 3 * (size_t)A->ncell;
    if (int e = stage_in(grad_u_final, dt, on_device, n3, A->gu, ct_double(A), A->stream)) return e;
    int rc = ct_double(A) ? run_backward<TF64>(A) : run_backward<TF32>(A);
    if (rc) return rc;
    if (grad_u0)
        if (int e = stage_out(A->gu, ct_double(A), n3, grad_u0, dt, on_device, A->stream)) return e;
    if (grad_force)
        if (int e = stage_out(A->gF, ct_double(A), n3, grad_force, dt, on_device, A->stream)) return e;
    D3Q19_HIPCHK(hipStreamSynchronize(A->stream));
    return D3Q19_OK;
}

int32_t d3q19_adjoint_checkpoint_every(const d3q19_adjoint* A) { return A ? A->K : -1; }
size_t d3q19_adjoint_memory_bytes(const d3q19_adjoint* A) { return A ? A->bytes : 0; }

}  // extern "C"
