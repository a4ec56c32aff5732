/*
 * d3q19.h -- C ABI of the MI355X D3Q19 lattice-Boltzmann engine (libd3q19.so).
 *
 * Drop-in for the step/read-back path of the reference package
 * packages/lattice-boltzmann-d3q19 (paths below are relative to the
 * reference root; NP = packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19).
 *
 * Conventions
 *  - Populations cross the ABI as (19, Nx, Ny, Nz) arrays, C order (z fastest),
 *    direction-major, in the reference's lexicographic velocity order
 *    (NP/reference/constants.py:39-59).  Vector fields (velocity, force) are
 *    (3, Nx, Ny, Nz); scalar fields (Nx, Ny, Nz).
 *  - dtype codes: D3Q19_F64, D3Q19_F32 for arrays crossing the ABI; the
 *    lattice storage type may also be D3Q19_F16 / D3Q19_BF16 / D3Q19_BF16SR
 *    (populations kept on the GPU as f_i - w_i, collision in FP32) or
 *    D3Q19_ME16 (the 19 monomial moments of f - w quantized to 16 bits,
 *    collision in FP32).
 *  - on_device != 0: the pointer is device memory on the lattice's device;
 *    otherwise host memory (the library stages it).
 *  - stream: a hipStream_t passed as void*; NULL is HIP's default stream.
 *    Calls are stream-ordered; read_* and diagnostics synchronise.
 *  - Every function returns a d3q19_status; d3q19_last_error() gives the
 *    message of the last failure on the calling thread.  No C++ exception
 *    crosses the ABI.  A handle must not be used by two threads at once.
 */
#ifndef D3Q19_H
#define D3Q19_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum d3q19_status {
    D3Q19_OK = 0,
    D3Q19_EINVAL = 1,  /* invalid argument or configuration */
    D3Q19_ENOMEM = 2,  /* device allocation failed */
    D3Q19_EHIP = 3,    /* HIP runtime error */
    D3Q19_ESTATE = 4   /* call not valid in the lattice's current state */
} d3q19_status;

/* Storage D3Q19_BF16SR is D3Q19_BF16 with stochastic rounding of every store
 * (unbiased; the rounding bits are a hash of cell and step, so runs repeat). */
/* Storage D3Q19_ME16_REF holds the same 16-bit moment codes as D3Q19_ME16,
 * moment-major (k * ncell + cell), and steps them in the moment-encoded
 * variant's own FP64 arithmetic (packages/lattice-boltzmann-d3q19-frontier-
 * moment-encoded/src/lbm_me.cpp:386-409 and its shaders): decode to full FP64
 * populations (lbm_decode.comp), BGK + Guo collision with the division-form
 * equilibrium (lbm_collide.comp), pull stream and (y-walls) half-way
 * bounce-back (lbm_stream_bounce.comp; moving walls take
 * apply_bounce_back_y_walls' momentum term), encode (lbm_encode.comp).
 * Periodic or y-wall boundaries; set_populations, set_equilibrium, step,
 * read_populations, read_moments, set_force_field and the raw state; the
 * reductions (diagnostics, health, enstrophy, moment_range), step_moments
 * and read_moments_noalloc return D3Q19_EINVAL.  Device memory: 2 B codes + 2 x 152 B
 * FP64 populations per cell. */
enum { D3Q19_F64 = 0, D3Q19_F32 = 1, D3Q19_F16 = 2, D3Q19_BF16 = 3, D3Q19_ME16 = 4, D3Q19_BF16SR = 5,
       D3Q19_ME16_REF = 6 };

enum {
    D3Q19_BC_PERIODIC = 0, /* periodic on all six faces (NP/reference/bgk.py:115-126) */
    D3Q19_BC_YWALLS = 1,   /* periodic, then apply_bounce_back_y_walls (NP/reference/bgk.py:145-216) */
    D3Q19_BC_HALFWAY = 2   /* textbook half-way bounce-back on selected faces + solid cells */
};

enum { D3Q19_FORCE_NONE = 0, D3Q19_FORCE_UNIFORM = 1, D3Q19_FORCE_FIELD = 2 };

/* Face indices for wall_faces / face_velocity (HALFWAY mode). */
enum { D3Q19_FACE_XM = 0, D3Q19_FACE_XP, D3Q19_FACE_YM, D3Q19_FACE_YP, D3Q19_FACE_ZM, D3Q19_FACE_ZP };

typedef struct d3q19_config {
    int32_t nx, ny, nz;
    double tau;                      /* BGK relaxation time, > 0.5 */
    int32_t storage;                 /* D3Q19_F32 (default) | F64 | F16 | BF16 */
    int32_t boundary;                /* D3Q19_BC_* */
    /* D3Q19_BC_YWALLS: the wall_velocity_top / wall_velocity_bottom
       arguments of apply_bounce_back_y_walls (NP/reference/bgk.py:145-149) */
    double wall_velocity_top[3];
    double wall_velocity_bottom[3];
    /* D3Q19_BC_HALFWAY: bit k set = face k is a wall (else periodic) */
    uint32_t wall_faces;
    double face_velocity[6][3];
    int32_t force_mode;              /* D3Q19_FORCE_*; FIELD uses d3q19_set_force_field */
    double force[3];                 /* D3Q19_FORCE_UNIFORM body force (Guo) */
    int32_t device;                  /* HIP device ordinal */
} d3q19_config;

typedef struct d3q19_lattice d3q19_lattice;

/* ---- library ---------------------------------------------------------- */
const char* d3q19_last_error(void);
const char* d3q19_version(void);
/* Fill cfg with: 0 extents, tau 0.6, F32 storage, periodic, no force, device 0. */
void d3q19_config_default(d3q19_config* cfg);

/* ---- stateful lattice (the step / read-back path) ---------------------
 * Replaces the runner loop of NP/sim.py:312-326 / :350-364:
 *   f = bgk_step(f, tau, force); f = apply_bounce_back_y_walls(f, top, bottom)
 * with one fused stream-collide kernel per step.  A lattice holds at most
 * 2^32 - 1 cells (nx, ny <= 65535); host arrays passed to it are staged
 * through device buffers of their size, device arrays are used in place.
 * A large lattice's step is launched as several kernels over runs of x
 * planes (one wavefront's results do not depend on it); the environment
 * variable D3Q19_STEP_CHUNKS, read here, forces the number of such launches
 * per step for the lattice (unset or 0: chosen from size and storage). */
int d3q19_create(const d3q19_config* cfg, void* stream, d3q19_lattice** out);
int d3q19_destroy(d3q19_lattice* lat);
/* Launch the lattice's later work on `stream` (NULL: the default stream).  Work
   already queued on the previous stream is waited for first, so that nothing
   queued on the new stream overtakes it. */
int d3q19_set_stream(d3q19_lattice* lat, void* stream);
/* Set the state to populations f (pre-collision, the reference's f^n). */
int d3q19_set_populations(d3q19_lattice* lat, const void* f, int32_t dtype, int32_t on_device);
/* Set the state to feq_field(rho, u) (NP/reference/equilibrium.py:71-98,
   NP/sim.py:281-285).  rho may be NULL for rho = 1.  One pass, no
   (19,Nx,Ny,Nz) staging. */
int d3q19_set_equilibrium(d3q19_lattice* lat, const void* rho, const void* u, int32_t dtype,
                          int32_t on_device);
/* Per-cell body force (3,Nx,Ny,Nz) for D3Q19_FORCE_FIELD (bgk.py:45-49 force_lattice). */
int d3q19_set_force_field(d3q19_lattice* lat, const void* force, int32_t dtype, int32_t on_device);
/* HALFWAY mode: solid cells (Nx,Ny,Nz) uint8, nonzero = solid; NULL clears. */
int d3q19_set_solid(d3q19_lattice* lat, const uint8_t* solid, int32_t on_device);
/* Brinkman solid fraction s in [0,1] (Nx,Ny,Nz): body force F - s * rho u. NULL clears. */
int d3q19_set_brinkman(d3q19_lattice* lat, const void* s, int32_t dtype, int32_t on_device);
/* Advance n steps. */
int d3q19_step(d3q19_lattice* lat, int32_t n);
/* Advance n >= 1 steps; the last step's own pass also writes rho (Nx,Ny,Nz)
   and u (3,Nx,Ny,Nz) of the state that step started from (f^{N-1} when the
   lattice ends at f^N): device FP32 pointers, either may be NULL. */
int d3q19_step_moments(d3q19_lattice* lat, int32_t n, float* rho_dev, float* u_dev);
/* Pre-collision populations f^n, as the reference holds them after step n. */
int d3q19_read_populations(d3q19_lattice* lat, void* f, int32_t dtype, int32_t on_device);
/* density_field(f^n) and macroscopic_velocity(f^n, force) (bgk.py:129-142).
   Either output may be NULL. */
int d3q19_read_moments(d3q19_lattice* lat, void* rho, void* u, int32_t dtype, int32_t on_device);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight capture/serving loops, harnesses that forbid hipMalloc):
 * d3q19_read_moments into HOST rho (Nx,Ny,Nz) and/or u (3,Nx,Ny,Nz), staged
 * through a fixed device buffer of the library (D3Q19_NOALLOC_SCRATCH_BYTES
 * per GPU, held until the process ends) instead of a per-call allocation;
 * it calls no hipMalloc/hipFree and no allocating routine.  Limit:
 * 4 * ncell * sizeof(dtype) <= D3Q19_NOALLOC_SCRATCH_BYTES (FP64: 524,288
 * cells, FP32: 1,048,576), else D3Q19_EINVAL.  All calls on one GPU share
 * the buffer: calls that overlap on one GPU (from several threads or
 * streams) race.  Synchronises the lattice's stream. */
#define D3Q19_NOALLOC_SCRATCH_BYTES (16u * 1024u * 1024u)
int d3q19_read_moments_noalloc(d3q19_lattice* lat, void* rho, void* u, int32_t dtype);
/* FP64 reductions over fluid cells of f^n: out[0] mass, out[1..3] momentum
   sum(c f), out[4] kinetic energy 0.5 sum(rho |u|^2) with the read-back u. */
int d3q19_diagnostics(d3q19_lattice* lat, double out[5]);
/* Health of f^n in one pass (the runners' tier-1 check, NP/sim.py:135-145):
 * out = [cells with a NaN/Inf population, rho or u; min rho; max rho;
 * max |u|] over the finite cells (solid cells excluded). */
int d3q19_health(d3q19_lattice* lat, double out[4]);
/* FP64 enstrophy 0.5 sum |curl u|^2 of the read-back u (Guo-corrected, in the
 * compute type: FP64 for F64 storage, else FP32), with periodic central
 * differences 0.5 (u(x+1) - u(x-1)).  Defined for periodic lattices only:
 * returns D3Q19_EINVAL for reference y-walls and for half-way lattices with a
 * wall face or solid cells.  (No reference counterpart: a diagnostic of the
 * decaying-turbulence workload.) */
int d3q19_enstrophy(d3q19_lattice* lat, double* out);
int d3q19_synchronize(d3q19_lattice* lat);
int64_t d3q19_step_count(const d3q19_lattice* lat);
/* Kernel launches one step makes once the state is streamed: the fused
 * step's plane chunks (see d3q19_create), 2 for D3Q19_ME16_REF (decode +
 * collide, stream + encode); -1 for a null lattice.  (No reference
 * counterpart: it tells a profile's per-launch figures from per-step ones.) */
int32_t d3q19_step_launches(const d3q19_lattice* lat);
/* Set the step counter (the BF16SR rounding bits are keyed on it). */
int d3q19_set_step_count(d3q19_lattice* lat, int64_t n);

/* Checkpoint / resume of the stored state itself (no reference counterpart:
 * the reference has no restart path).  The state is the lattice's population
 * buffer in its storage format and internal layout -- an opaque blob of
 * d3q19_state_bytes() bytes, valid for a lattice of the same shape, storage
 * (and, for ME16, moment ranges) -- plus whether it holds f^n in place
 * (`local`) and the step count.  Writing it back and stepping repeats the
 * uninterrupted run bit for bit in every storage.  Host memory. */
size_t d3q19_state_bytes(const d3q19_lattice* lat);
int d3q19_read_state(d3q19_lattice* lat, void* dst, int32_t* local, int64_t* steps);
int d3q19_write_state(d3q19_lattice* lat, const void* src, int32_t local, int64_t steps);
size_t d3q19_device_bytes(const d3q19_lattice* lat);

/* ---- 16-bit moment-encoded storage (D3Q19_ME16) -------------------------
 * The frontier-moment-encoded variant's codec
 * (packages/lattice-boltzmann-d3q19-frontier-moment-encoded: src/lbm_me.cpp:83-150
 * build_moment_basis, :222-253 calibrate_ranges, shaders/lbm_encode.comp,
 * shaders/lbm_decode.comp) as a population storage of the engine: per cell
 * the 19 moments m = M f over the velocity-set monomials (rows 0..3: density
 * and momentum), each as code = nearest((m - lo) / (hi - lo) * 65535),
 * clamped; f = M^-1 m on use.  The codes hold what every storage holds: the
 * post-collision populations of the last step (f^n in place after a load).
 * One fused pass per step: each pulled population is decoded from its
 * source cell's codes (row i of M^-1, 3-5 codes), the cell collides in FP32
 * and its 19 codes are stored.  Device memory: on lattices with rows of 64
 * cells or more, one ring of nx + gap x planes of codes (38 B per cell and
 * a gap of a few planes: about 40 B per cell at 256^3) -- the step writes
 * each plane of the new state into the slot of a plane gap planes back, in
 * place, once the pulls from that plane are consumed; the environment
 * variable D3Q19_RING_GAP, read by d3q19_create, sets a smaller gap (a test
 * knob, at least 2).  A wait that gives up (it should not) makes every
 * synchronising call return D3Q19_EHIP until a state is loaded again
 * (set_populations, set_equilibrium, write_state); D3Q19_RING_FAULT_TEST=1
 * (a test knob) makes every wait give up at once.  Shorter rows: 2 x 38 B
 * per cell.
 * The ranges must cover the moments of every stored state: calibrate them with
 * d3q19_moment_range_stored.  Storage D3Q19_ME16_REF (above) keeps the
 * variant's own quantization point (f^n) and FP64 arithmetic instead. */
/* M (moments = M f, row-major 19 x 19, entries 0/+-1) and 4 M^-1 (integers). */
int d3q19_moment_basis(int32_t m[361], int32_t minv4[361]);
/* Quantization ranges of the 19 moments of f (not of f - w); ME16 and
 * ME16_REF lattices only, before the first set_populations /
 * set_equilibrium / write_state. */
int d3q19_set_moment_ranges(d3q19_lattice* lat, const double lo[19], const double hi[19]);
/* FP64 minimum and maximum over all cells of the 19 moments of the current
 * state f^n (any storage): one frame of the range calibration. */
int d3q19_moment_range(d3q19_lattice* lat, double lo[19], double hi[19]);
/* The same over the stored state as it lies in the lattice's buffer: the
 * post-collision populations of the last step, or the loaded f^n before the
 * first step (any storage but ME16_REF).  A D3Q19_ME16 lattice encodes
 * exactly these states, so an FP64 run's ranges from this call, frame 0 ..
 * T, cover a T-step D3Q19_ME16 run of the same configuration.  No reference
 * counterpart: the variant quantizes f^n. */
int d3q19_moment_range_stored(d3q19_lattice* lat, double lo[19], double hi[19]);

/* ---- stateless drop-ins (NP/reference/__init__.py:43-71) ---------------
 * Device pointers, dtype D3Q19_F64 or D3Q19_F32, (19,Nx,Ny,Nz) layout,
 * stream-ordered on `stream`.  FP64 results follow the reference's operation
 * order (no FMA contraction). */
/* bgk_step (NP/reference/bgk.py:45-112); force may be NULL. */
int d3q19_bgk_step(const void* f, void* out, const void* force, int32_t nx, int32_t ny, int32_t nz,
                   double tau, int32_t dtype, void* stream);
/* stream (NP/reference/bgk.py:115-126). */
int d3q19_stream(const void* f, void* out, int32_t nx, int32_t ny, int32_t nz, int32_t dtype,
                 void* stream);
/* apply_bounce_back_y_walls (NP/reference/bgk.py:145-216). */
int d3q19_bounce_back_y_walls(const void* f, void* out, const double wall_velocity_top[3],
                              const double wall_velocity_bottom[3], int32_t nx, int32_t ny,
                              int32_t nz, int32_t dtype, void* stream);
/* feq_field (NP/reference/equilibrium.py:71-98). */
int d3q19_feq_field(const void* rho, const void* u, void* out, int64_t ncell, int32_t dtype,
                    void* stream);
/* feq (NP/reference/equilibrium.py:34-53), the point function, at each of
 * ncell (rho, u) pairs: rho (ncell) or NULL for rho = 1, u (3, ncell), out
 * (19, ncell).  It divides by c_s^2, 2 c_s^4 and 2 c_s^2 as the point function
 * does, where feq_field multiplies by their reciprocals, so in FP64 each
 * value equals the reference's feq(rho, u) to the bit. */
int d3q19_feq(const void* rho, const void* u, void* out, int64_t ncell, int32_t dtype, void* stream);
/* density_field (NP/reference/equilibrium.py:101-103). */
int d3q19_density_field(const void* f, void* rho, int64_t ncell, int32_t dtype, void* stream);
/* momentum_field (NP/reference/equilibrium.py:106-113). */
int d3q19_momentum_field(const void* f, void* mom, int64_t ncell, int32_t dtype, void* stream);
/* macroscopic_velocity (NP/reference/bgk.py:129-142); force may be NULL. */
int d3q19_macroscopic_velocity(const void* f, const void* force, void* u, int64_t ncell,
                               int32_t dtype, void* stream);

/* ---- differentiable step (no reference counterpart) ------------------
 * Forward map with a per-cell Guo force F and an optional Brinkman solid
 * fraction s (body force F - s m):
 *   f^0 = feq_field(1, u0);  f^{n+1} = bgk_step(f^n, tau, F - s m) [+ walls]
 *   u^T = macroscopic_velocity(f^T, F - s m)
 * The backward pass propagates dL/du^T to dL/du0, dL/dF and dL/ds with a
 * hand-written adjoint kernel (reverse streaming with the bounce-back links
 * reversed + transposed BGK/Guo/Brinkman Jacobian), recomputing the states
 * between checkpoints.  cfg: force_mode D3Q19_FORCE_FIELD; boundary PERIODIC,
 * YWALLS (the reference loop bgk_step + apply_bounce_back_y_walls with the
 * wall velocities, NP/sim.py:312-326) or HALFWAY (stationary or moving
 * faces, no solid cells); storage F32 or F64 (the adjoint runs in the
 * storage precision).  Fields are (3,Nx,Ny,Nz) or (Nx,Ny,Nz) device or host
 * arrays of dtype F32/F64. */
typedef struct d3q19_adjoint d3q19_adjoint;

typedef struct d3q19_adjoint_options {
    int32_t steps;               /* T >= 1 */
    int32_t checkpoint_every;    /* K; <= 0: the K that needs the least memory */
    int32_t checkpoint_storage;  /* the lattice storage (default), or D3Q19_F16 (f - w) */
    int32_t brinkman;            /* nonzero: s is a parameter of the forward map */
    uint64_t memory_budget;      /* bytes the adjoint may hold, 0 = unlimited */
} d3q19_adjoint_options;

void d3q19_adjoint_options_default(d3q19_adjoint_options* opt);
int d3q19_adjoint_create(const d3q19_config* cfg, const d3q19_adjoint_options* opt, void* stream,
                         d3q19_adjoint** out);
int d3q19_adjoint_destroy(d3q19_adjoint* adj);
/* As d3q19_set_stream: work queued on the previous stream is waited for first. */
int d3q19_adjoint_set_stream(d3q19_adjoint* adj, void* stream);
/* Run the forward map from (u0, F, s), keep checkpoints, write u^T.  s is
   ignored (may be NULL) unless the options enable the Brinkman parameter. */
int d3q19_adjoint_forward(d3q19_adjoint* adj, const void* u0, const void* force, const void* brinkman,
                          void* u_final, int32_t dtype, int32_t on_device);
/* Given dL/du^T, write dL/du0, dL/dF and dL/ds (any may be NULL). Needs a
   forward pass that completed: after a failed one it returns D3Q19_ESTATE. */
int d3q19_adjoint_backward(d3q19_adjoint* adj, const void* grad_u_final, void* grad_u0,
                           void* grad_force, void* grad_brinkman, int32_t dtype, int32_t on_device);
int32_t d3q19_adjoint_checkpoint_every(const d3q19_adjoint* adj);
size_t d3q19_adjoint_memory_bytes(const d3q19_adjoint* adj);

/* ---- calibration (no reference counterpart) ----------------------------
 * Plain device copy of `bytes` (multiple of 16) with 16-byte nontemporal
 * loads/stores, stream-ordered: the measured HBM copy peak and the known
 * byte count that scales rocprof FETCH_SIZE / WRITE_SIZE. */
int d3q19_copy_kernel(const void* src, void* dst, size_t bytes, void* stream);
/* Current shader clock in MHz: one wavefront counts shader-clock ticks
 * against the device's constant real-time counter for about 20 us.  `ticks`
 * is a caller-owned device buffer of 16 bytes; the call synchronises the
 * stream.  Read between timed launches to see power-management clock
 * transients. */
int d3q19_shader_clock(void* ticks, double* mhz, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* D3Q19_H */
