/*
 * d3q19_oracle.c -- FP64 CPU restatement of the reference D3Q19 BGK step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is built into
 * oracle/_build/libd3q19_oracle.so and is loaded only by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  The
 * product path (bit-physics_amd/) never links, loads or calls it.
 *
 * What it restates (paths relative to the reference root):
 *   NP = packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/
 *   - constants          NP/reference/constants.py:39-83   (lex velocity set, weights, c_s^2)
 *   - opposite map       NP/reference/bgk.py:164-187
 *   - feq (point)        NP/reference/equilibrium.py:34-53  (division operand form)
 *   - feq_field          NP/reference/equilibrium.py:71-98  (reciprocal operand form)
 *   - density_field      NP/reference/equilibrium.py:101-103
 *   - momentum_field     NP/reference/equilibrium.py:106-113 (lex 19-sum incl. zero terms,
 *                        the order the Warp port uses bit-exactly:
 *                        packages/lattice-boltzmann-d3q19-stack-e/.../d3q19_warp.py:123-141)
 *   - bgk_step           NP/reference/bgk.py:45-112          (collide with Guo, then stream)
 *   - stream             NP/reference/bgk.py:115-126         (np.roll per direction)
 *   - macroscopic_velocity NP/reference/bgk.py:129-142
 *   - apply_bounce_back_y_walls NP/reference/bgk.py:145-216 (post-stream copy from the
 *                        opposite direction at y=0 / y=Ny-1, moving-wall term
 *                        -2 w_i rho_wall (c_i.u_w)/c_s^2, rho_wall summed over all 19
 *                        post-stream values of the wall cell before the copy)
 *   - step composition   NP/sim.py:312-326 (Poiseuille) and :350-364 (Couette)
 *
 * Operation order follows the NumPy expressions term by term (left-to-right
 * evaluation of each expression, sequential lex sums), so with
 * -ffp-contract=off the results agree with the NumPy reference to the bit
 * wherever libm is not involved.
 *
 * Additions with no reference counterpart (closed-form pins only, stated in
 * the function comments): textbook half-way bounce-back walls with moving
 * faces and solid cells (lid-driven cavity, obstacles), the Brinkman
 * solid-fraction force, FP64 diagnostics (mass, momentum, kinetic energy),
 * and the FP64 discrete adjoint of the periodic Guo-forced step.
 *
 * Planted mistakes (test builds only).  Compiled with -DORC_PLANT_<NAME> the
 * oracle carries one deliberate, plausible mistake, so that a test can show
 * that its independent check rejects it (oracle/__init__.py: planted()):
 *   KE_NO_HALF        kinetic energy without the 1/2
 *   KE_UNIT_RHO       kinetic energy with rho taken as 1
 *   KE_NO_GUO         kinetic energy from the unshifted velocity m / rho
 *   SOLID_SAME_DIR    bounce from a solid source returns h_i instead of h_opp(i)
 *   SOLID_WRONG_SIDE  the solid test probes x + c_i instead of the source x - c_i
 *   ME_FIELD_FEQ      the moment-encoded variant's collision uses the field-form
 *                     equilibrium (reciprocal multiplies) instead of the division form
 * The default build defines none of them.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define Q 19

static const int CV[Q][3] = {
    {0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},  {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1}, {1, 0, -1},
    {-1, 0, 1}, {0, 1, 1},   {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};

static const int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

static double WT[Q];
static double CS2, INV_CS2, INV_CS4, INV_TWO_CS4, INV_TWO_CS2;
static int g_init = 0;

static void init_constants(void) {
    if (g_init) return;
    /* constants.py:62-69: 1/3, 1/18 x6, 1/36 x12; c_s^2 = 1/3 */
    WT[0] = 1.0 / 3.0;
    for (int i = 1; i <= 6; ++i) WT[i] = 1.0 / 18.0;
    for (int i = 7; i < Q; ++i) WT[i] = 1.0 / 36.0;
    CS2 = 1.0 / 3.0;
    /* equilibrium.py:87-89 and bgk.py:92-93 */
    INV_CS2 = 1.0 / CS2;
    INV_CS4 = 1.0 / (CS2 * CS2);
    INV_TWO_CS4 = 1.0 / (2.0 * CS2 * CS2);
    INV_TWO_CS2 = 1.0 / (2.0 * CS2);
    g_init = 1;
}

/* --------------------------------------------------------------------- */
/* constants accessors                                                   */
/* --------------------------------------------------------------------- */
void orc_constants(int32_t* c_out /*19x3*/, double* w_out /*19*/, int32_t* opp_out /*19*/,
                   double* cs2_out) {
    init_constants();
    for (int i = 0; i < Q; ++i) {
        for (int a = 0; a < 3; ++a) c_out[3 * i + a] = CV[i][a];
        w_out[i] = WT[i];
        opp_out[i] = OPP[i];
    }
    *cs2_out = CS2;
}

/* equilibrium.py:34-53 -- point form, division operands as written there. */
static inline double feq_point_one(int i, double r, double ux, double uy, double uz, double u_sq) {
    double cu = (double)CV[i][0] * ux + (double)CV[i][1] * uy + (double)CV[i][2] * uz;
    return WT[i] * r * (1.0 + cu / CS2 + (cu * cu) / (2.0 * CS2 * CS2) - u_sq / (2.0 * CS2));
}

void orc_feq_point(double rho, double ux, double uy, double uz, double* out) {
    init_constants();
    double u_sq = ux * ux + uy * uy + uz * uz;
    for (int i = 0; i < Q; ++i) out[i] = feq_point_one(i, rho, ux, uy, uz, u_sq);
}

/* equilibrium.py:71-98 -- field form, reciprocal operands. */
static inline double feq_one(int i, double r, double ux, double uy, double uz, double u_sq) {
    double cu = (double)CV[i][0] * ux + (double)CV[i][1] * uy + (double)CV[i][2] * uz;
    return WT[i] * r * (1.0 + cu * INV_CS2 + cu * cu * INV_TWO_CS4 - u_sq * INV_TWO_CS2);
}

void orc_feq_field(int64_t ncell, const double* rho, const double* u, double* out) {
    init_constants();
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < ncell; ++c) {
        double ux = u[c], uy = u[ncell + c], uz = u[2 * ncell + c];
        /* (u * u).sum(axis=0): sequential over the 3 components */
        double u_sq = ux * ux + uy * uy + uz * uz;
        for (int i = 0; i < Q; ++i) out[i * ncell + c] = feq_one(i, rho[c], ux, uy, uz, u_sq);
    }
}

/* equilibrium.py:101-103 and :106-113 */
static inline void moments_one(const double* fc, double* r, double* mx, double* my, double* mz) {
    double a = 0.0, bx = 0.0, by = 0.0, bz = 0.0;
    for (int d = 0; d < Q; ++d) {
        double fd = fc[d];
        a = a + fd;
        bx = bx + (double)CV[d][0] * fd;
        by = by + (double)CV[d][1] * fd;
        bz = bz + (double)CV[d][2] * fd;
    }
    *r = a;
    *mx = bx;
    *my = by;
    *mz = bz;
}

static inline void gather_cell(const double* f, int64_t ncell, int64_t c, double* fc) {
    for (int d = 0; d < Q; ++d) fc[d] = f[d * ncell + c];
}

void orc_density_field(int64_t ncell, const double* f, double* rho) {
    init_constants();
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < ncell; ++c) {
        double fc[Q], r, mx, my, mz;
        gather_cell(f, ncell, c, fc);
        moments_one(fc, &r, &mx, &my, &mz);
        rho[c] = r;
    }
}

void orc_momentum_field(int64_t ncell, const double* f, double* mom) {
    init_constants();
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < ncell; ++c) {
        double fc[Q], r, mx, my, mz;
        gather_cell(f, ncell, c, fc);
        moments_one(fc, &r, &mx, &my, &mz);
        mom[c] = mx;
        mom[ncell + c] = my;
        mom[2 * ncell + c] = mz;
    }
}

static inline double rho_floor(double r) {
    /* np.maximum(rho, 1e-30): NaN propagates */
    if (r != r) return r;
    return r > 1e-30 ? r : 1e-30;
}

/* bgk.py:129-142.  force may be NULL. */
void orc_macroscopic_velocity(int64_t ncell, const double* f, const double* force, double* u) {
    init_constants();
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < ncell; ++c) {
        double fc[Q], r, mx, my, mz;
        gather_cell(f, ncell, c, fc);
        moments_one(fc, &r, &mx, &my, &mz);
        double rs = rho_floor(r);
        if (force) {
            mx = mx + 0.5 * force[c];
            my = my + 0.5 * force[ncell + c];
            mz = mz + 0.5 * force[2 * ncell + c];
        }
        u[c] = mx / rs;
        u[ncell + c] = my / rs;
        u[2 * ncell + c] = mz / rs;
    }
}

/*
 * Collision of one cell: bgk.py:72-110 term by term.
 * has_force selects the Guo branch exactly as `force_lattice is not None`.
 * The Brinkman term (brink != NULL) is an addition with no reference
 * counterpart: the body force at the cell becomes F - s * rho * u_eq0 with
 * s in [0,1] the solid fraction and u_eq0 = m/rho the unforced velocity,
 * the usual penalised (Brinkman) drag.
 */
/* point_feq selects the equilibrium's form: 0 the field form of bgk.py
 * (reciprocal multiplies, equilibrium.py:87-98), 1 the division form of the
 * point function (equilibrium.py:34-53), which is the moment-encoded
 * variant's collision shader's (ME/shaders/lbm_collide.comp:58-65). */
static inline void collide_one_form(const double* fc, double* out, double tau, int has_force, double fx,
                                    double fy, double fz, int point_feq) {
    double r, mx, my, mz;
    moments_one(fc, &r, &mx, &my, &mz);
    double rs = rho_floor(r);
    double ux = mx / rs, uy = my / rs, uz = mz / rs;
    if (has_force) {
        ux = ux + 0.5 * fx / rs;
        uy = uy + 0.5 * fy / rs;
        uz = uz + 0.5 * fz / rs;
    }
    double u_sq = ux * ux + uy * uy + uz * uz;
    double pref = 1.0 - 0.5 / tau;
    for (int i = 0; i < Q; ++i) {
#ifdef ORC_PLANT_ME_FIELD_FEQ
        point_feq = 0;
#endif
        double feq = point_feq ? feq_point_one(i, r, ux, uy, uz, u_sq) : feq_one(i, r, ux, uy, uz, u_sq);
        double fp = fc[i] - (fc[i] - feq) / tau;
        if (has_force) {
            double c0 = (double)CV[i][0], c1 = (double)CV[i][1], c2 = (double)CV[i][2];
            double cdu = c0 * ux + c1 * uy + c2 * uz;
            double tx = (c0 - ux) * INV_CS2 + cdu * c0 * INV_CS4;
            double ty = (c1 - uy) * INV_CS2 + cdu * c1 * INV_CS4;
            double tz = (c2 - uz) * INV_CS2 + cdu * c2 * INV_CS4;
            double ft = pref * WT[i] * (tx * fx + ty * fy + tz * fz);
            fp = fp + ft;
        }
        out[i] = fp;
    }
}

static inline void collide_one(const double* fc, double* out, double tau, int has_force, double fx,
                               double fy, double fz) {
    collide_one_form(fc, out, tau, has_force, fx, fy, fz, 0);
}

static void collide_field(int64_t ncell, const double* f, double* out, double tau, const double* force,
                          int point_feq) {
    init_constants();
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < ncell; ++c) {
        double fc[Q], oc[Q];
        gather_cell(f, ncell, c, fc);
        double fx = 0, fy = 0, fz = 0;
        if (force) {
            fx = force[c];
            fy = force[ncell + c];
            fz = force[2 * ncell + c];
        }
        collide_one_form(fc, oc, tau, force != NULL, fx, fy, fz, point_feq);
        for (int i = 0; i < Q; ++i) out[i * ncell + c] = oc[i];
    }
}

/* The moment-encoded variant's collision (ME/shaders/lbm_collide.comp:36-76):
 * bgk.py's collision with the equilibrium in the division form.  The shader's
 * force is (force_x, 0, 0), entering u_eq as mx / rho + 0.5 force_x / rho and
 * the Guo term as prefactor w_i (term_x force_x); each component of `force`
 * enters as the x component does there, which for (fx, 0, 0) gives the
 * shader's values. */
void orc_me_collide(int64_t ncell, const double* f, double* out, double tau, const double* force) {
    collide_field(ncell, f, out, tau, force, 1);
}

/* The collision half of bgk.py:45-110 over a whole field (no streaming). */
void orc_collide(int64_t ncell, const double* f, double* out, double tau, const double* force) {
    init_constants();
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < ncell; ++c) {
        double fc[Q], oc[Q];
        gather_cell(f, ncell, c, fc);
        double fx = 0, fy = 0, fz = 0;
        if (force) {
            fx = force[c];
            fy = force[ncell + c];
            fz = force[2 * ncell + c];
        }
        collide_one(fc, oc, tau, force != NULL, fx, fy, fz);
        for (int i = 0; i < Q; ++i) out[i * ncell + c] = oc[i];
    }
}

static inline int64_t wrap(int64_t a, int64_t n) {
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}

/* bgk.py:115-126: out[i, x] = in[i, (x - c_i) mod N] (np.roll by +c_i). */
void orc_stream(int32_t nx, int32_t ny, int32_t nz, const double* in, double* out) {
    init_constants();
    int64_t ncell = (int64_t)nx * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t x = 0; x < nx; ++x)
        for (int64_t y = 0; y < ny; ++y)
            for (int64_t z = 0; z < nz; ++z) {
                int64_t c = (x * ny + y) * nz + z;
                for (int i = 0; i < Q; ++i) {
                    int64_t sx = wrap(x - CV[i][0], nx), sy = wrap(y - CV[i][1], ny),
                            sz = wrap(z - CV[i][2], nz);
                    out[i * ncell + c] = in[i * ncell + (sx * ny + sy) * nz + sz];
                }
            }
}

/* bgk.py:45-112: collide (with optional Guo force) then periodic stream. */
void orc_bgk_step(int32_t nx, int32_t ny, int32_t nz, const double* f, double* out, double tau,
                  const double* force, double* scratch /* 19*ncell */) {
    int64_t ncell = (int64_t)nx * ny * nz;
    orc_collide(ncell, f, scratch, tau, force);
    orc_stream(nx, ny, nz, scratch, out);
}

/* bgk.py:145-216, applied to a post-stream field f, result in out. */
void orc_bounce_back_y_walls(int32_t nx, int32_t ny, int32_t nz, const double* f, double* out,
                             const double* uw_top, const double* uw_bot) {
    init_constants();
    int64_t ncell = (int64_t)nx * ny * nz;
    memcpy(out, f, sizeof(double) * Q * ncell);
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t x = 0; x < nx; ++x)
        for (int64_t z = 0; z < nz; ++z) {
            int64_t cb = (x * ny + 0) * nz + z;
            int64_t ct = (x * ny + (ny - 1)) * nz + z;
            for (int i = 0; i < Q; ++i) {
                if (CV[i][1] > 0) {
                    double cdu = (double)CV[i][0] * uw_bot[0] + (double)CV[i][1] * uw_bot[1] +
                                 (double)CV[i][2] * uw_bot[2];
                    double rw = 0.0;
                    for (int d = 0; d < Q; ++d) rw = rw + f[d * ncell + cb];
                    double inj = -2.0 * WT[i] * rw * cdu / CS2;
                    out[i * ncell + cb] = f[OPP[i] * ncell + cb] + inj;
                }
                if (CV[i][1] < 0) {
                    double cdu = (double)CV[i][0] * uw_top[0] + (double)CV[i][1] * uw_top[1] +
                                 (double)CV[i][2] * uw_top[2];
                    double rw = 0.0;
                    for (int d = 0; d < Q; ++d) rw = rw + f[d * ncell + ct];
                    double inj = -2.0 * WT[i] * rw * cdu / CS2;
                    out[i * ncell + ct] = f[OPP[i] * ncell + ct] + inj;
                }
            }
        }
}

/*
 * Reference step composition, NP/sim.py:312-326 / :350-364:
 *   f <- bgk_step(f, tau, force); if ywalls: f <- apply_bounce_back_y_walls(f, top, bottom)
 * Runs `steps` steps in place on f (19*ncell).  work needs 2*19*ncell doubles.
 */
void orc_run_reference(int32_t nx, int32_t ny, int32_t nz, double* f, double tau,
                       const double* force, int32_t ywalls, const double* uw_top,
                       const double* uw_bot, int32_t steps, double* work) {
    int64_t n = (int64_t)Q * nx * ny * nz;
    double* a = work;
    double* b = work + n;
    for (int32_t s = 0; s < steps; ++s) {
        orc_bgk_step(nx, ny, nz, f, a, tau, force, b);
        if (ywalls)
            orc_bounce_back_y_walls(nx, ny, nz, a, f, uw_top, uw_bot);
        else
            memcpy(f, a, sizeof(double) * n);
    }
}

/* ===================================================================== */
/* Additions with no reference counterpart (pinned by closed forms).      */
/* ===================================================================== */

/*
 * Textbook half-way bounce-back (Krueger et al. 2017, sec. 5.3.3-5.3.4),
 * applied as the streaming of a post-collision field h:
 *   g_i(x) = h_i(x - c_i)                                 if the source is a fluid cell
 *   g_i(x) = h_opp(i)(x) + 2 w_i rho0 (c_i . u_w) / c_s^2  if the link crosses a wall face
 *   g_i(x) = h_opp(i)(x)                                  if the source is a solid cell
 * with rho0 = 1.  wall_mask bit k marks face k (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z)
 * as a wall; faces that are not walls are periodic.  A link that crosses
 * several wall faces (edge and corner links) takes the velocity of the
 * crossed face with the highest index.  Solid cells (solid[c] != 0) are
 * set to the rest equilibrium w_i.  face_u is 6x3.
 */
static inline int halfway_source(int64_t x, int64_t y, int64_t z, int i, int32_t nx, int32_t ny,
                                 int32_t nz, uint32_t wall_mask, int64_t* src, int* face) {
    int64_t s[3] = {x - CV[i][0], y - CV[i][1], z - CV[i][2]};
    int64_t n[3] = {nx, ny, nz};
    int crossed = -1;
    for (int a = 0; a < 3; ++a) {
        if (s[a] < 0) {
            if (wall_mask & (1u << (2 * a))) crossed = 2 * a;
            s[a] += n[a];
        } else if (s[a] >= n[a]) {
            if (wall_mask & (1u << (2 * a + 1))) crossed = 2 * a + 1;
            s[a] -= n[a];
        }
    }
    *src = (s[0] * ny + s[1]) * nz + s[2];
    *face = crossed;
    return crossed >= 0;
}

/* The cell whose solid flag decides a pull from src (the source itself). */
static inline int64_t solid_probe(int64_t x, int64_t y, int64_t z, int i, int32_t nx, int32_t ny,
                                  int32_t nz, uint32_t wall_mask, int64_t src) {
#ifdef ORC_PLANT_SOLID_WRONG_SIDE
    int64_t probe;
    int face;
    halfway_source(x, y, z, OPP[i], nx, ny, nz, wall_mask, &probe, &face);
    return probe;
#else
    (void)x, (void)y, (void)z, (void)i, (void)nx, (void)ny, (void)nz, (void)wall_mask;
    return src;
#endif
}

#ifdef ORC_PLANT_SOLID_SAME_DIR
#define SOLID_BOUNCE_DIR(i) (i)
#else
#define SOLID_BOUNCE_DIR(i) OPP[i]
#endif

void orc_stream_halfway(int32_t nx, int32_t ny, int32_t nz, const double* h, double* out,
                        uint32_t wall_mask, const double* face_u, const uint8_t* solid) {
    init_constants();
    int64_t ncell = (int64_t)nx * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t x = 0; x < nx; ++x)
        for (int64_t y = 0; y < ny; ++y)
            for (int64_t z = 0; z < nz; ++z) {
                int64_t c = (x * ny + y) * nz + z;
                if (solid && solid[c]) {
                    for (int i = 0; i < Q; ++i) out[i * ncell + c] = WT[i];
                    continue;
                }
                for (int i = 0; i < Q; ++i) {
                    int64_t src;
                    int face;
                    if (halfway_source(x, y, z, i, nx, ny, nz, wall_mask, &src, &face)) {
                        const double* uw = face_u + 3 * face;
                        double cdu = (double)CV[i][0] * uw[0] + (double)CV[i][1] * uw[1] +
                                     (double)CV[i][2] * uw[2];
                        out[i * ncell + c] = h[OPP[i] * ncell + c] + 2.0 * WT[i] * cdu * INV_CS2;
                    } else if (solid && solid[solid_probe(x, y, z, i, nx, ny, nz, wall_mask, src)]) {
                        out[i * ncell + c] = h[SOLID_BOUNCE_DIR(i) * ncell + c];
                    } else {
                        out[i * ncell + c] = h[i * ncell + src];
                    }
                }
            }
}

/*
 * Brinkman-penalised Guo force for one cell: F_eff = F - s * m, with m the
 * unforced momentum of the pre-collision populations, so that the Guo
 * equilibrium velocity becomes u_eq = (m + F_eff/2)/rho = ((1 - s/2) m + F/2)/rho.
 * s = 1 halves the momentum each step and drives u to F/(2 rho) damping;
 * s = 0 is the plain Guo step.
 */
static inline void brinkman_force(const double* fc, double s, double* fx, double* fy, double* fz) {
    double r, mx, my, mz;
    moments_one(fc, &r, &mx, &my, &mz);
    *fx = *fx - s * mx;
    *fy = *fy - s * my;
    *fz = *fz - s * mz;
}

/*
 * General step of the engine's half-way mode: collide every fluid cell
 * (Guo with per-cell force field `force` (3 x ncell) or uniform force `fu`
 * (3), optional Brinkman field `brink` (ncell) in [0,1]), then
 * orc_stream_halfway.  force, fu, brink and solid may be NULL.
 * work needs 19*ncell doubles.  f holds pre-collision populations.
 */
void orc_run_halfway(int32_t nx, int32_t ny, int32_t nz, double* f, double tau,
                     const double* force, const double* fu, const double* brink,
                     uint32_t wall_mask, const double* face_u, const uint8_t* solid,
                     int32_t steps, double* work) {
    init_constants();
    int64_t ncell = (int64_t)nx * ny * nz;
    int has_force = (force != NULL) || (fu != NULL) || (brink != NULL);
    for (int32_t s = 0; s < steps; ++s) {
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < ncell; ++c) {
            double fc[Q], oc[Q];
            gather_cell(f, ncell, c, fc);
            double fx = 0, fy = 0, fz = 0;
            if (force) {
                fx = force[c];
                fy = force[ncell + c];
                fz = force[2 * ncell + c];
            } else if (fu) {
                fx = fu[0];
                fy = fu[1];
                fz = fu[2];
            }
            if (brink) brinkman_force(fc, brink[c], &fx, &fy, &fz);
            collide_one(fc, oc, tau, has_force, fx, fy, fz);
            for (int i = 0; i < Q; ++i) work[i * ncell + c] = oc[i];
        }
        orc_stream_halfway(nx, ny, nz, work, f, wall_mask, face_u, solid);
    }
}

/*
 * FP64 diagnostics over pre-collision populations: out[0] = total mass
 * sum(rho), out[1..3] = total momentum sum(m) (no force shift),
 * out[4] = kinetic energy 0.5 * sum(rho |u|^2) with u the Guo-corrected
 * velocity (bgk.py:129-142) under the uniform force fu (may be NULL).
 * Sequential sums in cell order.
 */
void orc_diagnostics(int64_t ncell, const double* f, const double* fu, const uint8_t* solid,
                     double* out) {
    init_constants();
    double mass = 0, px = 0, py = 0, pz = 0, ke = 0;
    for (int64_t c = 0; c < ncell; ++c) {
        if (solid && solid[c]) continue;
        double fc[Q], r, mx, my, mz;
        gather_cell(f, ncell, c, fc);
        moments_one(fc, &r, &mx, &my, &mz);
        mass += r;
        px += mx;
        py += my;
        pz += mz;
        double ux = mx, uy = my, uz = mz;
#ifndef ORC_PLANT_KE_NO_GUO
        if (fu) {
            ux += 0.5 * fu[0];
            uy += 0.5 * fu[1];
            uz += 0.5 * fu[2];
        }
#endif
        double rs = rho_floor(r);
        ux /= rs;
        uy /= rs;
        uz /= rs;
#if defined(ORC_PLANT_KE_NO_HALF)
        ke += r * (ux * ux + uy * uy + uz * uz);
#elif defined(ORC_PLANT_KE_UNIT_RHO)
        ke += 0.5 * (ux * ux + uy * uy + uz * uz);
#else
        ke += 0.5 * r * (ux * ux + uy * uy + uz * uz);
#endif
    }
    out[0] = mass;
    out[1] = px;
    out[2] = py;
    out[3] = pz;
    out[4] = ke;
}

/* ===================================================================== */
/* FP64 discrete adjoint of the periodic Guo step (no reference          */
/* counterpart; pinned by central finite differences of the oracle's     */
/* own forward map in the tests).                                         */
/* ===================================================================== */

/*
 * Transposed Jacobian of collide_one (Guo branch, has_force = 1) at the
 * pre-collision populations fc and force (fx,fy,fz):
 *   lam_out_j = sum_i mu_i d(out_i)/d(fc_j)
 *   gF_a     += sum_i mu_i d(out_i)/d(F_a)
 * out_i = fc_i (1 - 1/tau) + feq_i(r, u)/tau + G_i(u, F),
 * u = (m + F/2)/r, feq_i = w_i r P_i(u), G_i = pref w_i sum_a T_ia(u) F_a.
 */
static void collide_one_vjp(const double* fc, double tau, double fx, double fy, double fz,
                            const double* mu, double* lam_out, double* gF) {
    double r, mx, my, mz;
    moments_one(fc, &r, &mx, &my, &mz);
    double F[3] = {fx, fy, fz};
    double u[3] = {(mx + 0.5 * fx) / r, (my + 0.5 * fy) / r, (mz + 0.5 * fz) / r};
    double u_sq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    double pref = 1.0 - 0.5 / tau;
    double itau = 1.0 / tau;
    double a_rho = 0.0;              /* d/d r of sum mu_i feq_i / tau at fixed u */
    double b_u[3] = {0.0, 0.0, 0.0}; /* d/d u_a of sum mu_i (feq_i/tau + G_i) */
    double g_F[3] = {0.0, 0.0, 0.0}; /* explicit d/d F_a of sum mu_i G_i at fixed u */
    for (int i = 0; i < Q; ++i) {
        double c[3] = {(double)CV[i][0], (double)CV[i][1], (double)CV[i][2]};
        double cu = c[0] * u[0] + c[1] * u[1] + c[2] * u[2];
        double P = 1.0 + cu * INV_CS2 + 0.5 * cu * cu * INV_CS4 - 0.5 * u_sq * INV_CS2;
        a_rho += mu[i] * itau * WT[i] * P;
        double cF = c[0] * F[0] + c[1] * F[1] + c[2] * F[2];
        double uF = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
        for (int a = 0; a < 3; ++a) {
            /* dP/du_a = c_a/cs2 + cu c_a/cs4 - u_a/cs2 */
            double dP = c[a] * INV_CS2 + cu * c[a] * INV_CS4 - u[a] * INV_CS2;
            /* G = pref w [ (cF - uF)/cs2 + cu cF/cs4 ];
               dG/du_a = pref w [ -F_a/cs2 + c_a cF/cs4 ] */
            double dG = pref * WT[i] * (-F[a] * INV_CS2 + c[a] * cF * INV_CS4);
            b_u[a] += mu[i] * (itau * WT[i] * r * dP + dG);
            /* dG/dF_a (explicit) = pref w [ (c_a - u_a)/cs2 + cu c_a/cs4 ] */
            g_F[a] += mu[i] * pref * WT[i] * ((c[a] - u[a]) * INV_CS2 + cu * c[a] * INV_CS4);
        }
        (void)uF;
    }
    /* du_a/dfc_j = (c_ja - u_a)/r ; du_a/dF_a = 0.5/r */
    for (int j = 0; j < Q; ++j) {
        double s = (1.0 - itau) * mu[j] + a_rho;
        for (int a = 0; a < 3; ++a) s += b_u[a] * ((double)CV[j][a] - u[a]) / r;
        lam_out[j] = s;
    }
    for (int a = 0; a < 3; ++a) gF[a] += g_F[a] + b_u[a] * 0.5 / r;
}

/*
 * Forward map of the differentiable problem (periodic, per-cell Guo force F):
 *   f^0 = feq_field(1, u0); f^{n+1} = bgk_step(f^n, tau, F)  (n < T)
 *   u^T = macroscopic_velocity(f^T, F)
 *   L = (1/ncell) sum_x |u^T(x) - u_tgt(x)|^2
 * orc_inverse_loss returns L.  orc_inverse_grad returns L and writes dL/du0
 * (3 x ncell) and dL/dF (3 x ncell), storing every forward state (small
 * problems only: (T+1)*19*ncell doubles are allocated).
 */
static double inverse_forward(int32_t nx, int32_t ny, int32_t nz, const double* u0,
                              const double* F, const double* u_tgt, double tau, int32_t T,
                              double* states /* (T+1)*19*ncell or NULL */) {
    int64_t ncell = (int64_t)nx * ny * nz, n = (int64_t)Q * ncell;
    double* rho = (double*)malloc(sizeof(double) * ncell);
    double* f = (double*)malloc(sizeof(double) * n);
    double* tmp = (double*)malloc(sizeof(double) * n);
    double* ut = (double*)malloc(sizeof(double) * 3 * ncell);
    for (int64_t c = 0; c < ncell; ++c) rho[c] = 1.0;
    orc_feq_field(ncell, rho, u0, f);
    if (states) memcpy(states, f, sizeof(double) * n);
    for (int32_t t = 0; t < T; ++t) {
        double* g = (double*)malloc(sizeof(double) * n);
        orc_bgk_step(nx, ny, nz, f, g, tau, F, tmp);
        memcpy(f, g, sizeof(double) * n);
        free(g);
        if (states) memcpy(states + (int64_t)(t + 1) * n, f, sizeof(double) * n);
    }
    orc_macroscopic_velocity(ncell, f, F, ut);
    double L = 0.0;
    for (int64_t c = 0; c < ncell; ++c)
        for (int a = 0; a < 3; ++a) {
            double d = ut[a * ncell + c] - u_tgt[a * ncell + c];
            L += d * d;
        }
    free(rho);
    free(f);
    free(tmp);
    free(ut);
    return L / (double)ncell;
}

double orc_inverse_loss(int32_t nx, int32_t ny, int32_t nz, const double* u0, const double* F,
                        const double* u_tgt, double tau, int32_t T) {
    init_constants();
    return inverse_forward(nx, ny, nz, u0, F, u_tgt, tau, T, NULL);
}

double orc_inverse_grad(int32_t nx, int32_t ny, int32_t nz, const double* u0, const double* F,
                        const double* u_tgt, double tau, int32_t T, double* g_u0, double* g_F) {
    init_constants();
    int64_t ncell = (int64_t)nx * ny * nz, n = (int64_t)Q * ncell;
    double* states = (double*)malloc(sizeof(double) * n * (T + 1));
    double L = inverse_forward(nx, ny, nz, u0, F, u_tgt, tau, T, states);
    double* lam = (double*)malloc(sizeof(double) * n);
    double* mu = (double*)malloc(sizeof(double) * n);
    memset(g_F, 0, sizeof(double) * 3 * ncell);
    /* seed: dL/df^T through u^T = (m + F/2)/r */
    const double* fT = states + (int64_t)T * n;
    for (int64_t c = 0; c < ncell; ++c) {
        double fc[Q], r, mx, my, mz;
        gather_cell(fT, ncell, c, fc);
        moments_one(fc, &r, &mx, &my, &mz);
        double m[3] = {mx, my, mz};
        double gu[3], u[3];
        for (int a = 0; a < 3; ++a) {
            u[a] = (m[a] + 0.5 * F[a * ncell + c]) / r;
            gu[a] = 2.0 * (u[a] - u_tgt[a * ncell + c]) / (double)ncell;
        }
        double gu_dot_u = gu[0] * u[0] + gu[1] * u[1] + gu[2] * u[2];
        for (int j = 0; j < Q; ++j) {
            double s = -gu_dot_u / r;
            for (int a = 0; a < 3; ++a) s += gu[a] * (double)CV[j][a] / r;
            lam[j * ncell + c] = s;
        }
        for (int a = 0; a < 3; ++a) g_F[a * ncell + c] += gu[a] * 0.5 / r;
    }
    /* reverse sweep: f^{t+1} = S(C(f^t)); lam^t = J_C(f^t)^T S^T lam^{t+1} */
    for (int32_t t = T - 1; t >= 0; --t) {
        const double* ft = states + (int64_t)t * n;
        /* S^T: mu_i(x) = lam_i(x + c_i) */
        for (int64_t x = 0; x < nx; ++x)
            for (int64_t y = 0; y < ny; ++y)
                for (int64_t z = 0; z < nz; ++z) {
                    int64_t c = (x * ny + y) * nz + z;
                    for (int i = 0; i < Q; ++i) {
                        int64_t sx = wrap(x + CV[i][0], nx), sy = wrap(y + CV[i][1], ny),
                                sz = wrap(z + CV[i][2], nz);
                        mu[i * ncell + c] = lam[i * ncell + (sx * ny + sy) * nz + sz];
                    }
                }
        for (int64_t c = 0; c < ncell; ++c) {
            double fc[Q], mc[Q], lc[Q], gF[3] = {0, 0, 0};
            gather_cell(ft, ncell, c, fc);
            gather_cell(mu, ncell, c, mc);
            collide_one_vjp(fc, tau, F[c], F[ncell + c], F[2 * ncell + c], mc, lc, gF);
            for (int j = 0; j < Q; ++j) lam[j * ncell + c] = lc[j];
            for (int a = 0; a < 3; ++a) g_F[a * ncell + c] += gF[a];
        }
    }
    /* f^0 = feq(1, u0): dL/du0_a = sum_i lam_i w_i dP_i/du_a (rho = 1) */
    for (int64_t c = 0; c < ncell; ++c) {
        double u[3] = {u0[c], u0[ncell + c], u0[2 * ncell + c]};
        double g[3] = {0, 0, 0};
        for (int i = 0; i < Q; ++i) {
            double cc[3] = {(double)CV[i][0], (double)CV[i][1], (double)CV[i][2]};
            double cu = cc[0] * u[0] + cc[1] * u[1] + cc[2] * u[2];
            for (int a = 0; a < 3; ++a)
                g[a] += lam[i * ncell + c] * WT[i] *
                        (cc[a] * INV_CS2 + cu * cc[a] * INV_CS4 - u[a] * INV_CS2);
        }
        for (int a = 0; a < 3; ++a) g_u0[a * ncell + c] = g[a];
    }
    free(states);
    free(lam);
    free(mu);
    return L;
}

/* ------------------------------------------------------------------------
 * 16-bit moment-encoded state (frontier-moment-encoded variant,
 * packages/lattice-boltzmann-d3q19-frontier-moment-encoded):
 *   basis    src/lbm_me.cpp:83-90 (monomials), :104-150 (Gauss-Jordan inverse)
 *   encode   shaders/lbm_encode.comp:37-51   m = M f, t = clamp((m-lo)/(hi-lo), 0, 1),
 *            code = min(round(t * 65535), 65535)
 *   decode   shaders/lbm_decode.comp:33-43   m = lo + code (hi-lo)/65535, f = M^-1 m
 *   step     src/lbm_me.cpp:376-393          decode, collide, stream + walls, encode
 * Codes are stored moment-major here: codes[k * ncell + cell].
 * ---------------------------------------------------------------------- */
static void me_monomials(int i, double* mono) {
    const double x = CV[i][0], y = CV[i][1], z = CV[i][2];
    const double v[Q] = {1.0,       x,         y,         z,         x * x,     y * y,     z * z,
                         x * y,     x * z,     y * z,     x * x * y, x * x * z, y * y * x, y * y * z,
                         z * z * x, z * z * y, x * x * y * y, x * x * z * z, y * y * z * z};
    for (int k = 0; k < Q; ++k) mono[k] = v[k];
}

/* M (row-major, moments = M f) and its inverse by Gauss-Jordan elimination
 * with partial pivoting; returns max |M M^-1 - I|. */
double orc_me_basis(double* m, double* minv) {
    double a[Q * Q];
    for (int i = 0; i < Q; ++i) {
        double mono[Q];
        me_monomials(i, mono);
        for (int k = 0; k < Q; ++k) m[k * Q + i] = mono[k];
    }
    memcpy(a, m, sizeof(a));
    for (int r = 0; r < Q; ++r)
        for (int c = 0; c < Q; ++c) minv[r * Q + c] = (r == c) ? 1.0 : 0.0;
    for (int col = 0; col < Q; ++col) {
        int piv = col;
        for (int r = col + 1; r < Q; ++r)
            if (fabs(a[r * Q + col]) > fabs(a[piv * Q + col])) piv = r;
        if (piv != col)
            for (int j = 0; j < Q; ++j) {
                double t = a[piv * Q + j];
                a[piv * Q + j] = a[col * Q + j];
                a[col * Q + j] = t;
                t = minv[piv * Q + j];
                minv[piv * Q + j] = minv[col * Q + j];
                minv[col * Q + j] = t;
            }
        const double d = a[col * Q + col];
        for (int j = 0; j < Q; ++j) {
            a[col * Q + j] /= d;
            minv[col * Q + j] /= d;
        }
        for (int r = 0; r < Q; ++r) {
            if (r == col) continue;
            const double f = a[r * Q + col];
            if (f == 0.0) continue;
            for (int j = 0; j < Q; ++j) {
                a[r * Q + j] -= f * a[col * Q + j];
                minv[r * Q + j] -= f * minv[col * Q + j];
            }
        }
    }
    double resid = 0.0;
    for (int r = 0; r < Q; ++r)
        for (int c = 0; c < Q; ++c) {
            double s = 0.0;
            for (int k = 0; k < Q; ++k) s += m[r * Q + k] * minv[k * Q + c];
            const double e = fabs(s - (r == c ? 1.0 : 0.0));
            if (e > resid) resid = e;
        }
    return resid;
}

void orc_me_encode(int64_t ncell, const double* f, const double* lo, const double* hi, uint16_t* codes) {
    double m[Q * Q], minv[Q * Q];
    orc_me_basis(m, minv);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < ncell; ++c)
        for (int k = 0; k < Q; ++k) {
            double mk = 0.0;
            for (int i = 0; i < Q; ++i) mk += m[k * Q + i] * f[(int64_t)i * ncell + c];
            double t = (mk - lo[k]) / (hi[k] - lo[k]);
            t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
            long code = lround(t * 65535.0);
            codes[(int64_t)k * ncell + c] = (uint16_t)(code > 65535 ? 65535 : code);
        }
}

void orc_me_decode(int64_t ncell, const uint16_t* codes, const double* lo, const double* hi, double* f) {
    double m[Q * Q], minv[Q * Q];
    orc_me_basis(m, minv);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < ncell; ++c) {
        double mk[Q];
        for (int k = 0; k < Q; ++k)
            mk[k] = lo[k] + (double)codes[(int64_t)k * ncell + c] * (hi[k] - lo[k]) / 65535.0;
        for (int i = 0; i < Q; ++i) {
            double s = 0.0;
            for (int k = 0; k < Q; ++k) s += minv[i * Q + k] * mk[k];
            f[(int64_t)i * ncell + c] = s;
        }
    }
}

/* The quantized trajectory continued from `codes` (ME/src/lbm_me.cpp:386-409):
 * `steps` times decode (lbm_decode.comp), collide (lbm_collide.comp: the
 * division-form equilibrium), stream and, with ywalls, bounce back
 * (lbm_stream_bounce.comp: a pull, then the half-way y-walls; for resting
 * walls this is bgk.py's stream + apply_bounce_back_y_walls to the bit, and
 * moving walls take apply_bounce_back_y_walls' momentum term), encode
 * (lbm_encode.comp).  The final codes stay in `codes` and f receives them
 * decoded.  work: 2 * 19 * ncell doubles. */
void orc_run_me_codes(int32_t nx, int32_t ny, int32_t nz, uint16_t* codes, double tau, const double* force,
                      int32_t ywalls, const double* uw_top, const double* uw_bot, int32_t steps, const double* lo,
                      const double* hi, double* f, double* work) {
    const int64_t ncell = (int64_t)nx * ny * nz;
    const int64_t n = (int64_t)Q * ncell;
    double* a = work;
    double* b = work + n;
    for (int32_t s = 0; s < steps; ++s) {
        orc_me_decode(ncell, codes, lo, hi, f);
        orc_me_collide(ncell, f, b, tau, force);
        orc_stream(nx, ny, nz, b, a);
        if (ywalls)
            orc_bounce_back_y_walls(nx, ny, nz, a, f, uw_top, uw_bot);
        else
            memcpy(f, a, sizeof(double) * n);
        orc_me_encode(ncell, f, lo, hi, codes);
    }
    orc_me_decode(ncell, codes, lo, hi, f);
}

/* The quantized trajectory from populations: encode f (ME/src/lbm_me.cpp:376-384),
 * then orc_run_me_codes. */
void orc_run_me(int32_t nx, int32_t ny, int32_t nz, double* f, double tau, const double* force,
                int32_t ywalls, const double* uw_top, const double* uw_bot, int32_t steps, const double* lo,
                const double* hi, uint16_t* codes, double* work) {
    orc_me_encode((int64_t)nx * ny * nz, f, lo, hi, codes);
    orc_run_me_codes(nx, ny, nz, codes, tau, force, ywalls, uw_top, uw_bot, steps, lo, hi, f, work);
}

/* Threads of the parallel loops (OpenMP): the current maximum, and a setter
 * so that a caller can time the single-thread restatement. */
int orc_max_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
