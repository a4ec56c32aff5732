// C++ interface of the D3Q19 moment-encoded (16-bit) Poiseuille solver, run on
// one MI355X through libd3q19.
//
// Same namespace, types, fields, defaults and functions as the reference's
// moment-encoded package
// (packages/lattice-boltzmann-d3q19-frontier-moment-encoded/include/bit_physics/
// lbm_d3q19_me/lbm_me.hpp:30-117), so a caller switches by pointing its include
// path at this directory and linking libd3q19_me.so instead of the Vulkan build.
//
// The case is a channel between the reference's static y-walls
// (apply_bounce_back_y_walls with zero wall velocity), periodic in x and z,
// driven by a uniform Guo body force along +x, from rest (f = w_i).
// quantize = true keeps the state as 19 moments m = M f per cell, each a
// 16-bit code over a range calibrated on an FP64 run of the same trajectory;
// quantize = false runs FP64 populations.
//
// All lattice work runs in libd3q19's HIP kernels.  Errors throw
// std::invalid_argument (bad configuration) or std::runtime_error (a failed
// library call, with d3q19_last_error()'s text; or runs that are not bit
// identical).
#pragma once

#include <array>
#include <cstdint>
#include <filesystem>
#include <string>
#include <vector>

namespace bit_physics::lbm_d3q19_me {

inline constexpr int kQ = 19;

// D3Q19 velocities in the lexicographic order of the NumPy reference
// (reference/constants.py:39-59): rest, +-x, +-y, +-z, then the xy, xz and
// yz edges.
inline constexpr std::array<std::array<int, 3>, kQ> kVelocities = {{
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},   {0, 0, -1},
    {1, 1, 0},   {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1}, {1, 0, -1},
    {-1, 0, 1},  {0, 1, 1},   {0, -1, -1}, {0, 1, -1}, {0, -1, 1},
}};

// kVelocities[kOpposite[i]] == -kVelocities[i]: odd i pairs with i + 1.
inline constexpr std::array<int, kQ> kOpposite = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10,
                                                  9, 12, 11, 14, 13, 16, 15, 18, 17};

struct LbmConfig {
    uint32_t nx = 64, ny = 32, nz = 3;
    uint32_t steps = 1000;
    uint32_t capture_interval = 1;  // frames at 0, every capture_interval steps, and the last step
    double tau = 0.7;
    double force_x = 1e-5;          // Guo body force along +x, lattice units
    int seed = 42;                  // written to the capture; the initial state is the rest state
    bool quantize = true;           // 16-bit moment codes (true) or FP64 populations (false)
    double range_margin = 0.25;     // calibrated ranges are widened by this fraction of their span
    uint32_t warmup_steps = 64;     // calibration covers max(warmup_steps, steps) steps
};

struct MomentBasis {
    std::array<double, kQ * kQ> m{};      // row-major: moment k = sum_i m[k*19 + i] f_i
    std::array<double, kQ * kQ> m_inv{};  // its inverse
    double inverse_residual = 0.0;        // max |M M^-1 - I|
};

struct QuantRanges {
    std::array<double, kQ> lo{};
    std::array<double, kQ> hi{};
    // Half a code step per moment, (hi - lo) / 2 / 65535: the largest
    // distance of a value in [lo, hi] from its decoded code.
    std::array<double, kQ> bound() const;
};

struct StepFrame {
    uint32_t step = 0;
    std::vector<double> rho;  // nx*ny*nz, cell (x*ny + y)*nz + z
    std::vector<double> u;    // 3*nx*ny*nz, component-major (3, nx, ny, nz), Guo-corrected
};

struct LbmResult {
    std::vector<StepFrame> frames;
    std::string determinism_witness_sha256;  // hex SHA-256 of the final FP64 populations
    double total_mass_initial = 0.0;
    double total_mass_final = 0.0;
    std::array<double, 3> total_momentum_initial{};
    std::array<double, 3> total_momentum_final{};
    QuantRanges ranges_used{};
};

// The monomial moment basis {1, cx, cy, cz, cx^2, cy^2, cz^2, cx cy, cx cz,
// cy cz, cx^2 cy, cx^2 cz, cy^2 cx, cy^2 cz, cz^2 cx, cz^2 cy, cx^2 cy^2,
// cx^2 cz^2, cy^2 cz^2} over the 19 velocities, and its inverse (entries are
// multiples of 1/4, held exactly).
MomentBasis build_moment_basis();

// Per-moment ranges over the FP64 trajectory of max(warmup_steps, steps)
// steps from rest (every frame, the initial one included), widened by
// range_margin times the span (a span below 1e-12 counts as 1e-12).  basis
// must be build_moment_basis()'s: the engine's codec is built on it.
QuantRanges calibrate_ranges(const LbmConfig& cfg, const MomentBasis& basis);

// One FP64 step (BGK + Guo force along x, periodic streaming, static y-wall
// bounce-back) of f, (19, nx, ny, nz) direction-major, in place.
void reference_step(std::vector<double>& f, const LbmConfig& cfg);

// f = feq(rho = 1, u = 0) = w_i in every cell, (19, nx, ny, nz).
std::vector<double> initial_rest_state(const LbmConfig& cfg);

// Run the trajectory twice and check that the final populations agree bit
// for bit (their SHA-256 is the witness), keeping the frames of the first
// run.  With capture_manifest, writes the capture: the manifest JSON at that
// path and the payload (an .npz archive with keys steps/{n}/state/{rho,u},
// steps/{n}/diagnostics/{rho_min,rho_max,u_max_lat}, metadata/...) beside it,
// named after the manifest's stem.
LbmResult run_lbm(const LbmConfig& cfg, const std::filesystem::path* capture_manifest);

// Not in the reference's interface: the arithmetic of a quantized run.
// reference_fp64 (what run_lbm above uses) steps the codes as the reference
// variant's shaders do, decoding to FP64 populations, colliding (division-form
// equilibrium), streaming with the y-wall bounce-back and encoding the result
// (D3Q19_ME16_REF); engine_fp32 uses the engine's fused FP32 moment-coded
// step (D3Q19_ME16: codes of the stored post-collision state, ranges
// calibrated over those states and reported in ranges_used), faster and
// within the lbm-quantized tolerance of the FP64 run.
enum class QuantArithmetic { reference_fp64, engine_fp32 };
LbmResult run_lbm(const LbmConfig& cfg, const std::filesystem::path* capture_manifest, QuantArithmetic arithmetic);

}  // namespace bit_physics::lbm_d3q19_me
