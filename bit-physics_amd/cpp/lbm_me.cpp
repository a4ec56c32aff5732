// The moment-encoded solver's C++ interface over libd3q19's C ABI.
//
// Follows the reference's moment-encoded package
// (packages/lattice-boltzmann-d3q19-frontier-moment-encoded/src/lbm_me.cpp):
// build_moment_basis :83-136, initial_rest_state :140-146, reference_step
// :148-206, calibrate_ranges :208-253, run_lbm :283-486.  Every lattice
// operation is a libd3q19 call on the default HIP stream: the trajectory, the
// FP64 calibration run and reference_step run in the engine's kernels (FP64
// storage; a quantized trajectory in D3Q19_ME16_REF codes, the variant's own
// FP64 step, or on request D3Q19_ME16); only the frame sums, the witness hash
// and the capture files are host work.
#include "bit_physics/lbm_d3q19_me/lbm_me.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <limits>
#include <stdexcept>

#include "capture_npz.hpp"
#include "d3q19.h"

namespace bit_physics::lbm_d3q19_me {

namespace {

namespace cap = d3q19_capture;

double weight(int i) { return i == 0 ? 1.0 / 3.0 : (i <= 6 ? 1.0 / 18.0 : 1.0 / 36.0); }

size_t ncell(const LbmConfig& c) { return (size_t)c.nx * c.ny * c.nz; }

void check(int rc, const char* what) {
    if (rc != D3Q19_OK) throw std::runtime_error(std::string(what) + ": " + d3q19_last_error());
}

void validate(const LbmConfig& c) {
    if (c.nx < 1 || c.ny < 1 || c.nz < 1) throw std::invalid_argument("nx, ny and nz must be >= 1");
    if (c.nx > 65535 || c.ny > 65535) throw std::invalid_argument("nx and ny must be <= 65535");
    if (!(c.tau > 0.5) || !std::isfinite(c.tau)) throw std::invalid_argument("tau must be a finite number > 0.5");
    if (!std::isfinite(c.force_x)) throw std::invalid_argument("force_x must be finite");
    if (c.capture_interval < 1) throw std::invalid_argument("capture_interval must be >= 1");
    if (!(c.range_margin >= 0.0) || !std::isfinite(c.range_margin))
        throw std::invalid_argument("range_margin must be a finite number >= 0");
}

// One engine lattice of the channel: reference y-walls at rest, periodic x
// and z, uniform Guo force (force_x, 0, 0).
struct Channel {
    d3q19_lattice* h = nullptr;
    Channel(const LbmConfig& c, int32_t storage) {
        d3q19_config cfg;
        d3q19_config_default(&cfg);
        cfg.nx = (int32_t)c.nx;
        cfg.ny = (int32_t)c.ny;
        cfg.nz = (int32_t)c.nz;
        cfg.tau = c.tau;
        cfg.storage = storage;
        cfg.boundary = D3Q19_BC_YWALLS;
        cfg.force_mode = D3Q19_FORCE_UNIFORM;
        cfg.force[0] = c.force_x;
        check(d3q19_create(&cfg, nullptr, &h), "d3q19_create");
    }
    ~Channel() {
        if (h) d3q19_destroy(h);
    }
    Channel(const Channel&) = delete;
    Channel& operator=(const Channel&) = delete;

    void load(const std::vector<double>& f) { check(d3q19_set_populations(h, f.data(), D3Q19_F64, 0), "set_populations"); }
    void step(uint64_t n) {
        while (n) {
            const int32_t k = (int32_t)std::min<uint64_t>(n, (uint64_t)std::numeric_limits<int32_t>::max());
            check(d3q19_step(h, k), "d3q19_step");
            n -= (uint64_t)k;
        }
    }
    void populations(std::vector<double>& f) { check(d3q19_read_populations(h, f.data(), D3Q19_F64, 0), "read_populations"); }
    // stored: the moments of the stored state (what D3Q19_ME16 encodes), else of f^n
    void moment_range(std::array<double, kQ>& lo, std::array<double, kQ>& hi, bool stored = false) {
        if (stored)
            check(d3q19_moment_range_stored(h, lo.data(), hi.data()), "moment_range_stored");
        else
            check(d3q19_moment_range(h, lo.data(), hi.data()), "moment_range");
    }
};

std::string utc_now() {
    const std::time_t now = std::time(nullptr);
    std::tm tm{};
    gmtime_r(&now, &tm);
    char buf[32];
    std::strftime(buf, sizeof buf, "%Y-%m-%dT%H:%M:%SZ", &tm);
    return buf;
}

// Sequential sums over cells, in cell order: mass = sum rho, momentum = sum rho u.
void totals(const StepFrame& fr, size_t n, double& mass, std::array<double, 3>& mom) {
    mass = 0.0;
    mom = {0.0, 0.0, 0.0};
    for (size_t c = 0; c < n; ++c) mass += fr.rho[c];
    for (size_t c = 0; c < n; ++c)
        for (int a = 0; a < 3; ++a) mom[a] += fr.rho[c] * fr.u[a * n + c];
}

void write_capture(const LbmConfig& cfg, const LbmResult& res, const std::filesystem::path& manifest_path,
                   double wall_seconds, const std::string& start_utc) {
    namespace fs = std::filesystem;
    const size_t n = ncell(cfg);
    fs::path dir = manifest_path.parent_path();
    if (dir.empty()) dir = ".";
    fs::create_directories(dir);
    const std::string payload_name = manifest_path.stem().string() + ".npz";
    const fs::path payload = dir / payload_name;
    const std::string variant = cfg.quantize ? "frontier-moment-encoded-16bit" : "stack-c-f64";
    const std::string sim_name = "lattice-boltzmann-d3q19";
    const std::string stack_name = "mi355x-hip";
    const int64_t nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;

    std::vector<std::pair<std::string, cap::NpyArray>> members;
    for (const StepFrame& fr : res.frames) {
        const std::string pre = "steps/" + std::to_string(fr.step) + "/";
        double rmin = std::numeric_limits<double>::infinity(), rmax = -rmin, umax = 0.0;
        for (size_t c = 0; c < n; ++c) {
            rmin = std::min(rmin, fr.rho[c]);
            rmax = std::max(rmax, fr.rho[c]);
        }
        for (double v : fr.u) umax = std::max(umax, std::fabs(v));  // largest |u| component
        members.emplace_back(pre + "state/rho", cap::npy_f64(fr.rho, {nx, ny, nz}));
        members.emplace_back(pre + "state/u", cap::npy_f64(fr.u, {3, nx, ny, nz}));
        members.emplace_back(pre + "diagnostics/rho_min", cap::npy_f64_scalar(rmin));
        members.emplace_back(pre + "diagnostics/rho_max", cap::npy_f64_scalar(rmax));
        members.emplace_back(pre + "diagnostics/u_max_lat", cap::npy_f64_scalar(umax));
    }
    members.emplace_back("metadata/schema_version", cap::npy_str_scalar("1.0.0"));
    members.emplace_back("metadata/sim_name", cap::npy_str_scalar(sim_name));
    members.emplace_back("metadata/sim_category", cap::npy_str_scalar("lattice"));
    members.emplace_back("metadata/sim_variant", cap::npy_str_scalar(variant));
    members.emplace_back("metadata/stack_name", cap::npy_str_scalar(stack_name));
    members.emplace_back("metadata/seed", cap::npy_i64_scalar(cfg.seed));
    cap::write_npz(payload, members);

    using cap::Json;
    Json m = Json::object();
    m["schema_version"] = Json::string("1.0.0");
    m["sim"]["name"] = Json::string(sim_name);
    m["sim"]["category"] = Json::string("lattice");
    m["sim"]["variant"] = Json::string(variant);
    m["stack"]["name"] = Json::string(stack_name);
    m["stack"]["version"] = Json::string("0.1.0");
    m["stack"]["build_id"] = Json::string("libd3q19-gfx950");
    Json& c = m["config"];
    c["tier"] = Json::string("test");
    c["dims"] = Json::array();
    for (int64_t d : {nx, ny, nz}) c["dims"].push(Json::integer(d));
    c["dtype"] = Json::string("f64");
    c["seed"] = Json::integer(cfg.seed);
    c["params"]["tau"] = Json::number(cfg.tau);
    c["params"]["force_x_lattice"] = Json::number(cfg.force_x);
    c["params"]["quantize_bits"] = Json::number(cfg.quantize ? 16.0 : 0.0);
    c["params"]["range_margin"] = Json::number(cfg.range_margin);
    c["params"]["warmup_steps"] = Json::number((double)cfg.warmup_steps);
    m["run"]["step_count"] = Json::integer(cfg.steps);
    m["run"]["capture_interval"] = Json::integer(cfg.capture_interval);
    m["run"]["wall_clock_seconds"] = Json::number(wall_seconds);
    m["run"]["start_utc"] = Json::string(start_utc);
    m["payload"]["format"] = Json::string("npz");
    m["payload"]["path"] = Json::string(payload_name);
    m["payload"]["checksum"] = Json::string("sha256:" + cap::sha256_file_hex(payload));
    m["determinism"]["claimed"] = Json::string("bit-exact-same-hw");
    m["determinism"]["atomic_ops"] = Json::boolean(false);
    m["determinism"]["subgroup_ops"] = Json::boolean(false);
    std::FILE* fh = std::fopen(manifest_path.c_str(), "wb");
    if (!fh) throw std::runtime_error("cannot write " + manifest_path.string());
    const std::string text = m.dump();
    const bool ok = std::fwrite(text.data(), 1, text.size(), fh) == text.size();
    if (std::fclose(fh) != 0 || !ok) throw std::runtime_error("write failed: " + manifest_path.string());
}

}  // namespace

std::array<double, kQ> QuantRanges::bound() const {
    std::array<double, kQ> b{};
    for (int k = 0; k < kQ; ++k) b[k] = (hi[k] - lo[k]) / 2.0 / 65535.0;
    return b;
}

MomentBasis build_moment_basis() {
    int32_t m[kQ * kQ], minv4[kQ * kQ];
    check(d3q19_moment_basis(m, minv4), "d3q19_moment_basis");
    MomentBasis b;
    for (int k = 0; k < kQ * kQ; ++k) {
        b.m[k] = (double)m[k];
        b.m_inv[k] = (double)minv4[k] / 4.0;  // exact: integers over 4
    }
    double worst = 0.0;
    for (int r = 0; r < kQ; ++r)
        for (int c = 0; c < kQ; ++c) {
            double s = 0.0;
            for (int k = 0; k < kQ; ++k) s += b.m[r * kQ + k] * b.m_inv[k * kQ + c];
            worst = std::max(worst, std::fabs(s - (r == c ? 1.0 : 0.0)));
        }
    b.inverse_residual = worst;
    return b;
}

std::vector<double> initial_rest_state(const LbmConfig& cfg) {
    const size_t n = ncell(cfg);
    std::vector<double> f((size_t)kQ * n);
    for (int i = 0; i < kQ; ++i) std::fill(f.begin() + (ptrdiff_t)(i * n), f.begin() + (ptrdiff_t)((i + 1) * n), weight(i));
    return f;
}

void reference_step(std::vector<double>& f, const LbmConfig& cfg) {
    validate(cfg);
    if (f.size() != (size_t)kQ * ncell(cfg))
        throw std::invalid_argument("reference_step: f must hold 19 * nx * ny * nz values");
    Channel lat(cfg, D3Q19_F64);
    lat.load(f);
    lat.step(1);
    lat.populations(f);
}

namespace {
// The variant's range rule over the FP64 run's f^n, or (stored) over its
// stored post-collision states, the states a D3Q19_ME16 lattice encodes.
QuantRanges calibrate(const LbmConfig& cfg, bool stored) {
    Channel lat(cfg, D3Q19_F64);
    lat.load(initial_rest_state(cfg));
    QuantRanges r;
    lat.moment_range(r.lo, r.hi, stored);
    std::array<double, kQ> lo, hi;
    // The ranges must cover the whole run, not only the warm-up: a start-up
    // flow keeps accelerating, and a clamped momentum moment biases u.
    const uint32_t horizon = std::max(cfg.warmup_steps, cfg.steps);
    for (uint32_t s = 0; s < horizon; ++s) {
        lat.step(1);
        lat.moment_range(lo, hi, stored);
        for (int k = 0; k < kQ; ++k) {
            r.lo[k] = std::min(r.lo[k], lo[k]);
            r.hi[k] = std::max(r.hi[k], hi[k]);
        }
    }
    for (int k = 0; k < kQ; ++k) {
        const double span = std::max(r.hi[k] - r.lo[k], 1e-12);
        r.lo[k] -= cfg.range_margin * span;
        r.hi[k] += cfg.range_margin * span;
    }
    return r;
}
}  // namespace

QuantRanges calibrate_ranges(const LbmConfig& cfg, const MomentBasis& basis) {
    validate(cfg);
    if (basis.m != build_moment_basis().m)
        throw std::invalid_argument("calibrate_ranges: the engine's codec needs build_moment_basis()'s basis");
    return calibrate(cfg, false);
}

LbmResult run_lbm(const LbmConfig& cfg, const std::filesystem::path* capture_manifest) {
    return run_lbm(cfg, capture_manifest, QuantArithmetic::reference_fp64);
}

LbmResult run_lbm(const LbmConfig& cfg, const std::filesystem::path* capture_manifest, QuantArithmetic arithmetic) {
    validate(cfg);
    const int32_t quantized_storage = arithmetic == QuantArithmetic::reference_fp64 ? D3Q19_ME16_REF : D3Q19_ME16;
    const auto t0 = std::chrono::steady_clock::now();
    const std::string start = utc_now();
    const size_t n = ncell(cfg);
    const MomentBasis basis = build_moment_basis();
    QuantRanges ranges{};
    if (cfg.quantize)
        ranges = arithmetic == QuantArithmetic::reference_fp64 ? calibrate_ranges(cfg, basis) : calibrate(cfg, true);
    const std::vector<double> f0 = initial_rest_state(cfg);

    // One trajectory; frames (0, every capture_interval, the last step) when
    // out is set.  Returns the SHA-256 of the final FP64 populations.
    auto trajectory = [&](LbmResult* out) -> std::string {
        Channel lat(cfg, cfg.quantize ? quantized_storage : D3Q19_F64);
        if (cfg.quantize) check(d3q19_set_moment_ranges(lat.h, ranges.lo.data(), ranges.hi.data()), "set_moment_ranges");
        lat.load(f0);  // ME16: the state starts as the codes of the rest state
        auto snapshot = [&](uint32_t s) {
            StepFrame fr;
            fr.step = s;
            fr.rho.resize(n);
            fr.u.resize(3 * n);
            check(d3q19_read_moments(lat.h, fr.rho.data(), fr.u.data(), D3Q19_F64, 0), "read_moments");
            out->frames.push_back(std::move(fr));
        };
        if (out) {
            snapshot(0);
            uint32_t s = 0;
            while (s < cfg.steps) {
                const uint64_t next_mult = ((uint64_t)s / cfg.capture_interval + 1) * cfg.capture_interval;
                const uint32_t next = (uint32_t)std::min<uint64_t>(next_mult, cfg.steps);
                lat.step(next - s);
                s = next;
                snapshot(s);
            }
        } else {
            lat.step(cfg.steps);
        }
        std::vector<double> f((size_t)kQ * n);
        lat.populations(f);
        return cap::sha256_hex(f.data(), f.size() * sizeof(double));
    };

    LbmResult res;
    const std::string first = trajectory(&res);
    const std::string second = trajectory(nullptr);
    if (first != second)
        throw std::runtime_error("run_lbm: two runs of the same configuration differ (sha256 " + first + " vs " +
                                 second + "): the run is not bit-reproducible");
    res.determinism_witness_sha256 = first;
    totals(res.frames.front(), n, res.total_mass_initial, res.total_momentum_initial);
    totals(res.frames.back(), n, res.total_mass_final, res.total_momentum_final);
    res.ranges_used = ranges;
    if (capture_manifest) {
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        write_capture(cfg, res, *capture_manifest, wall, start);
    }
    return res;
}

}  // namespace bit_physics::lbm_d3q19_me
