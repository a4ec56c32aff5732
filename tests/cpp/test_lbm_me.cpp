// Tests of the moment-encoded solver's C++ interface (libd3q19_me.so).
//
// Default mode (needs the GPU): the reference's acceptance cases for this
// interface (packages/lattice-boltzmann-d3q19-frontier-moment-encoded/tests/
// test_lbm_me.cpp:47-164) -- constants, basis inverse, code round trip within
// the closed-form bound, rest state fixed under reference_step, FP64 mass and
// momentum conservation, quantized mass within the quantization budget, a
// (tau, force) sweep, quantized vs FP64 -- plus the frame schedule, the
// witness, error paths, and dumps for the Python side to compare against the
// oracle:  test_lbm_me [dump_dir]
//
// --host-only <dir> (no GPU): SHA-256 known answers (FIPS 180-2 appendix B)
// and hashes, an .npz archive and a JSON document written to <dir> for the
// Python side to read back with hashlib / numpy / json.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "bit_physics/lbm_d3q19_me/lbm_me.hpp"
#include "capture_npz.hpp"

using namespace bit_physics::lbm_d3q19_me;

namespace {

int g_failed = 0;

#define CHECK(cond)                                                                     \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            ++g_failed;                                                                 \
            std::fprintf(stderr, "%s:%d: CHECK failed: %s\n", __FILE__, __LINE__, #cond); \
        }                                                                               \
    } while (0)

template <class E>
bool throws(const std::function<void()>& fn) {
    try {
        fn();
    } catch (const E&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

LbmConfig small(bool quantize, double tau = 0.7, double fx = 0.0) {
    LbmConfig c;
    c.nx = 16;
    c.ny = 8;
    c.nz = 3;
    c.steps = 24;
    c.capture_interval = 24;
    c.tau = tau;
    c.force_x = fx;
    c.quantize = quantize;
    c.warmup_steps = 24;
    return c;
}

size_t cells(const LbmConfig& c) { return (size_t)c.nx * c.ny * c.nz; }

void t_constants() {
    const LbmConfig c = small(false);
    const std::vector<double> f = initial_rest_state(c);
    const size_t n = cells(c);
    CHECK(f.size() == 19 * n);
    double w = 0.0;
    for (int i = 0; i < kQ; ++i) w += f[(size_t)i * n + 5];
    CHECK(std::fabs(w - 1.0) <= 1e-15);
    for (int i = 0; i < kQ; ++i) {
        CHECK(kOpposite[kOpposite[i]] == i);
        for (int a = 0; a < 3; ++a) CHECK(kVelocities[i][a] + kVelocities[kOpposite[i]][a] == 0);
    }
}

void t_basis() {
    const MomentBasis b = build_moment_basis();
    CHECK(b.inverse_residual <= 1e-12);
    CHECK(b.inverse_residual == 0.0);  // M^-1 has quarter-integer entries, held exactly
    for (int i = 0; i < kQ; ++i) {
        CHECK(b.m[i] == 1.0);
        for (int a = 0; a < 3; ++a) CHECK(b.m[(1 + a) * kQ + i] == (double)kVelocities[i][a]);
        // the last monomial row: cy^2 cz^2
        const double y = kVelocities[i][1], z = kVelocities[i][2];
        CHECK(b.m[18 * kQ + i] == y * y * z * z);
    }
}

void t_round_trip() {
    const MomentBasis b = build_moment_basis();
    const QuantRanges r = calibrate_ranges(small(true, 0.7, 1e-5), b);
    const auto bound = r.bound();
    for (int k = 0; k < kQ; ++k) {
        CHECK(r.hi[k] > r.lo[k]);
        for (int t = 0; t <= 10; ++t) {
            const double m = r.lo[k] + (r.hi[k] - r.lo[k]) * (t / 10.0);
            const uint32_t code = (uint32_t)std::lround((m - r.lo[k]) / (r.hi[k] - r.lo[k]) * 65535.0);
            const double back = r.lo[k] + code * (r.hi[k] - r.lo[k]) / 65535.0;
            CHECK(std::fabs(back - m) <= bound[k] + 1e-15 * (std::fabs(r.hi[k]) + std::fabs(r.lo[k])));
        }
    }
    CHECK(throws<std::invalid_argument>([&] {
        MomentBasis other = b;
        other.m[0] = 2.0;
        calibrate_ranges(small(true), other);
    }));
}

void t_rest_fixed_point() {
    const LbmConfig c = small(false, 0.7, 0.0);
    std::vector<double> f = initial_rest_state(c);
    const std::vector<double> f0 = f;
    reference_step(f, c);
    double d = 0.0;
    for (size_t k = 0; k < f.size(); ++k) d = std::max(d, std::fabs(f[k] - f0[k]));
    CHECK(d <= 1e-15);
    std::vector<double> bad(7);
    CHECK(throws<std::invalid_argument>([&] { reference_step(bad, c); }));
}

void t_f64_conserves() {
    const LbmConfig c = small(false, 0.8, 0.0);
    const LbmResult r = run_lbm(c, nullptr);
    CHECK(std::fabs(r.total_mass_final - r.total_mass_initial) <= 1e-12 * std::fabs(r.total_mass_initial));
    for (int a = 0; a < 3; ++a) CHECK(std::fabs(r.total_momentum_final[a] - r.total_momentum_initial[a]) <= 1e-12);
    CHECK(r.determinism_witness_sha256.size() == 64);
    CHECK(r.ranges_used.lo == QuantRanges{}.lo && r.ranges_used.hi == QuantRanges{}.hi);
}

void t_quantized_mass_budget() {
    const LbmConfig c = small(true, 0.8, 0.0);
    const LbmResult r = run_lbm(c, nullptr);
    const double budget = (c.steps + 1.0) * (double)cells(c) * r.ranges_used.bound()[0];
    CHECK(std::fabs(r.total_mass_final - r.total_mass_initial) <= budget);
}

void t_sweep() {
    for (double tau : {0.6, 0.7, 1.0})
        for (double fx : {0.0, 1e-5}) {
            LbmConfig c = small(false, tau, fx);
            c.steps = 12;
            const LbmResult r = run_lbm(c, nullptr);
            CHECK(std::fabs(r.total_mass_final - r.total_mass_initial) <= 1e-11 * std::fabs(r.total_mass_initial));
            for (const StepFrame& fr : r.frames)
                for (double rho : fr.rho) CHECK(std::isfinite(rho) && rho > 0.0);
        }
}

void t_quantized_tracks_f64() {
    const LbmResult q = run_lbm(small(true, 0.7, 1e-5), nullptr);
    const LbmResult f = run_lbm(small(false, 0.7, 1e-5), nullptr);
    CHECK(q.frames.size() == f.frames.size());
    double e = 0.0, eu = 0.0;
    for (size_t k = 0; k < q.frames.back().rho.size(); ++k)
        e = std::max(e, std::fabs(q.frames.back().rho[k] - f.frames.back().rho[k]));
    for (size_t k = 0; k < q.frames.back().u.size(); ++k)
        eu = std::max(eu, std::fabs(q.frames.back().u[k] - f.frames.back().u[k]));
    CHECK(e <= 1e-2);
    // The codes do quantize; the witness is u (this channel's density stays 1
    // to within its 1e-12 range floor).
    CHECK(eu > 0.0);
    CHECK(eu <= 1e-2);
    CHECK(q.determinism_witness_sha256 != f.determinism_witness_sha256);
}

// Both arithmetics of a quantized run stay within the lbm-quantized
// tolerance (err <= 1e-5 + 5e-2 |parent|) of the FP64 run on every frame.
void t_quantized_arithmetics() {
    LbmConfig c = small(true, 0.7, 1e-5);
    c.capture_interval = 6;
    const LbmResult r = run_lbm(c, nullptr, QuantArithmetic::reference_fp64);
    const LbmResult e = run_lbm(c, nullptr, QuantArithmetic::engine_fp32);
    const LbmResult d = run_lbm(c, nullptr);  // the two-argument form is reference_fp64
    c.quantize = false;
    const LbmResult f = run_lbm(c, nullptr);
    CHECK(d.determinism_witness_sha256 == r.determinism_witness_sha256);
    CHECK(e.determinism_witness_sha256 != r.determinism_witness_sha256);
    CHECK(r.frames.size() == f.frames.size() && e.frames.size() == f.frames.size());
    bool ok_r = true, ok_e = true;
    for (size_t n = 0; n < f.frames.size() && n < r.frames.size() && n < e.frames.size(); ++n)
        for (size_t k = 0; k < f.frames[n].u.size(); ++k) {
            const double p = f.frames[n].u[k];
            ok_r = ok_r && std::fabs(r.frames[n].u[k] - p) <= 1e-5 + 5e-2 * std::fabs(p);
            ok_e = ok_e && std::fabs(e.frames[n].u[k] - p) <= 1e-5 + 5e-2 * std::fabs(p);
        }
    CHECK(ok_r);
    CHECK(ok_e);
}

void t_frames_and_witness() {
    LbmConfig c = small(true, 0.7, 1e-5);
    c.steps = 10;
    c.capture_interval = 4;
    const LbmResult a = run_lbm(c, nullptr);
    CHECK(a.frames.size() == 4);
    const uint32_t want[4] = {0, 4, 8, 10};
    for (size_t k = 0; k < a.frames.size() && k < 4; ++k) {
        CHECK(a.frames[k].step == want[k]);
        CHECK(a.frames[k].rho.size() == cells(c) && a.frames[k].u.size() == 3 * cells(c));
    }
    const LbmResult b = run_lbm(c, nullptr);
    CHECK(a.determinism_witness_sha256 == b.determinism_witness_sha256);
    c.steps = 0;
    const LbmResult z = run_lbm(c, nullptr);
    CHECK(z.frames.size() == 1 && z.frames[0].step == 0);
}

void t_errors() {
    LbmConfig c = small(false);
    c.capture_interval = 0;
    CHECK(throws<std::invalid_argument>([&] { run_lbm(c, nullptr); }));
    c = small(false);
    c.tau = 0.5;
    CHECK(throws<std::invalid_argument>([&] { run_lbm(c, nullptr); }));
    c = small(false);
    c.ny = 0;
    CHECK(throws<std::invalid_argument>([&] { run_lbm(c, nullptr); }));
}

// Files for the Python side: a reference_step of a perturbed state, and
// captures of a short run in both modes.
void t_dump(const std::filesystem::path& dir) {
    namespace cap = d3q19_capture;
    std::filesystem::create_directories(dir);
    LbmConfig c = small(false, 0.65, 2e-5);
    c.nx = 6;
    c.ny = 5;
    c.nz = 4;
    std::vector<double> f = initial_rest_state(c);
    for (size_t k = 0; k < f.size(); ++k) f[k] *= 1.0 + 0.01 * std::sin(0.37 * (double)k);
    const std::vector<double> f_in = f;
    reference_step(f, c);
    const std::vector<int64_t> shape = {19, (int64_t)c.nx, (int64_t)c.ny, (int64_t)c.nz};
    cap::write_npz(dir / "reference_step.npz", {{"f_in", cap::npy_f64(f_in, shape)},
                                                 {"f_out", cap::npy_f64(f, shape)},
                                                 {"tau", cap::npy_f64_scalar(c.tau)},
                                                 {"force_x", cap::npy_f64_scalar(c.force_x)}});
    LbmConfig p;  // the default channel, shortened
    p.steps = 60;
    p.capture_interval = 7;
    const std::filesystem::path mq = dir / "me16_capture.json", mf = dir / "f64_capture.json";
    const LbmResult rq = run_lbm(p, &mq);
    p.quantize = false;
    const LbmResult rf = run_lbm(p, &mf);
    CHECK(std::filesystem::exists(dir / "me16_capture.npz") && std::filesystem::exists(dir / "f64_capture.npz"));
    std::FILE* fh = std::fopen((dir / "results.txt").c_str(), "w");
    if (fh) {
        for (const auto* r : {&rq, &rf}) {
            std::fprintf(fh, "%s %.17g %.17g %.17g %.17g %zu\n", r->determinism_witness_sha256.c_str(),
                         r->total_mass_initial, r->total_mass_final, r->total_momentum_final[0],
                         r->total_momentum_initial[0], r->frames.size());
        }
        for (int k = 0; k < kQ; ++k) std::fprintf(fh, "%.17g %.17g\n", rq.ranges_used.lo[k], rq.ranges_used.hi[k]);
        std::fclose(fh);
    }
}

int host_only(const std::filesystem::path& dir) {
    namespace cap = d3q19_capture;
    const std::string abc = "abc", two = "abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq";
    const std::string mil(1000000, 'a');
    CHECK(cap::sha256_hex("", 0) == "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855");
    CHECK(cap::sha256_hex(abc.data(), abc.size()) == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad");
    CHECK(cap::sha256_hex(two.data(), two.size()) == "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1");
    CHECK(cap::sha256_hex(mil.data(), mil.size()) == "cdc76e5c9914fb9281a1c7e284d73e67f1809a48a497200e046d39ccc7112cd0");
    std::filesystem::create_directories(dir);
    // hashes at every length 0..200 of a byte ramp, for hashlib to check
    std::vector<unsigned char> ramp(200);
    for (size_t k = 0; k < ramp.size(); ++k) ramp[k] = (unsigned char)(k * 7 + 3);
    std::FILE* fh = std::fopen((dir / "sha256.txt").c_str(), "w");
    if (!fh) return 1;
    for (size_t len = 0; len <= ramp.size(); ++len) std::fprintf(fh, "%zu %s\n", len, cap::sha256_hex(ramp.data(), len).c_str());
    std::fclose(fh);
    std::vector<double> a(6);
    for (int k = 0; k < 6; ++k) a[k] = 0.5 * k - 1.25;
    cap::write_npz(dir / "host.npz", {{"steps/3/state/a", cap::npy_f64(a, {2, 3})},
                                       {"line", cap::npy_f64(a, {6})},
                                       {"scalar", cap::npy_f64_scalar(3.25)},
                                       {"count", cap::npy_i64_scalar(-42)},
                                       {"name", cap::npy_str_scalar("lattice-boltzmann-d3q19")},
                                       {"empty", cap::npy_f64({}, {0})}});
    using cap::Json;
    Json j = Json::object();
    j["b"]["x"] = Json::number(0.7);
    j["b"]["y"] = Json::number(16.0);
    j["b"]["z"] = Json::number(1e-5);
    j["a"] = Json::array();
    j["a"].push(Json::integer(64));
    j["a"].push(Json::boolean(true));
    j["a"].push(Json::string("quote \" back \\ tab \t"));
    j["e"] = Json::object();
    fh = std::fopen((dir / "host.json").c_str(), "w");
    if (!fh) return 1;
    const std::string text = j.dump();
    std::fwrite(text.data(), 1, text.size(), fh);
    std::fclose(fh);
    std::printf("host-only: %s\n", g_failed ? "FAILED" : "ok");
    return g_failed ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 3 && std::strcmp(argv[1], "--host-only") == 0) return host_only(argv[2]);
    struct Case {
        const char* name;
        std::function<void()> fn;
    };
    std::vector<Case> cases = {
        {"constants", t_constants},
        {"moment basis", t_basis},
        {"16-bit round trip within the closed-form bound", t_round_trip},
        {"rest state is a fixed point of reference_step", t_rest_fixed_point},
        {"FP64 run conserves mass and momentum (no force)", t_f64_conserves},
        {"quantized mass within the quantization budget", t_quantized_mass_budget},
        {"(tau, force) sweep: mass conserved, rho finite and positive", t_sweep},
        {"quantized run tracks the FP64 run", t_quantized_tracks_f64},
        {"reference-FP64 and engine-FP32 quantized arithmetic", t_quantized_arithmetics},
        {"frame schedule and witness", t_frames_and_witness},
        {"configuration errors", t_errors},
    };
    if (argc >= 2) {
        const std::filesystem::path dir = argv[1];
        cases.push_back({"dump for the oracle comparison", [dir] { t_dump(dir); }});
    }
    for (const Case& c : cases) {
        std::printf("[case] %s\n", c.name);
        std::fflush(stdout);
        const int before = g_failed;
        try {
            c.fn();
        } catch (const std::exception& e) {
            ++g_failed;
            std::fprintf(stderr, "  exception: %s\n", e.what());
        }
        std::printf("  %s\n", g_failed == before ? "ok" : "FAILED");
        std::fflush(stdout);
    }
    std::printf("%s (%d failed checks)\n", g_failed ? "FAILED" : "all passed", g_failed);
    return g_failed ? 1 : 0;
}
