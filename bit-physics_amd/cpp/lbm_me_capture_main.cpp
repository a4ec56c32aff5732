// lbm_me_capture <out_manifest.json> [--f64] [--steps N]
//
// Writes the moment-encoded Poiseuille capture at the default configuration
// (64 x 32 x 3, tau 0.7, force_x 1e-5, 1000 steps, every step captured), as the
// reference's capture tool does
// (packages/lattice-boltzmann-d3q19-frontier-moment-encoded/src/lbm_me_capture_main.cpp:10-31):
// --f64 runs FP64 populations instead of the 16-bit codes, --steps changes
// the horizon.  Prints frames, witness and the initial and final mass.
#include <cstdlib>
#include <cstring>
#include <exception>
#include <iostream>

#include "bit_physics/lbm_d3q19_me/lbm_me.hpp"

int main(int argc, char** argv) {
    namespace me = bit_physics::lbm_d3q19_me;
    if (argc < 2) {
        std::cerr << "usage: lbm_me_capture <out_manifest.json> [--f64] [--steps N]\n";
        return 2;
    }
    me::LbmConfig cfg;
    const std::filesystem::path out = argv[1];
    for (int i = 2; i < argc; ++i) {
        if (std::strcmp(argv[i], "--f64") == 0) {
            cfg.quantize = false;
        } else if (std::strcmp(argv[i], "--steps") == 0 && i + 1 < argc) {
            char* end = nullptr;
            const long v = std::strtol(argv[++i], &end, 10);
            if (!end || *end || v < 0 || v > 0x7fffffffL) {
                std::cerr << "lbm_me_capture: --steps needs a count in [0, 2^31)\n";
                return 2;
            }
            cfg.steps = (uint32_t)v;
        } else {
            std::cerr << "lbm_me_capture: unknown argument " << argv[i] << "\n";
            return 2;
        }
    }
    try {
        const me::LbmResult r = me::run_lbm(cfg, &out);
        std::cout.precision(17);
        std::cout << "frames=" << r.frames.size() << " witness=" << r.determinism_witness_sha256
                  << " mass0=" << r.total_mass_initial << " massT=" << r.total_mass_final << "\n";
    } catch (const std::exception& e) {
        std::cerr << "lbm_me_capture: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
