// Host-only checks of the launch-shape arithmetic (d3q19_launch.hpp): a
// reservation of reserve_lds(L, k) bytes per block leaves room for exactly k
// blocks on a CU with L bytes of LDS, and stays within 64 KiB per block; the
// plane chunks of a fused step; the ring gap of the moment-coded step; the
// adjoint's memory plan and its automatic checkpoint interval.
#include <cstdio>
#include <initializer_list>

#include "d3q19_launch.hpp"

int main() {
    int bad = 0;
    const unsigned sizes[] = {65536u, 98304u, 163840u};
    for (unsigned lds : sizes) {
        for (unsigned k = 1; k <= 16; ++k) {
            const unsigned b = d3q19::reserve_lds(lds, k);
            if (b == 0) continue;
            if (lds / b != k || b > 65536u) {
                std::printf("lds %u k %u: %u bytes fit %u blocks\n", lds, k, b, lds / b);
                ++bad;
            }
        }
    }
    // the caps the fused step uses on a 160 KiB CU (2 and 3 blocks) are reservable
    for (unsigned k : {2u, 3u})
        if (d3q19::reserve_lds(163840u, k) == 0) {
            std::printf("no reservation for %u blocks on 160 KiB\n", k);
            ++bad;
        }
    // one block per CU would need more than 64 KiB per block: no reservation
    if (d3q19::reserve_lds(163840u, 1) != 0 || d3q19::reserve_lds(0, 3) != 0 || d3q19::reserve_lds(163840u, 0) != 0) {
        std::printf("reservation where none is possible or asked for\n");
        ++bad;
    }
    // plane chunks of a fused step: about target_us of work each once the
    // step takes min_us or more, never more chunks than planes; the chunks
    // of chunk_planes planes cover the x extent exactly once
    struct Case { unsigned long long ncell; unsigned nx; double glups, min_us; unsigned want; };
    const Case cases[] = {
        {256ull * 256 * 256, 256, 41.0, 300.0, 2},    // FP32 256^3: 409 us
        {256ull * 256 * 256, 256, 20.0, 300.0, 4},    // FP64 256^3: 839 us
        {512ull * 512 * 512, 512, 41.0, 300.0, 16},   // FP32 512^3
        {384ull * 384 * 384, 384, 41.0, 300.0, 7},    // FP32 384^3
        {128ull * 128 * 128, 128, 41.0, 300.0, 1},    // FP32 128^3: 51 us
        {256ull * 256 * 256, 256, 82.0, 1000.0, 1},   // FP16 256^3: 205 us
        {384ull * 384 * 384, 384, 82.0, 1000.0, 1},   // FP16 384^3: 690 us
        {512ull * 512 * 512, 512, 82.0, 1000.0, 8},   // FP16 512^3
        {4000000000ull, 3, 20.0, 300.0, 3},           // more work than planes: one chunk per plane
        {256ull * 256 * 256, 256, 0.0, 300.0, 1},     // no rate: one launch
    };
    for (const Case& c : cases) {
        const unsigned n = d3q19::step_chunks(c.ncell, c.nx, c.glups, c.min_us, 210.0);
        if (n != c.want) {
            std::printf("step_chunks(%llu cells, nx %u, %.0f GLUPS) = %u, want %u\n", c.ncell, c.nx, c.glups, n, c.want);
            ++bad;
        }
    }
    if (d3q19::step_chunks(256ull * 256 * 256, 256, 41.0, 300.0, 210.0, 5) != 5 ||
        d3q19::step_chunks(1000, 10, 41.0, 300.0, 210.0, 64) != 10) {
        std::printf("step_chunks ignores or does not clamp the override\n");
        ++bad;
    }
    for (unsigned nx = 1; nx <= 600; ++nx) {
        for (unsigned n = 1; n <= nx; ++n) {
            const unsigned S = d3q19::chunk_planes(nx, n);
            unsigned covered = 0, launches = 0;
            for (unsigned x0 = 0; x0 < nx; x0 += S) {
                covered += (S < nx - x0 ? S : nx - x0);
                ++launches;
            }
            if (covered != nx || launches > n || S == 0) {
                std::printf("chunk_planes(%u, %u) = %u covers %u planes in %u launches\n", nx, n, S, covered, launches);
                ++bad;
            }
        }
    }
    // ring gap of the moment-coded step: twice the planes the resident blocks
    // cover plus two, at most nx
    struct Gap { unsigned nx; unsigned long long bpp, resident; unsigned want; };
    const Gap gaps[] = {
        {256, 128, 1024, 18},   // 256^3: 128 blocks per plane, 256 CUs x 4 blocks
        {96, 32, 1024, 66},     // 96 x 128 x 128
        {512, 512, 1024, 6},    // 512^3
        {64, 2, 1024, 64},      // more resident blocks than the lattice holds: nx (two-lattice footprint)
        {10, 0, 1024, 10},      // no blocks per plane: nx
        {3, 1000, 1, 3},        // clamp at nx
    };
    for (const Gap& g : gaps) {
        const unsigned v = d3q19::ring_gap(g.nx, g.bpp, g.resident);
        if (v != g.want) {
            std::printf("ring_gap(%u, %llu, %llu) = %u, want %u\n", g.nx, g.bpp, g.resident, v, g.want);
            ++bad;
        }
    }
    // the adjoint's memory plan: ceil(T / K) checkpoints, K moment slots, four
    // working states and the fields; the automatic K needs the least of it
    struct Plan { int T; unsigned long long ck, mom, st, fields; int wantK; unsigned long long wantBytes; };
    const Plan plans[] = {
        // 8^3, FP32, 100 steps: rows of 64 padded cells (311,296 B states), 512 cells x 16 B moments
        {100, 311296, 8192, 311296, 4 * 3 * 512 * 4, 50, 2 * 311296ull + 50 * 8192ull + 4 * 311296ull + 24576},
        // 16^3, FP16 checkpoints of FP32 states, 400 steps
        {400, 622592, 65536, 1245184, 4 * 3 * 4096 * 4, 67, 6 * 622592ull + 67 * 65536ull + 4 * 1245184ull + 196608},
        // one step: one checkpoint, one slot
        {1, 1000, 10, 1000, 0, 1, 1000 + 10 + 4000},
    };
    for (const Plan& q : plans) {
        const int K = d3q19::adjoint_least_memory_K(q.T, q.ck, q.mom, q.st, q.fields);
        const unsigned long long b = d3q19::adjoint_plan_bytes(q.T, K, q.ck, q.mom, q.st, q.fields);
        bool least = true;
        for (int k = 1; k <= q.T; ++k)
            if (d3q19::adjoint_plan_bytes(q.T, k, q.ck, q.mom, q.st, q.fields) < b) least = false;
        if (K != q.wantK || b != q.wantBytes || !least) {
            std::printf("adjoint plan T %d: K %d (want %d), %llu B (want %llu), least %d\n", q.T, K, q.wantK, b,
                        q.wantBytes, least);
            ++bad;
        }
    }
    // the benchmarked inverse problem: 128^3 FP32, 2000 steps, K = 50 -> 8.79 GB
    {
        const unsigned long long st = 128ull * 128 * 19 * 128 * 4, mom = 16ull * 128 * 128 * 128;
        const unsigned long long b = d3q19::adjoint_plan_bytes(2000, 50, st, mom, st, 4 * 3 * 4 * 128ull * 128 * 128);
        if (b != 8791261184ull) {
            std::printf("adjoint plan 128^3: %llu B\n", b);
            ++bad;
        }
    }
    std::printf(bad ? "FAIL\n" : "OK\n");
    return bad ? 1 : 0;
}
