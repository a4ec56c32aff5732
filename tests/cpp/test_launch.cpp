// Host-only checks of the launch-shape arithmetic (d3q19_launch.hpp): a
// reservation of reserve_lds(L, k) bytes per block leaves room for exactly k
// blocks on a CU with L bytes of LDS, and stays within 64 KiB per block.
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
    std::printf(bad ? "FAIL\n" : "OK\n");
    return bad ? 1 : 0;
}
