// Host check of sarf::HostCopier: random sizes and offsets against memcpy,
// back-to-back jobs, and copiers created and destroyed repeatedly.  Built by
// tests/test_hostcopy.py (optionally under ThreadSanitizer).
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "sarf_hostcopy.hpp"

int main() {
    std::mt19937_64 rng(12345);
    std::vector<unsigned char> src(1 << 21), dst(1 << 21), ref(1 << 21);
    for (auto &b : src) b = (unsigned char)rng();
    for (int round = 0; round < 6; round++) {
        sarf::HostCopier cp(1 + round % 8, 1024 + 4096 * (round % 3));
        for (int it = 0; it < 300; it++) {
            const size_t n = rng() % (1 << 20), so = rng() % (1 << 20), d0 = rng() % (1 << 20);
            for (size_t i = 0; i < src.size(); i += 4096) src[i] = (unsigned char)rng();
            ref = dst;
            std::memcpy(ref.data() + d0, src.data() + so, n);
            cp.copy(dst.data() + d0, src.data() + so, n);
            if (std::memcmp(ref.data(), dst.data(), dst.size()) != 0) {
                std::fprintf(stderr, "mismatch: round %d it %d n %zu\n", round, it, n);
                return 1;
            }
        }
    }
    std::printf("hostcopy ok\n");
    return 0;
}
