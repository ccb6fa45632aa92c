// Writes Philox4x32-10 vectors produced by rocRAND's engine (an implementation
// independent of this repository's) for tests/test_oracle.py.
// rocrand_philox4x32_10 maps (seed, subsequence, offset) to key = seed,
// counter = (offset/4 low, offset/4 high, subsequence low, subsequence high).
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

int main() {
    const unsigned cases[][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {1u, 0u, 0u, 0u, 1729u, 0u},
        {7u, 12345u, 0u, 0u, 1729u, 0u},
        {0xFFFFFFFFu, 99u, 0u, 0u, 1729u, 0u},
        {3u, 5u, 0u, 0u, 0xDEADBEEFu, 0x01234567u},
        {0xFFFFFFFFu, 0x3FFFFFFFu, 0xA5A5A5A5u, 0x5A5A5A5Au, 0xFFFFFFFFu, 0xFFFFFFFFu},
    };
    for (const auto& c : cases) {
        const unsigned long long seed = (unsigned long long)c[4] | ((unsigned long long)c[5] << 32);
        const unsigned long long sub = (unsigned long long)c[2] | ((unsigned long long)c[3] << 32);
        const unsigned long long off = 4ull * ((unsigned long long)c[0] | ((unsigned long long)c[1] << 32));
        rocrand_device::philox4x32_10_engine eng(seed, sub, off);
        uint4 r = eng.next4();
        std::printf("%u %u %u %u %u %u %u %u %u %u\n", c[0], c[1], c[2], c[3], c[4], c[5], r.x, r.y, r.z, r.w);
    }
    return 0;
}
