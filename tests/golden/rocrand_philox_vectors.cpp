// Writes Philox4x32-10 known answers from rocRAND's host-callable
// philox4x32_10_engine (ROCm /opt/rocm/include/rocrand) as JSON on stdout.
// engine(seed=key, subsequence=(c3<<32|c2), offset=4*((c1<<32)|c0)) yields
// philox(counter=(c0,c1,c2,c3), key) as its next four outputs.
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

int main()
{
    const uint32_t cases[][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {1u, 0u, 0u, 0u, 0u, 0u},
        {0u, 0u, 4u, 0u, 20221u, 0u},
        {999u, 123u, 4u, 0u, 20221u, 0u},
        {12345u, 678u, 2u, 0u, 0xDEADBEEFu, 0x01234567u},
        {0xFFFFFFFEu, 0u, 3u, 0u, 7u, 9u},
        {77u, 0x3FFFFFFFu, 1u, 0u, 0xFFFFFFFFu, 0xFFFFFFFFu},
        {0xFFFFFFFFu, 5u, 0xFFFFFFFFu, 0xFFFFFFFFu, 1u, 2u},
    };
    // offset is 4 * ((c1 << 32) | c0) in 64 bits, so c1 must stay below 2^30.
    printf("[\n");
    const int nc = sizeof(cases) / sizeof(cases[0]);
    for (int k = 0; k < nc; ++k) {
        const uint32_t *c = cases[k];
        const unsigned long long seed = ((unsigned long long)c[5] << 32) | c[4];
        const unsigned long long sub = ((unsigned long long)c[3] << 32) | c[2];
        const unsigned long long off = 4ull * (((unsigned long long)c[1] << 32) | c[0]);
        rocrand_device::philox4x32_10_engine e(seed, sub, off);
        uint32_t o[4];
        for (int i = 0; i < 4; ++i) o[i] = e.next();
        printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
               c[0], c[1], c[2], c[3], c[4], c[5], o[0], o[1], o[2], o[3], k + 1 < nc ? "," : "");
    }
    printf("]\n");
    return 0;
}
