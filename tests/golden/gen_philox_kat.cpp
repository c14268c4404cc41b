// Writes Philox4x32-10 known-answer vectors computed by PyTorch's
// at::philox_engine (ATen/core/PhiloxRNGEngine.h, a source header) as JSON.
// Test-time generator only; nothing from the header enters the build.
// Build and run: tests/golden/gen_philox_kat.sh
#include <ATen/core/PhiloxRNGEngine.h>

#include <cstdint>
#include <cstdio>

int main()
{
    const uint32_t cases[][6] = {
        // ctr0..3, key0..1
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
        {7u, 0u, 12345u, 16u, 0x5eed0001u, 0x00c0ffeeu},
        {1u, 2u, 3u, 4u, 5u, 6u},
        {0u, 307199u, 33u, 2u, 0x5eed0001u, 0x00c0ffeeu},
        {0x80000000u, 0x7fffffffu, 1u, 0xfffffffeu, 0x9e3779b9u, 0xbb67ae85u},
    };
    std::printf("[\n");
    const int n = sizeof(cases) / sizeof(cases[0]);
    for (int i = 0; i < n; ++i) {
        const uint32_t* c = cases[i];
        const uint64_t seed = (uint64_t)c[4] | ((uint64_t)c[5] << 32);
        const uint64_t subseq = (uint64_t)c[2] | ((uint64_t)c[3] << 32);
        const uint64_t offset = (uint64_t)c[0] | ((uint64_t)c[1] << 32);
        at::philox_engine eng(seed, subseq, offset);
        uint32_t out[4];
        for (int k = 0; k < 4; ++k) {
            out[k] = eng();
        }
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n", c[0], c[1],
                    c[2], c[3], c[4], c[5], out[0], out[1], out[2], out[3], i + 1 < n ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
