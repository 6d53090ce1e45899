// Known-answer vectors of Philox4x32-10 from ROCm's open rocrand implementation
// (rocrand/rocrand_philox4x32_10.h), an implementation independent of this repository's.
// Build and run: tests/golden/gen_philox_kat.sh  ->  philox_kat.json
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

struct Engine : rocrand_device::philox4x32_10_engine {
    uint4 rounds(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main()
{
    const unsigned ctr[3][4] = {{0, 0, 0, 0},
                                {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
                                {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u}};
    const unsigned key[3][2] = {{0, 0}, {0xffffffffu, 0xffffffffu}, {0xa4093822u, 0x299f31d0u}};
    Engine e;
    std::printf("[\n");
    for (int t = 0; t < 3; ++t) {
        uint4 r = e.rounds(uint4{ctr[t][0], ctr[t][1], ctr[t][2], ctr[t][3]}, uint2{key[t][0], key[t][1]});
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
                    ctr[t][0], ctr[t][1], ctr[t][2], ctr[t][3], key[t][0], key[t][1], r.x, r.y, r.z, r.w,
                    t < 2 ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
