// Writes Philox4x32-10 vectors computed by ROCm's open-source rocrand header
// (rocrand_philox4x32_10.h, ten_rounds) for seeded pseudo-random counters and
// keys.  Host code only; it writes philox_rocrand.json:
//   hipcc -std=c++17 -O1 -o xcheck tests/golden/philox_rocrand_xcheck.cpp
//   ./xcheck > tests/golden/philox_rocrand.json
#include <cstdio>
#include <cstdint>
#include <rocrand/rocrand_philox4x32_10.h>

struct probe : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main() {
    probe p;
    uint64_t s = 0x9E3779B97F4A7C15ull;
    auto next = [&]() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return (uint32_t)(s >> 11); };
    printf("{\"source\": \"rocrand_philox4x32_10.h ten_rounds (ROCm 7.2)\", \"vectors\": [\n");
    for (int i = 0; i < 64; i++) {
        uint4 c{next(), next(), next(), next()};
        uint2 k{next(), next()};
        if (i == 0) { c = uint4{0, 0, 0, 0}; k = uint2{0, 0}; }
        uint4 o = p.run(c, k);
        printf("  {\"ctr\": [\"%08x\", \"%08x\", \"%08x\", \"%08x\"], \"key\": [\"%08x\", \"%08x\"], "
               "\"out\": [\"%08x\", \"%08x\", \"%08x\", \"%08x\"]}%s\n",
               c.x, c.y, c.z, c.w, k.x, k.y, o.x, o.y, o.z, o.w, i == 63 ? "" : ",");
    }
    printf("]}\n");
    return 0;
}
