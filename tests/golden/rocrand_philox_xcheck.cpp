// Host-side cross-check of Philox4x32-10 against rocRAND's open-source header:
// reads lines "c0 c1 c2 c3 k0 k1" (hex) on stdin, prints ten_rounds(counter, key).
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

struct Engine : rocrand_device::philox4x32_10_engine {
    using rocrand_device::philox4x32_10_engine::ten_rounds;
};

int main() {
    Engine eng;
    unsigned c0, c1, c2, c3, k0, k1;
    while (std::scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        uint4 r = eng.ten_rounds(uint4{c0, c1, c2, c3}, uint2{k0, k1});
        std::printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
