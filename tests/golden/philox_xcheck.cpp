// Independent check of Philox4x32-10: prints rocRAND's (open-source, header-only)
// ten_rounds() output for counter/key pairs read from stdin, one per line:
//   c0 c1 c2 c3 k0 k1   ->   r0 r1 r2 r3
// Compiled host-only with hipcc by tests/test_oracle.py.
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>

struct Engine : rocrand_device::philox4x32_10_engine {
    uint4 rounds(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main() {
    Engine e;
    unsigned c0, c1, c2, c3, k0, k1;
    while (std::scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        const uint4 r = e.rounds(make_uint4(c0, c1, c2, c3), make_uint2(k0, k1));
        std::printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
