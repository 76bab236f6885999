// Host-side cross-check of the Philox4x32-10 used by the kernels
// (pimc.jl_amd/csrc/philox.hpp) against rocRAND's header-only engine.
// Reads "c0 c1 c2 c3 k0 k1" lines (hex) on stdin, prints both outputs.
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>

#include "philox.hpp"

struct Engine : rocrand_device::philox4x32_10_engine {
    uint4 rounds(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main()
{
    Engine e;
    unsigned c0, c1, c2, c3, k0, k1;
    while (std::scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        const uint4 r = e.rounds(uint4{c0, c1, c2, c3}, uint2{k0, k1});
        const pimc::U4 p = pimc::philox4x32_10(pimc::U4{c0, c1, c2, c3}, k0, k1);
        std::printf("%08x %08x %08x %08x %08x %08x %08x %08x\n", r.x, r.y, r.z, r.w, p.x, p.y, p.z, p.w);
    }
    return 0;
}
