// Counter-based random numbers for the sampling kernels.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with the
// published multipliers and Weyl key increments.  Every draw is keyed by
// (counter = {item, sweep, chain, purpose}, key = seed), so a draw depends only
// on where it is used, never on launch shape or thread order: runs are
// reproducible and a chain can be resumed from (seed, sweep).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pimc {

enum Purpose : uint32_t {
    kLevy = 1,       // normals of the interior beads (i, i+1), i odd, of a staging segment,
                     // item = particle * M + slice of bead i: words x, y for bead i, z, w for i + 1
    kLevyAccept = 2, // Metropolis uniform of one staging segment, item = particle * M + start slice
    kCom = 3,        // displacement (words 0..dim-1) and Metropolis uniform (word 2), item = particle
    kOffset = 4,     // staging offset of one chain in one sweep, item = 0
    kInitStart = 5,  // first bead of a particle's initial path (words 0..dim-1), item = particle
    kInitLevy = 6,   // normals of the initial path, pairs (i, i+1) with i odd, item = particle * M + i
    kSwap = 7,       // swap try a of a sweep: slice, first particle, partner uniform, Metropolis uniform; item = a
    kSwapLevy = 8,   // normals of swap try a's path w (0, 1), pairs (i, i+1) with i odd, item = (2a + w) M + i
    kHardSphere = 9, // redraw t >= 2 of a bridge bead (words x, y), item = particle * M + slice; the purpose
                     // word is kHardSphere | t << 8
    kHardSphereCom = 10, // redraw t >= 2 of a COM shift (words 0..dim-1), item = particle; kHardSphereCom | t << 8
    kHardSphereSwap = 11, // redraw t >= 2 of bead i of swap try a's path w (words x, y), item = (2a + w) M + i;
                          // kHardSphereSwap | t << 8
};

struct U4 {
    uint32_t x, y, z, w;
};

// a ^ b ^ c; on the device one v_bitop3_b32 (truth table 0x96) instead of two XORs.
__host__ __device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

__host__ __device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c.x;
        const uint64_t p1 = (uint64_t)M1 * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = U4{xor3(hi1, c.y, k0), lo1, xor3(hi0, c.w, k1), lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

__host__ __device__ __forceinline__ U4 draw(uint64_t seed, uint32_t item, uint32_t sweep, uint32_t chain,
                                            uint32_t purpose)
{
    return philox4x32_10(U4{item, sweep, chain, purpose}, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// (x + 1/2) 2^-32: a uniform in (0, 1) that is exact in double precision.
__host__ __device__ __forceinline__ double u01d(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }
__device__ __forceinline__ float u01f(uint32_t x) { return ((float)x + 0.5f) * (1.0f / 4294967296.0f); }

template <typename T> struct Uniform;
template <> struct Uniform<double> {
    __device__ __forceinline__ static double get(uint32_t x) { return u01d(x); }
};
template <> struct Uniform<float> {
    __device__ __forceinline__ static float get(uint32_t x) { return u01f(x); }
};

// Box-Muller: two uniforms -> two independent standard normals.
__device__ __forceinline__ void box_muller(uint32_t x0, uint32_t x1, double &n0, double &n1)
{
    const double u0 = u01d(x0), u1 = u01d(x1);
    const double rad = sqrt(-2.0 * log(u0));
    double s, c;
    sincospi(2.0 * u1, &s, &c);
    n0 = rad * c;
    n1 = rad * s;
}

__device__ __forceinline__ void box_muller(uint32_t x0, uint32_t x1, float &n0, float &n1)
{
    const float u0 = u01f(x0), u1 = u01f(x1);
    // v_log_f32 is log2; v_sin/v_cos take revolutions, so cos(2 pi u1) needs no scaling.
    const float rad = __builtin_amdgcn_sqrtf(-2.0f * 0.69314718055994530942f * __builtin_amdgcn_logf(u0));
    n0 = rad * __builtin_amdgcn_cosf(u1);
    n1 = rad * __builtin_amdgcn_sinf(u1);
}

} // namespace pimc
