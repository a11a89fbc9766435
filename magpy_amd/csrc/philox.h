// philox.h -- Philox4x32-10 counter-based generator and the Wiener stream.
//
// Philox4x32-10 follows Salmon, Moraes, Dror & Shaw, "Parallel random numbers:
// as easy as 1, 2, 3" (SC'11).  It replaces the reference's sequential
// std::mt19937_64 stream (lib/rng.cpp:14-24): a particle's draw for any step
// is a pure function of (seed, particle, step), so no generator state lives
// in memory and any step block can be replayed or resumed.
//
// Stream definition (shared by host and device): normal j of particle p of a
// member with seed s is component j%4 of
//   BoxMuller(Philox4x32_10(ctr = (j/4 lo, j/4 hi, p, tag), key = (s lo, s hi)))
// where the four words (w0,w1,w2,w3) give the pairs (w0,w1), (w2,w3) and a
// pair (a,b) gives (r cos 2 pi u(b), r sin 2 pi u(b)), r = sqrt(-2 ln u(a)),
// u(x) = (x + 1/2) 2^-32.  Step n uses normals 3n, 3n+1, 3n+2.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define MP_HD __host__ __device__ inline
#else
#define MP_HD inline
#endif

namespace mp {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

struct u32x4 {
    uint32_t x, y, z, w;
};

// Round keys of one key: k_r = key + r * W, r = 0..9.
struct PhiloxKeys {
    uint32_t k0[10], k1[10];
    MP_HD void init(uint32_t key0, uint32_t key1)
    {
#pragma unroll
        for (int r = 0; r < 10; ++r) {
            k0[r] = key0 + (uint32_t)r * kPhiloxW0;
            k1[r] = key1 + (uint32_t)r * kPhiloxW1;
        }
    }
};

// a ^ b ^ c; on the device one v_bitop3_b32 (truth table 0x96) instead of
// two v_xor_b32, which the compiler does not merge by itself.
MP_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

MP_HD u32x4 philox4x32_10(u32x4 c, const PhiloxKeys &k)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c.x;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c.z;
        u32x4 n;
        n.x = xor3((uint32_t)(p1 >> 32), c.y, k.k0[r]);
        n.y = (uint32_t)p1;
        n.z = xor3((uint32_t)(p0 >> 32), c.w, k.k1[r]);
        n.w = (uint32_t)p0;
        c = n;
    }
    return c;
}

}  // namespace mp
