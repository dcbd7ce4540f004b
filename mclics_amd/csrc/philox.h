// Philox4x32-10 counter-based generator (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11), written from the paper. The generator has no state: each Metropolis trial
// draws its uniforms from (lattice key, site, sweep), so a sweep reads and writes nothing but the directors.
// Used by the device kernels and by the host program (random initial condition).
#ifndef MCL_PHILOX_H_
#define MCL_PHILOX_H_

#include <stdint.h>

#if defined(__HIPCC__)
#define MCL_HD __host__ __device__ __forceinline__
#else
#define MCL_HD inline
#endif

namespace mcl {

struct Philox4 {
  uint32_t v[4];
};

MCL_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

// a ^ b ^ c; on gfx950 one v_bitop3_b32 (truth table 0x96) instead of two v_xor_b32.
MCL_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

// Ten rounds; round: (hi0, lo0) = M0 * c0, (hi1, lo1) = M1 * c2, c = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0);
// the key advances by the Weyl constants between rounds. Each 32x32 -> 64-bit product is one multiply (on gfx950
// a single v_mad_u64_u32 gives both halves, where a separate high and low multiply would cost two).
MCL_HD Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = xor3(hi1, c1, k0);
    c1 = lo1;
    c2 = xor3(hi0, c3, k1);
    c3 = lo0;
    k0 += W0;
    k1 += W1;
  }
  Philox4 out;
  out.v[0] = c0;
  out.v[1] = c1;
  out.v[2] = c2;
  out.v[3] = c3;
  return out;
}

// Uniform in the open interval (0, 1): (2m + 1) 2^-24 for the top 23 bits m of x. Every value is exact in FP32,
// the smallest is 2^-24 and the largest 1 - 2^-24, and 2u - 1 is symmetric about 0.
MCL_HD float u01_open(uint32_t x) { return (float)(x >> 9) * (1.0f / 8388608.0f) + (1.0f / 16777216.0f); }

// Uniform in (0, 1) from 16 bits h: (2h + 1) 2^-17, exact in FP32 and symmetric under u -> 1 - u.
MCL_HD float u01_open16(uint32_t h) { return (float)h * (1.0f / 65536.0f) + (1.0f / 131072.0f); }

}  // namespace mcl

#endif
