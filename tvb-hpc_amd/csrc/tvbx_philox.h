// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) and the Box-Muller
// transform used for every normal variate on the device.  Shared by the
// emitted step kernels and the runtime's RNG kernel; the CPU oracle restates
// the same functions.
#ifndef TVBX_PHILOX_H
#define TVBX_PHILOX_H

__device__ __forceinline__ uint4 tvbx_philox10(uint4 c, uint2 k) {
  // Philox4x32-10 (Salmon et al., SC'11)
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const unsigned hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    const unsigned hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    k.x += 0x9E3779B9u;
    k.y += 0xBB67AE85u;
  }
  return c;
}

__device__ __forceinline__ void tvbx_boxmuller(unsigned a, unsigned b,
                                               float &z0, float &z1) {
  // u in (0, 1], v in [0, 1), exact multiples of 2^-24
  const float u = (float)((a >> 8) + 1u) * 0x1p-24f;
  const float v = (float)(b >> 8) * 0x1p-24f;
  const float r = sqrtf(-2.0f * logf(u));
  float s, c;
  sincospif(2.0f * v, &s, &c);
  z0 = r * c;
  z1 = r * s;
}

#endif
