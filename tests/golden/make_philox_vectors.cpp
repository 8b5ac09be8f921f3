// Writes Philox4x32-10 known-answer vectors from rocRAND's host-callable
// engine (an implementation independent of this repository's), used to pin
// the oracle's and the kernels' Philox.  Build and run:
//   hipcc -O1 -o /tmp/mkphilox tests/golden/make_philox_vectors.cpp
//   /tmp/mkphilox > tests/golden/philox4x32_10_rocrand.txt
// Each line: ctr0 ctr1 ctr2 ctr3 key0 key1 out0 out1 out2 out3 (hex).
// rocRAND's engine keys with (seed lo, seed hi); its counter is
// (offset/4 as 64 bits, subsequence as 64 bits), so engine(seed, subseq,
// 4*offset).next4() = philox4x32_10(ctr, key).
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

int main() {
  const uint32_t cases[][6] = {
      {0, 0, 0, 0, 0, 0},
      {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
      {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
      {1, 0, 0, 0, 0, 0},
      {0, 1, 0, 0, 0, 0},
      {0, 0, 1, 0, 0, 0},
      {0, 0, 0, 1, 0, 0},
      {0, 0, 0, 0, 1, 0},
      {0, 0, 0, 0, 0, 1},
      {12345, 7, 99, 3, 42, 0},
      {0xdeadbeefu, 0x01000005u, 8191, 0, 0x9e3779b9u, 0x7f4a7c15u},
  };
  for (const auto &c : cases) {
    unsigned long long seed = (unsigned long long)c[4] | ((unsigned long long)c[5] << 32);
    unsigned long long subseq = (unsigned long long)c[2] | ((unsigned long long)c[3] << 32);
    unsigned long long off = (unsigned long long)c[0] | ((unsigned long long)c[1] << 32);
    rocrand_device::philox4x32_10_engine eng(seed, 0, 0);
    // position the counter exactly: restart at subsequence, then skip whole
    // 4-word blocks (discard_subsequence / discard operate on the counter)
    eng.discard_subsequence(subseq);
    // discard(k) advances the 128-bit counter by k/4; add `off` in chunks
    // small enough that 4*chunk does not wrap 64 bits
    const unsigned long long chunk = (1ull << 62) - 1;
    for (unsigned long long left = off; left > 0;) {
      unsigned long long step = left < chunk ? left : chunk;
      eng.discard(step * 4ull);
      left -= step;
    }
    uint4 r = eng.next4();
    printf("%08x %08x %08x %08x %08x %08x %08x %08x %08x %08x\n", c[0], c[1],
           c[2], c[3], c[4], c[5], r.x, r.y, r.z, r.w);
  }
  return 0;
}
