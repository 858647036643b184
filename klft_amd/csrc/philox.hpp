// Counter-based random numbers for the klft MI355X path.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).  Every draw is keyed, so a sweep is reproducible
// for any launch shape:
//   key   = (seed lo32, seed hi32)
//   block j of stream (tag, c0, c1, c2, chain, mu) uses the counter
//         { c0, c1, c2, tag<<28 | chain<<18 | mu<<16 | j }.
// Streams used by the library:
//   tag 1  Metropolis hit      c0 = lexicographic site, c1 = sweep, c2 = hit
//   tag 2  HMC momenta         c0 = site, c1 = trajectory
//   tag 3  random-algebra start c0 = site
//   tag 4  Haar start          c0 = site
//   tag 5  randSUN(delta) start c0 = site
//   tag 6  random-complex start c0 = site
//   tag 7  HMC accept          c0 = 0, c1 = trajectory
//   tag 8  replica swap        c0 = replica pair, c1 = step
// The reference (Metropolis.hpp:34, :76) draws from Kokkos'
// Random_XorShift64_Pool, whose stream depends on the backend and thread
// count; this generator replaces it.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define KLFT_HD __host__ __device__ __forceinline__
#else
#define KLFT_HD inline
#endif

namespace klft_rng {

enum : uint32_t {
  TAG_METROPOLIS = 1,
  TAG_MOMENTA = 2,
  TAG_ALGEBRA = 3,
  TAG_HAAR = 4,
  TAG_RANDSUN = 5,
  TAG_COMPLEX = 6,
  TAG_ACCEPT = 7,
  TAG_SWAP = 8,
  TAG_METROPOLIS_SYM = 9
};

// a ^ b ^ c; on gfx950 one v_bitop3_b32 (truth table 0x96) instead of two XORs
KLFT_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

KLFT_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                           uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
    const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// A stream of 32-bit words.  The four buffered words live in named registers
// and are selected by comparisons, so no private (scratch) array is needed.
struct Stream {
  uint32_t k0, k1, c0, c1, c2, hi, block;
  uint32_t b0, b1, b2, b3;
  int pos;

  KLFT_HD Stream(uint64_t seed, uint32_t tag, uint32_t s0, uint32_t s1, uint32_t s2,
                 uint32_t chain, uint32_t mu)
      : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), c0(s0), c1(s1), c2(s2),
        hi((tag << 28) | (chain << 18) | (mu << 16)), block(0), b0(0), b1(0), b2(0), b3(0),
        pos(4) {}

  // the four words of block j of this stream (independent of the buffer)
  KLFT_HD void block_words(uint32_t j, uint32_t o[4]) const {
    philox4x32_10(c0, c1, c2, hi | (j & 0xffffu), k0, k1, o);
  }

  KLFT_HD void refill() {
    uint32_t o[4];
    philox4x32_10(c0, c1, c2, hi | (block & 0xffffu), k0, k1, o);
    b0 = o[0];
    b1 = o[1];
    b2 = o[2];
    b3 = o[3];
    ++block;
    pos = 0;
  }

  KLFT_HD uint32_t word() {
    if (pos == 4) refill();
    const uint32_t w = pos == 0 ? b0 : (pos == 1 ? b1 : (pos == 2 ? b2 : b3));
    ++pos;
    return w;
  }

  // uniform double in [0, 1) with 53 random bits from two words
  KLFT_HD double u01() {
    const uint32_t a = word() >> 5;
    const uint32_t b = word() >> 6;
    return ((double)a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
  }

  // uniform float in [0, 1) with 24 random bits from one word
  KLFT_HD float u01f() { return (float)(word() >> 8) * (1.0f / 16777216.0f); }
};

}  // namespace klft_rng
