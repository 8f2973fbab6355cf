// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror and Shaw,
// "Parallel Random Numbers: As Easy as 1, 2, 3", SC'11), usable from host
// and device code.  It replaces the reference's Boost mixmax stream
// (src/stan/services/util/create_rng.hpp:8, 28-34), which is absent here.
//
// Stream convention: key = {seed, chain}; variate k (uniform or normal)
// is generated from counter {lo32(k), hi32(k), 0, 0}, so variate k of a
// chain can be produced independently by any thread.
#ifndef STANX_PHILOX_HPP
#define STANX_PHILOX_HPP

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define SX_HD __host__ __device__
#else
#define SX_HD
#endif

namespace stanx {

struct philox_block {
  uint32_t w[4];
};

SX_HD inline philox_block philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                        uint32_t c3, uint32_t k0,
                                        uint32_t k1) {
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
    const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c3 ^ k1;
    c1 = static_cast<uint32_t>(p1);
    c3 = static_cast<uint32_t>(p0);
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  philox_block b;
  b.w[0] = c0;
  b.w[1] = c1;
  b.w[2] = c2;
  b.w[3] = c3;
  return b;
}

SX_HD inline double philox_bits53(uint32_t a, uint32_t b) {
  return static_cast<double>(
      ((static_cast<uint64_t>(a) << 32) | static_cast<uint64_t>(b)) >> 11);
}

// uniform [0,1) from variate k of stream (seed, chain)
SX_HD inline double philox_uniform(uint32_t seed, uint32_t chain, uint64_t k) {
  philox_block b = philox4x32_10(static_cast<uint32_t>(k),
                                 static_cast<uint32_t>(k >> 32), 0u, 0u, seed,
                                 chain);
  return philox_bits53(b.w[0], b.w[1]) * 0x1.0p-53;
}

// standard normal from variate k: Box-Muller (cos branch), u1 in (0,1]
SX_HD inline double philox_normal(uint32_t seed, uint32_t chain, uint64_t k) {
  philox_block b = philox4x32_10(static_cast<uint32_t>(k),
                                 static_cast<uint32_t>(k >> 32), 0u, 0u, seed,
                                 chain);
  const double u1 = (philox_bits53(b.w[0], b.w[1]) + 1.0) * 0x1.0p-53;
  const double u2 = philox_bits53(b.w[2], b.w[3]) * 0x1.0p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
}

// Host-side sequential stream: the rng_t of this library.
class rng_t {
 public:
  rng_t(uint32_t seed = 0, uint32_t chain = 0)
      : seed_(seed), chain_(chain), counter_(0) {}
  double uniform() { return philox_uniform(seed_, chain_, counter_++); }
  double normal() { return philox_normal(seed_, chain_, counter_++); }
  uint64_t counter() const { return counter_; }
  void set_counter(uint64_t c) { counter_ = c; }
  uint32_t seed() const { return seed_; }
  uint32_t chain() const { return chain_; }

 private:
  uint32_t seed_, chain_;
  uint64_t counter_;
};

}  // namespace stanx
#endif
