/*
 * rng_oracle.c -- test infrastructure (checker only).
 *
 * Philox4x32-10 counter-based generator, restated from Salmon, Moraes, Dror
 * and Shaw, "Parallel Random Numbers: As Easy as 1, 2, 3" (SC'11): ten
 * rounds of two 32x32->64 multiplies with the published multipliers and
 * Weyl key increments.  It stands in for the reference's Boost mixmax
 * stream (src/stan/services/util/create_rng.hpp:28-34), which is absent.
 *
 * Stream convention (restated identically in the product):
 *   key   = {seed, chain}
 *   draw k uses counter {lo32(k), hi32(k), 0, 0}, k = 0, 1, 2, ...
 *   uniform: 53 high bits of (w0:w1) * 2^-53                 in [0,1)
 *   normal : u1 = (53 bits of w0:w1 + 1) * 2^-53  in (0,1],
 *            u2 = 53 bits of (w2:w3) * 2^-53,
 *            z  = sqrt(-2 log u1) cos(2 pi u2)        (Box and Muller 1958)
 */
#include <math.h>
#include "oracle.h"

static void mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                       uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int round = 0; round < 10; ++round) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
    mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    if (round < 9) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_rng_init(orc_rng* r, uint32_t seed, uint32_t chain) {
  r->key[0] = seed;
  r->key[1] = chain;
  r->counter = 0;
}

static void next_block(orc_rng* r, uint32_t w[4]) {
  uint32_t ctr[4] = {(uint32_t)r->counter, (uint32_t)(r->counter >> 32), 0, 0};
  r->counter++;
  orc_philox4x32_10(ctr, r->key, w);
}

static uint64_t bits53(uint32_t a, uint32_t b) {
  return (((uint64_t)a << 32) | (uint64_t)b) >> 11;
}

double orc_rng_uniform(orc_rng* r) {
  uint32_t w[4];
  next_block(r, w);
  return (double)bits53(w[0], w[1]) * 0x1.0p-53;
}

double orc_rng_normal(orc_rng* r) {
  uint32_t w[4];
  next_block(r, w);
  double u1 = ((double)bits53(w[0], w[1]) + 1.0) * 0x1.0p-53;
  double u2 = (double)bits53(w[2], w[3]) * 0x1.0p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
}
