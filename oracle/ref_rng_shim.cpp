// TEST INFRASTRUCTURE ONLY.  Exposes the reference's own RNG (compiled from
// src/rng.h, which includes src/rng.c and src/isaac64.{h,c}) through a C
// interface, so that the oracle's restatement can be checked against it.
// Built by `make -C oracle ref` into oracle/_ref/ (never shipped, never
// part of the product path).
#include <cstdint>

#include "rng.h"

extern "C" {

// first n uint32 draws of isaac64 seeded with `seed` (rng_init)
void ref_isaac64_u32(uint32_t seed, uint64_t n, uint32_t* out) {
  struct rng_state s;
  rng_init(&s, seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = rng_uint(&s);
}

// first n rng_dbl draws
void ref_isaac64_dbl(uint32_t seed, uint64_t n, double* out) {
  struct rng_state s;
  rng_init(&s, seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = rng_dbl(&s);
}

// first n Gaussians (rng_gauss, src/rng.c:173-218)
void ref_rng_gauss(uint32_t seed, uint64_t n, double* out) {
  struct rng_state s;
  rng_init(&s, seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = rng_gauss(&s);
}

}  // extern "C"
