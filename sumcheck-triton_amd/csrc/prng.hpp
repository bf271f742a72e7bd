// Counter-based input generator for synthetic tables (host and device).
//
// Every element is a pure function of (seed, instance, table, index), so the device fills
// tables in parallel and any CPU checker regenerates exactly the same values. The mixing
// function is the SplitMix64 finalizer (Steele, Lea, Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014).
//
// uniform: L 32-bit words from ceil(L/2) draws, masked to bitlen(p), rejection-sampled into
//          [0, p) with at most 32 attempts (then reduced mod p; never reached in practice:
//          the worst rejection rate, BLS12-381, is 0.094 per attempt).
// selector: one draw, low bit, so the entry is 0 or 1.
// affine:  the reference's seed + 7919(v+1) + 104729(o+1) (cuda/src/field_sumcheck.cu:123-128),
//          reduced mod p.
#pragma once

#include <stdint.h>

#include "field.hpp"

namespace sc {

constexpr uint64_t PRNG_GAMMA = 0x9E3779B97F4A7C15ull;
constexpr int PRNG_MAX_ATTEMPTS = 32;

SC_HD uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

SC_HD uint64_t stream_key(uint64_t seed, uint32_t instance, uint32_t table) {
  const uint64_t k = mix64(seed ^ 0x5851F42D4C957F2Dull);
  return mix64(k ^ (((uint64_t)instance << 32) | table));
}

SC_HD uint64_t prng_draw(uint64_t key, uint64_t counter) { return mix64(key + counter * PRNG_GAMMA); }

template <class F>
SC_HD Elem<F::L> uniform_draw(uint64_t key, uint64_t index, int attempt) {
  constexpr int L = F::L;
  Elem<L> x;
#pragma unroll
  for (int w = 0; w < (L + 1) / 2; ++w) {
    const uint64_t d = prng_draw(key, index * 128u + (uint64_t)attempt * 4u + (uint64_t)w);
    x.v[2 * w] = (uint32_t)d;
    if (2 * w + 1 < L) x.v[2 * w + 1] = (uint32_t)(d >> 32);
  }
  constexpr int top_bits = F::BITS - 32 * (L - 1);
  if (top_bits < 32) x.v[L - 1] &= (1u << top_bits) - 1u;
  return x;
}

// The first attempt whose masked draw is below p. Each attempt draws the word holding the top
// limb first: a top limb above p's rejects the attempt without the other draws, one below
// accepts it, and only a tie needs the whole comparison (the same values as drawing every word
// of every attempt, with a third of the draws at 255 bits, where 9.4% of attempts fail).
template <class F>
SC_HD Elem<F::L> uniform_element(uint64_t key, uint64_t index) {
  constexpr int L = F::L;
  constexpr int TW = (L + 1) / 2 - 1;  // the draw holding limb L - 1
  constexpr int top_bits = F::BITS - 32 * (L - 1);
  constexpr uint32_t top_mask = top_bits < 32 ? (1u << top_bits) - 1u : 0xFFFFFFFFu;
  for (int attempt = 0; attempt < PRNG_MAX_ATTEMPTS; ++attempt) {
    const uint64_t d = prng_draw(key, index * 128u + (uint64_t)attempt * 4u + (uint64_t)TW);
    const uint32_t top = ((L & 1) ? (uint32_t)d : (uint32_t)(d >> 32)) & top_mask;
    if (top > F::p(L - 1)) continue;  // x > p
    const Elem<L> x = uniform_draw<F>(key, index, attempt);
    if (top < F::p(L - 1)) return x;  // x < p
    Elem<L> t;
    if (sub_p<F>(t, x)) return x;  // x < p
  }
  // Unreachable for the four fields in practice; keep the value in range deterministically.
  const Elem<L> x = uniform_draw<F>(key, index, PRNG_MAX_ATTEMPTS - 1);
  Elem<L> t;
  sub_p<F>(t, x);
  return t;
}

template <class F>
SC_HD Elem<F::L> selector_element(uint64_t key, uint64_t index) {
  Elem<F::L> x = zero<F>();
  x.v[0] = (uint32_t)(prng_draw(key, index * 128u) & 1u);
  return x;
}

template <class F>
SC_HD Elem<F::L> affine_element(uint64_t seed, uint32_t table, uint64_t offset) {
  // Below p for every table and size that fits in memory, except at 32 bits.
  const uint64_t v = seed + 7919ull * (table + 1ull) + 104729ull * (offset + 1ull);
  Elem<F::L> x = zero<F>();
  if (F::L == 1) {
    x.v[0] = (uint32_t)(v % (uint64_t)F::p(0));
  } else {
    x.v[0] = (uint32_t)v;
    x.v[1] = (uint32_t)(v >> 32);
  }
  return x;
}

}  // namespace sc
