// Prime-field arithmetic over 32-bit limbs in Montgomery form, for device and host.
//
// The four fields of src/zkduel/fields.py:258-294 (2^32-5, Goldilocks, 2^127-1, the BLS12-381
// scalar field) are compile-time policies, so moduli and Montgomery constants become
// instruction immediates. Values are little-endian limb arrays in [0, p).
//
// mont_mul is CIOS Montgomery multiplication (a*b*R^-1 mod p, R = 2^(32L)) that keeps the
// full carry word, so it is exact for every odd p < R, not only below the bound that
// src/zkduel/fields.py:118-121 proves for the reference's truncated routine. Because the
// Montgomery image of a value is unique in [0, p), results are bit-identical to the
// reference kernels' (cuda/src/field_sumcheck.cu:277-319) for these four moduli.
//
// 32x32+64 -> 64 products are written as (uint64_t)a*b + c, which the AMDGPU backend selects
// as v_mad_u64_u32; carries go through 64-bit adds that lower to v_add_co/v_addc_co chains.
// Final corrections are computed unconditionally and chosen with selects, so no lane of a
// wave diverges inside the arithmetic.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define SC_HD __host__ __device__ __forceinline__
#else
#define SC_HD inline
#endif

namespace sc {

template <int L>
struct Elem {
  uint32_t v[L];
};

// ---------------------------------------------------------------------------------------------
// Field policies. P = modulus limbs, NPRIME = -p^-1 mod 2^32, BITS = bit length of p.
// ---------------------------------------------------------------------------------------------
struct FieldP32 {
  static constexpr int L = 1;
  static constexpr int WIDTH = 32;
  static constexpr int BITS = 32;
  static constexpr uint32_t NPRIME = 0xCCCCCCCDu;
  SC_HD static constexpr uint32_t p(int i) { return i == 0 ? 0xFFFFFFFBu : 0u; }
};

struct FieldP64 {  // Goldilocks 2^64 - 2^32 + 1
  static constexpr int L = 2;
  static constexpr int WIDTH = 64;
  static constexpr int BITS = 64;
  static constexpr uint32_t NPRIME = 0xFFFFFFFFu;
  SC_HD static constexpr uint32_t p(int i) { return i == 0 ? 0x00000001u : (i == 1 ? 0xFFFFFFFFu : 0u); }
};

struct FieldP128 {  // Mersenne 2^127 - 1
  static constexpr int L = 4;
  static constexpr int WIDTH = 128;
  static constexpr int BITS = 127;
  static constexpr uint32_t NPRIME = 0x00000001u;
  SC_HD static constexpr uint32_t p(int i) { return i < 3 ? 0xFFFFFFFFu : (i == 3 ? 0x7FFFFFFFu : 0u); }
};

struct FieldP256 {  // BLS12-381 scalar field
  static constexpr int L = 8;
  static constexpr int WIDTH = 256;
  static constexpr int BITS = 255;
  static constexpr uint32_t NPRIME = 0xFFFFFFFFu;
  SC_HD static constexpr uint32_t p(int i) {
    return i == 0   ? 0x00000001u
           : i == 1 ? 0xFFFFFFFFu
           : i == 2 ? 0xFFFE5BFEu
           : i == 3 ? 0x53BDA402u
           : i == 4 ? 0x09A1D805u
           : i == 5 ? 0x3339D808u
           : i == 6 ? 0x299D7D48u
           : i == 7 ? 0x73EDA753u
                    : 0u;
  }
};

// ---------------------------------------------------------------------------------------------
// Arithmetic
// ---------------------------------------------------------------------------------------------
template <class F>
SC_HD Elem<F::L> zero() {
  Elem<F::L> z;
#pragma unroll
  for (int i = 0; i < F::L; ++i) z.v[i] = 0u;
  return z;
}

// s = a - p (with borrow out); used by both add and the Montgomery final step.
template <class F>
SC_HD uint32_t sub_p(Elem<F::L>& out, const Elem<F::L>& a) {
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < F::L; ++i) {
    uint64_t d = (uint64_t)a.v[i] - F::p(i) - borrow;
    out.v[i] = (uint32_t)d;
    borrow = (uint32_t)(d >> 63);
  }
  return borrow;
}

template <class F>
SC_HD Elem<F::L> add(const Elem<F::L>& a, const Elem<F::L>& b) {
  Elem<F::L> s, t;
  uint32_t carry = 0;
#pragma unroll
  for (int i = 0; i < F::L; ++i) {
    uint64_t x = (uint64_t)a.v[i] + b.v[i] + carry;
    s.v[i] = (uint32_t)x;
    carry = (uint32_t)(x >> 32);
  }
  uint32_t borrow = sub_p<F>(t, s);
  // a + b >= p  <=>  carry out of the sum, or no borrow out of (sum - p)
  const bool use_t = carry | (borrow ^ 1u);
#pragma unroll
  for (int i = 0; i < F::L; ++i) s.v[i] = use_t ? t.v[i] : s.v[i];
  return s;
}

template <class F>
SC_HD Elem<F::L> sub(const Elem<F::L>& a, const Elem<F::L>& b) {
  Elem<F::L> d, t;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < F::L; ++i) {
    uint64_t x = (uint64_t)a.v[i] - b.v[i] - borrow;
    d.v[i] = (uint32_t)x;
    borrow = (uint32_t)(x >> 63);
  }
  uint32_t carry = 0;
#pragma unroll
  for (int i = 0; i < F::L; ++i) {
    uint64_t x = (uint64_t)d.v[i] + F::p(i) + carry;
    t.v[i] = (uint32_t)x;
    carry = (uint32_t)(x >> 32);
  }
#pragma unroll
  for (int i = 0; i < F::L; ++i) d.v[i] = borrow ? t.v[i] : d.v[i];
  return d;
}

// CIOS Montgomery product a*b*R^-1 mod p; a, b < p (a < R suffices when b < p).
template <class F>
SC_HD Elem<F::L> mont_mul(const Elem<F::L>& a, const Elem<F::L>& b) {
  constexpr int L = F::L;
  uint32_t t[L + 2];
#pragma unroll
  for (int i = 0; i < L + 2; ++i) t[i] = 0u;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    // t += a * b[i]
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < L; ++j) {
      uint64_t uv = (uint64_t)a.v[j] * b.v[i] + (uint64_t)t[j] + c;
      t[j] = (uint32_t)uv;
      c = uv >> 32;
    }
    uint64_t s = (uint64_t)t[L] + c;
    t[L] = (uint32_t)s;
    t[L + 1] = (uint32_t)(s >> 32);
    // t = (t + m*p) / 2^32 with m = t[0] * nprime mod 2^32
    const uint32_t m = t[0] * F::NPRIME;
    uint64_t uv = (uint64_t)m * F::p(0) + (uint64_t)t[0];
    c = uv >> 32;
#pragma unroll
    for (int j = 1; j < L; ++j) {
      uv = (uint64_t)m * F::p(j) + (uint64_t)t[j] + c;
      t[j - 1] = (uint32_t)uv;
      c = uv >> 32;
    }
    s = (uint64_t)t[L] + c;
    t[L - 1] = (uint32_t)s;
    t[L] = t[L + 1] + (uint32_t)(s >> 32);
  }
  Elem<L> r, q;
#pragma unroll
  for (int i = 0; i < L; ++i) r.v[i] = t[i];
  uint32_t borrow = sub_p<F>(q, r);
  const bool use_q = (t[L] != 0u) | (borrow ^ 1u);
#pragma unroll
  for (int i = 0; i < L; ++i) r.v[i] = use_q ? q.v[i] : r.v[i];
  return r;
}

template <class F>
SC_HD bool is_zero(const Elem<F::L>& a) {
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < F::L; ++i) acc |= a.v[i];
  return acc == 0u;
}

}  // namespace sc
