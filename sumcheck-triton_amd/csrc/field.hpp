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

#include <math.h>
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
SC_HD Elem<F::L> add_limbs(const Elem<F::L>& a, const Elem<F::L>& b) {
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

// V mod p for V = sum_k col_k 2^(32k) given as L column sums of 32-bit words (each column
// below 2^63, V below 2^20 p): a sum of canonical values added word by word with no carries.
// The carries are resolved once, the quotient q = floor(V / p - 2^-20) comes from the top
// words in double precision (its error is far below 2^-20, so q is floor(V / p) or one less),
// and V - q p in [0, 2p) (L words and a top bit: 2p can exceed 2^(32L)) takes one conditional
// subtraction.
template <class F>
SC_HD Elem<F::L> reduce_word_columns(const uint64_t (&col)[F::L]) {
  constexpr int L = F::L;
  uint32_t w[L];
  uint64_t c = 0;
#pragma unroll
  for (int k = 0; k < L; ++k) {
    c += col[k];
    w[k] = (uint32_t)c;
    c >>= 32;
  }
  const uint32_t top = (uint32_t)c;  // V = w + top 2^(32L), top < 2^21
  double pd = 0.0;                   // p / 2^(32(L-1)) as a double
#pragma unroll
  for (int k = 0; k < L; ++k) pd = pd * 0x1p-32 + (double)F::p(k);
  double v = (double)top * 0x1p32 + (double)w[L - 1];
  if (L > 1) v += (double)w[L > 1 ? L - 2 : 0] * 0x1p-32;
  double q = floor(fmax(v / pd - 0x1p-20, 0.0));
  const uint32_t qi = (uint32_t)q;
  // x = V - q p: q p as L + 1 words, subtracted with borrow
  Elem<L> x;
  uint64_t m = 0;
  uint32_t borrow = 0;
#pragma unroll
  for (int k = 0; k < L; ++k) {
    m += (uint64_t)qi * F::p(k);
    const uint64_t d = (uint64_t)w[k] - (uint32_t)m - borrow;
    x.v[k] = (uint32_t)d;
    borrow = (uint32_t)(d >> 63);
    m >>= 32;
  }
  const uint32_t xtop = top - (uint32_t)m - borrow;  // 0 or 1: x + xtop 2^(32L) < 2p
  Elem<L> t;
  const uint32_t b2 = sub_p<F>(t, x);  // t = x - p (mod 2^(32L)), the value when x >= p
  const bool ge = xtop != 0u || b2 == 0u;
#pragma unroll
  for (int k = 0; k < L; ++k) x.v[k] = ge ? t.v[k] : x.v[k];
  return x;
}

template <class F>
SC_HD Elem<F::L> sub_limbs(const Elem<F::L>& a, const Elem<F::L>& b) {
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
SC_HD Elem<F::L> mont_mul_cios(const Elem<F::L>& a, const Elem<F::L>& b) {
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

// ---------------------------------------------------------------------------------------------
// Goldilocks on 64-bit words. p = 2^64 - 2^32 + 1, so 2^64 = 2^32 - 1 (mod p): a carry out of
// bit 63 is folded back by adding 2^32 - 1, a borrow by subtracting it.
// ---------------------------------------------------------------------------------------------
constexpr uint64_t GL_P = 0xFFFFFFFF00000001ull;
constexpr uint64_t GL_EPS = 0xFFFFFFFFull;  // 2^64 mod p

SC_HD uint64_t gl_u64(const Elem<2>& a) { return (uint64_t)a.v[0] | ((uint64_t)a.v[1] << 32); }
SC_HD Elem<2> gl_elem(uint64_t x) {
  Elem<2> e;
  e.v[0] = (uint32_t)x;
  e.v[1] = (uint32_t)(x >> 32);
  return e;
}
// x in [0, 2^64) -> [0, p) (2^64 < 2p)
SC_HD uint64_t gl_canon(uint64_t x) { return x >= GL_P ? x - GL_P : x; }
// a, b < p -> (a + b) mod p
SC_HD uint64_t gl_add(uint64_t a, uint64_t b) {
  const uint64_t s = a + b;
  const uint64_t t = s + GL_EPS;  // s - p mod 2^64
  return (s < a || s >= GL_P) ? t : s;
}
// a, b < p -> (a - b) mod p
SC_HD uint64_t gl_sub(uint64_t a, uint64_t b) {
  const uint64_t d = a - b;
  return a < b ? d - GL_EPS : d;  // d + p mod 2^64
}
// One operand < p, the other < 2^64: a value < 2^64 congruent to a + b (not reduced).
SC_HD uint64_t gl_add_lazy(uint64_t a, uint64_t b) {
  const uint64_t s = a + b;
  return s < a ? s + GL_EPS : s;
}

// ---------------------------------------------------------------------------------------------
// 2^32 - 5 on one word: 2^32 = 5 (mod p), so subtracting p is adding 5 modulo 2^32.
// ---------------------------------------------------------------------------------------------
constexpr uint32_t P32 = 0xFFFFFFFBu;

// a, b < p -> (a + b) mod p: a + b >= p exactly when a + b or a + b + 5 carries out of 32
// bits (a + b < 2p), and then a + b - p = (a + b + 5) mod 2^32.
SC_HD uint32_t p32_add(uint32_t a, uint32_t b) {
  const uint32_t s = a + b;
  const uint32_t t = s + 5u;
  return (s < a || t < s) ? t : s;
}
// a, b < p -> (a - b) mod p: on a borrow, (a - b) mod 2^32 - 5 = a - b + p (no wrap: >= 1).
SC_HD uint32_t p32_sub(uint32_t a, uint32_t b) {
  const uint32_t d = a - b;
  return a < b ? d - 5u : d;
}
// Montgomery product a b 2^-32 mod p in [0, p) for a, b < p: with m = x NPRIME mod 2^32,
// r = (x + m p) / 2^32 < 2p is the high word h of the 64-bit sum plus its carry c times
// 2^32; r >= p exactly when c is set or h + 5 carries, and then r - p = (h + 5) mod 2^32.
SC_HD uint32_t p32_mont_mul(uint32_t a, uint32_t b) {
  const uint64_t x = (uint64_t)a * b;
  const uint32_t m = (uint32_t)x * 0xCCCCCCCDu;
  const uint64_t s = (uint64_t)m * P32 + x;
  const uint32_t h = (uint32_t)(s >> 32);
  const uint32_t t = h + 5u;
  return (s < x || t < h) ? t : h;
}

// Canonical add/sub for every field (limb chains; 64-bit words for Goldilocks; one word with
// the 2^32 = 5 identity for 2^32 - 5).
template <class F>
SC_HD Elem<F::L> add(const Elem<F::L>& a, const Elem<F::L>& b) {
  if constexpr (F::WIDTH == 64) {
    return gl_elem(gl_add(gl_u64(a), gl_u64(b)));
  } else if constexpr (F::WIDTH == 32) {
    Elem<1> s;
    s.v[0] = p32_add(a.v[0], b.v[0]);
    return s;
  } else {
    return add_limbs<F>(a, b);
  }
}

template <class F>
SC_HD Elem<F::L> sub(const Elem<F::L>& a, const Elem<F::L>& b) {
  if constexpr (F::WIDTH == 64) {
    return gl_elem(gl_sub(gl_u64(a), gl_u64(b)));
  } else if constexpr (F::WIDTH == 32) {
    Elem<1> d;
    d.v[0] = p32_sub(a.v[0], b.v[0]);
    return d;
  } else {
    return sub_limbs<F>(a, b);
  }
}

// The interpolation step x(t+1) = x(t) + d with d < p. For Goldilocks x(t) may be left
// unreduced in [0, 2^64) (the Montgomery product accepts any 64-bit operands); other fields
// reduce.
template <class F>
SC_HD Elem<F::L> step_add(const Elem<F::L>& x, const Elem<F::L>& d) {
  if constexpr (F::WIDTH == 64) {
    return gl_elem(gl_add_lazy(gl_u64(x), gl_u64(d)));
  } else {
    return add<F>(x, d);
  }
}

// x -> canonical form (identity except for unreduced Goldilocks values).
template <class F>
SC_HD Elem<F::L> canon(const Elem<F::L>& x) {
  if constexpr (F::WIDTH == 64) {
    return gl_elem(gl_canon(gl_u64(x)));
  } else {
    return x;
  }
}

// Running sum of field elements. Generic fields add modulo p each time; Goldilocks keeps a
// 64-bit sum plus a count of carries out of bit 63 (each worth 2^32 - 1) and reduces once.
template <class F>
struct Accum {
  Elem<F::L> v;
  SC_HD void clear() { v = zero<F>(); }
  SC_HD void add(const Elem<F::L>& x) { v = sc::add<F>(v, canon<F>(x)); }
  SC_HD void sub(const Elem<F::L>& x) { v = sc::sub<F>(v, canon<F>(x)); }
  SC_HD Elem<F::L> value() const { return v; }
};

// 2^32 - 5: a 64-bit sum of canonical words (at most 2^32 of them), reduced once with
// 2^32 = 5 (mod p): 5 hi + lo < 6 2^32, once more below 2^32 + 30 < 2p, then one subtraction.
template <>
struct Accum<FieldP32> {
  uint64_t s;
  SC_HD void clear() { s = 0; }
  SC_HD void add(const Elem<1>& x) { s += x.v[0]; }
  SC_HD void sub(const Elem<1>& x) { s += P32 - x.v[0]; }  // x < p, so p - x in [1, p]
  SC_HD Elem<1> value() const {
    const uint64_t y = (s >> 32) * 5u + (uint32_t)s;
    const uint64_t z = (y >> 32) * 5u + (uint32_t)y;
    Elem<1> r;
    r.v[0] = z >= P32 ? (uint32_t)(z - P32) : (uint32_t)z;
    return r;
  }
};

SC_HD uint64_t gl_product(const Elem<2>& a, const Elem<2>& b, uint64_t& hi);
SC_HD uint64_t gl_reduce(uint64_t lo, uint64_t hi);

// Goldilocks: the sum is held as an integer top 2^128 + W (W = w_hi 2^64 + w_lo) whose
// value is (top 2^128 + W) R^-1 mod p, R = 2^64. A Montgomery value X enters as X 2^64 (an
// add to w_hi); a product A B of two Montgomery values enters as the raw 128-bit product, so
// its reduction by R^-1 happens once for the whole sum, in value(), instead of once per
// product. 2^128 R^-1 = 2^64 = 2^32 - 1 (mod p) prices each carry out of bit 127.
template <>
struct Accum<FieldP64> {
  uint64_t w_lo, w_hi;
  uint32_t top;
  SC_HD void clear() {
    w_lo = 0;
    w_hi = 0;
    top = 0;
  }
  SC_HD void add_hi(uint64_t v) {
    const uint64_t s = w_hi + v;
    top += s < v ? 1u : 0u;
    w_hi = s;
  }
  SC_HD void add(const Elem<2>& x) { add_hi(gl_u64(x)); }  // any x < 2^64
  SC_HD void sub(const Elem<2>& x) {  // adds p - canon(x) (0 stays 0)
    const uint64_t c = gl_canon(gl_u64(x));
    add_hi(c == 0 ? 0 : GL_P - c);
  }
  // W += a b for any a, b < 2^64 (the product's high word is at most 2^64 - 2, so it absorbs
  // the carry of the low-word add)
  SC_HD void add_product(const Elem<2>& a, const Elem<2>& b) {
    uint64_t hi;
    const uint64_t lo = gl_product(a, b, hi);
    const uint64_t s = w_lo + lo;
    w_lo = s;
    add_hi(hi + (s < lo ? 1u : 0u));
  }
  SC_HD Elem<2> value() const {
    const uint64_t r = gl_reduce(w_lo, w_hi);  // W R^-1 mod p
    return gl_elem(gl_add(r, gl_canon((uint64_t)top * GL_EPS)));
  }
};

// acc (64-bit) + top (32-bit) += a * b. On the device this is v_mad_u64_u32 with its carry
// out, then v_addc_co_u32 into top: two instructions per multiply-accumulate, with the
// accumulator pair used in place as the 64-bit addend (a C++ formulation makes the compiler
// re-pair zero-extended halves with extra moves). Arithmetic only, no memory operands. The
// s_nop gives the two wait states gfx950 needs between a VALU write of an SGPR (the carry)
// and a VALU read of it, which the compiler does not insert inside an asm statement.
SC_HD void mac(uint64_t& acc, uint32_t& top, uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t carry;
  asm("v_mad_u64_u32 %0, %1, %3, %4, %0\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32_e64 %2, %1, 0, %2, %1"
      : "+v"(acc), "=&s"(carry), "+v"(top)
      : "v"(a), "v"(b));
#else
  const uint64_t s = (uint64_t)a * b + acc;
  top += s < acc ? 1u : 0u;
  acc = s;
#endif
}

// Montgomery product by finely integrated product scanning (Koc, Acar, Kaliski, "Analyzing
// and comparing Montgomery multiplication algorithms", 1996): column k of a*b and of m*p is
// summed into a 96-bit accumulator, m_k chosen so the column's low word vanishes. 2L^2 + L
// multiply-accumulates; the result is < 2p and is brought into [0, p) by one subtraction.
template <class F>
SC_HD Elem<F::L> mont_mul_fips(const Elem<F::L>& a, const Elem<F::L>& b) {
  constexpr int L = F::L;
  uint32_t m[L];
  uint64_t acc = 0;
  uint32_t top = 0;
#pragma unroll
  for (int i = 0; i < L; ++i) {
#pragma unroll
    for (int j = 0; j < i; ++j) {
      mac(acc, top, a.v[j], b.v[i - j]);
      mac(acc, top, m[j], F::p(i - j));
    }
    mac(acc, top, a.v[i], b.v[0]);
    m[i] = (uint32_t)acc * F::NPRIME;
    mac(acc, top, m[i], F::p(0));
    acc = (acc >> 32) | ((uint64_t)top << 32);
    top = 0;
  }
  Elem<L> r;
#pragma unroll
  for (int i = L; i < 2 * L; ++i) {
#pragma unroll
    for (int j = i - L + 1; j < L; ++j) {
      mac(acc, top, a.v[j], b.v[i - j]);
      mac(acc, top, m[j], F::p(i - j));
    }
    r.v[i - L] = (uint32_t)acc;
    acc = (acc >> 32) | ((uint64_t)top << 32);
    top = 0;
  }
  Elem<L> q;
  const uint32_t borrow = sub_p<F>(q, r);
  const bool use_q = (acc != 0u) | (borrow ^ 1u);
#pragma unroll
  for (int i = 0; i < L; ++i) r.v[i] = use_q ? q.v[i] : r.v[i];
  return r;
}

// Goldilocks (p = 2^64 - 2^32 + 1, R = 2^64). 2^96 = -1 and 2^192 = 1 mod p, so
// R^-1 = 2^128 = -2^32 and, for x = a*b = x3:x2:x1:x0 (32-bit words),
//   x * R^-1 = (x3 - x1 - x0) * 2^32 + x1 + x2   (mod p).
// The 64-bit sums below fold each wrap-around back in with 2^64 = 2^32 - 1 (mod p).
// Any 64-bit operands (x < 2^128 always).

// The 128-bit product a*b as lo (x1:x0) and hi (x3:x2).
SC_HD uint64_t gl_product(const Elem<2>& a, const Elem<2>& b, uint64_t& hi) {
  const uint64_t p00 = (uint64_t)a.v[0] * b.v[0];
  const uint64_t t0 = (uint64_t)a.v[0] * b.v[1] + (p00 >> 32);
  const uint64_t t1 = (uint64_t)a.v[1] * b.v[0] + (uint32_t)t0;
  hi = (uint64_t)a.v[1] * b.v[1] + ((t0 >> 32) + (t1 >> 32));
  return (uint32_t)p00 | ((uint64_t)(uint32_t)t1 << 32);
}

// x * R^-1 mod p in [0, p) for x = hi:lo < 2^128.
SC_HD uint64_t gl_reduce(uint64_t lo, uint64_t hi) {
  const uint64_t x0 = (uint32_t)lo, x1 = lo >> 32;
  // s = x3 * 2^32 + x2 + x1 (hi is x3 * 2^32 + x2)
  const uint64_t add = hi + x1;
  const uint64_t s = add + ((add < hi) ? 0xFFFFFFFFull : 0ull);
  // s -= x0 * 2^32, s -= x1 * 2^32; a borrow of 2^64 is worth -(2^32 - 1)
  uint64_t d = s - (x0 << 32);
  d -= (d > s) ? 0xFFFFFFFFull : 0ull;
  const uint64_t e = d - (x1 << 32);
  const uint64_t f = e - ((e > d) ? 0xFFFFFFFFull : 0ull);
  return f >= 0xFFFFFFFF00000001ull ? f - 0xFFFFFFFF00000001ull : f;
}

SC_HD Elem<2> goldilocks_mont_mul(const Elem<2>& a, const Elem<2>& b) {
  uint64_t hi;
  const uint64_t lo = gl_product(a, b, hi);
  return gl_elem(gl_reduce(lo, hi));
}

// (a1 b1 + a2 b2) R^-1 with one reduction: the two 128-bit products are summed first; a
// carry out of bit 128 is worth 2^128 R^-1 = 2^64 = 2^32 - 1 (mod p). Any 64-bit operands;
// returns a value < 2^64 congruent to the result (not reduced below p).
SC_HD Elem<2> goldilocks_mont_mul_sum(const Elem<2>& a1, const Elem<2>& b1, const Elem<2>& a2, const Elem<2>& b2) {
  uint64_t h1, h2;
  const uint64_t l1 = gl_product(a1, b1, h1);
  const uint64_t l2 = gl_product(a2, b2, h2);
  const uint64_t lo = l1 + l2;
  const uint64_t h = h1 + h2;
  const uint64_t hi = h + (lo < l1 ? 1ull : 0ull);
  const bool carry = h < h1 || hi < h;
  return gl_elem(gl_reduce(lo, hi) + (carry ? 0xFFFFFFFFull : 0ull));
}

// Mersenne p = 2^127 - 1, R = 2^128: 2^127 = 1, so R = 2 and R^-1 = 2^126 (mod p), and
// x * R^-1 is x mod p rotated right by one bit within 127 bits. x mod p folds the product
// at bit 127 twice (x < 2^254).
SC_HD Elem<4> mersenne127_mont_mul(const Elem<4>& a, const Elem<4>& b) {
  uint32_t t[8];
  uint64_t acc = 0;
  uint32_t top = 0;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
#pragma unroll
    for (int i = (k < 4 ? 0 : k - 3); i <= (k < 4 ? k : 3); ++i) mac(acc, top, a.v[i], b.v[k - i]);
    t[k] = (uint32_t)acc;
    acc = (acc >> 32) | ((uint64_t)top << 32);
    top = 0;
  }
  t[7] = (uint32_t)acc;
  // lo = x mod 2^127, hi = x >> 127 (both < 2^127)
  uint32_t hi[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) hi[i] = (t[i + 3] >> 31) | (t[i + 4] << 1);
  uint32_t lo[4] = {t[0], t[1], t[2], t[3] & 0x7FFFFFFFu};
  // s = lo + hi < 2^128; fold bit 127 once more (s < 2^127 + 1 afterwards: s <= p)
  uint32_t sum[4];
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    c += (uint64_t)lo[i] + hi[i];
    sum[i] = (uint32_t)c;
    c >>= 32;
  }
  const uint32_t bit127 = sum[3] >> 31;
  sum[3] &= 0x7FFFFFFFu;
  c = bit127;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    c += sum[i];
    sum[i] = (uint32_t)c;
    c >>= 32;
  }
  // s == p (all 127 bits set) is 0 mod p
  const bool is_p = (sum[0] & sum[1] & sum[2]) == 0xFFFFFFFFu && sum[3] == 0x7FFFFFFFu;
  // rotate right by one within 127 bits: (s >> 1) | ((s & 1) << 126)
  Elem<4> out;
  out.v[0] = (sum[0] >> 1) | (sum[1] << 31);
  out.v[1] = (sum[1] >> 1) | (sum[2] << 31);
  out.v[2] = (sum[2] >> 1) | (sum[3] << 31);
  out.v[3] = (sum[3] >> 1) | ((sum[0] & 1u) << 30);
  if (is_p) {
    out.v[0] = out.v[1] = out.v[2] = out.v[3] = 0u;
  }
  return out;
}

// Montgomery product a*b*R^-1 mod p in [0, p) for a, b in [0, p). Field-specific
// reductions where the modulus has structure; the result is the same unique value.
template <class F>
SC_HD Elem<F::L> mont_mul(const Elem<F::L>& a, const Elem<F::L>& b) {
  if constexpr (F::WIDTH == 64) {
    return goldilocks_mont_mul(a, b);
  } else if constexpr (F::WIDTH == 32) {
    Elem<1> r;
    r.v[0] = p32_mont_mul(a.v[0], b.v[0]);
    return r;
  } else if constexpr (F::WIDTH == 128) {
    return mersenne127_mont_mul(a, b);
  } else {
#if defined(SC_MONT_CIOS)
    return mont_mul_cios<F>(a, b);
#else
    return mont_mul_fips<F>(a, b);
#endif
  }
}

template <class F>
SC_HD bool is_zero(const Elem<F::L>& a) {
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < F::L; ++i) acc |= a.v[i];
  return acc == 0u;
}

}  // namespace sc
