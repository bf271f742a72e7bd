// Working form of field elements inside the prover kernels.
//
// Tables arrive as Montgomery limbs in [0, p) (R = 2^(32L)), and round polynomials and
// challenges leave in that form. In between, a kernel may hold values in any representation
// that maps back to the same residues; `Work<F>` names the one each field uses:
//
//   from_mont(w)      caller's Montgomery limbs -> working value
//   from_stored(w)    fold-buffer words (written by store) -> working value
//   store(w, x)       working value -> fold-buffer words (same size as a Montgomery element)
//   coeff(w)          composition coefficient, already in working form on the host (work_coeff)
//   mul(a, b)         product
//   mul_sum(a1, b1, a2, b2)  a1 b1 + a2 b2 (one reduction where the form allows it)
//   diff(hi, lo)      hi - lo, the interpolation step of a pair
//   step(x, d)        x + d
//   fold(lo, hi, r)   lo + r (hi - lo), the folded entry, ready for store
//   at(lo, d, t)      lo + t d for a small per-lane t (point-parallel evaluation)
//   Acc               running sum: add, sub, value() -> canonical Montgomery limbs
//
// The 2^32-5 and Goldilocks fields work on their Montgomery limbs directly. 2^127-1 works on
// plain residues in radix 2^26 (R26) and the BLS12-381 scalar field on Montgomery images with
// R = 2^261 in radix 2^29 (R29): with limbs narrower than 32 bits, 32x32-bit multiply-adds
// never overflow a 64-bit column sum, so products need no carry chains.
#pragma once

#include <math.h>

#include "field.hpp"

namespace sc {

template <class F>
struct Work {
  static constexpr int L = F::L;
  static constexpr bool kMontIsStored = true;  // from_mont and from_stored read the same form
  // The 255-bit form reads Montgomery input more cheaply in two places (Work<FieldP256>):
  // folding straight from it, and reading it scaled in round 0 of compiled compositions.
  static constexpr bool kCheapMontInput = false;
  // The 255-bit form folds by a per-block table of the challenge (fold2_tab, fold1_tab).
  static constexpr bool kFoldTable = false;
  static constexpr bool kLazyStep = false;  // step_lazy: an interpolation step left unnormalized
  using V = Elem<L>;
  SC_HD static V load(const uint32_t* w) {
    V x;
#pragma unroll
    for (int i = 0; i < L; ++i) x.v[i] = w[i];
    return x;
  }
  SC_HD static V from_mont(const uint32_t* w) { return load(w); }
  SC_HD static V from_stored(const uint32_t* w) { return load(w); }
  SC_HD static V coeff(const uint32_t* w) { return load(w); }
  SC_HD static void store(uint32_t* w, const V& x) {
#pragma unroll
    for (int i = 0; i < L; ++i) w[i] = x.v[i];
  }
  SC_HD static V mul(const V& a, const V& b) { return mont_mul<F>(a, b); }
  // a1 b1 + a2 b2 (Goldilocks: one reduction for both products)
  SC_HD static V mul_sum(const V& a1, const V& b1, const V& a2, const V& b2) {
    if constexpr (F::WIDTH == 64) {
      return goldilocks_mont_mul_sum(a1, b1, a2, b2);
    } else {
      return step_add<F>(mont_mul<F>(a1, b1), mont_mul<F>(a2, b2));
    }
  }
  SC_HD static V diff(const V& hi, const V& lo) { return sub<F>(hi, lo); }
  SC_HD static V step(const V& x, const V& d) { return step_add<F>(x, d); }
  SC_HD static V fold(const V& lo, const V& hi, const V& r) { return add<F>(lo, mont_mul<F>(r, sub<F>(hi, lo))); }
  // lo + t d for a per-lane 0 <= t < 8
  SC_HD static V at(const V& lo, const V& d, int t) {
    V x = lo;
    for (int s = 0; s < t; ++s) x = step_add<F>(x, d);
    return x;
  }
  using Acc = Accum<F>;
};

// ---------------------------------------------------------------------------------------------
// Goldilocks working form: the caller's Montgomery words (R = 2^64), with products and running
// sums built so that no 64-bit carry has to be detected.
//
// gl_mad(a, b, c) = a b + c for 32-bit a, b and a 64-bit c is one v_mad_u64_u32 (written as a
// one-instruction asm statement: in C++ the backend rewrites sums of zero-extended words into
// 64-bit adds that need extra register moves). With b = 1 it adds a 32-bit word to a 64-bit
// one. The 128-bit product of two 64-bit words a = a1:a0, b = b1:b0 then takes 8 of them,
// none of which can overflow:
//   L = a0 b0,  T = a1 b0 + L1,  U = a0 b1 + T0,  V = a1 b1 + T1 + U1   (T, U < 2^64, V <= 2^64 - 1)
//   a b = L0 + U0 2^32 + V 2^64.
// Montgomery reduction of x = V 2^64 + x1 2^32 + x0 uses 2^96 = -1, 2^64 = 2^32 - 1 (mod p):
//   x 2^-64 = V + x1 - (x0 + x1) 2^32 = V - D,  D = (s0 + s1) 2^32 - (x1 + s1),
// with s = x0 + x1 = s1 2^32 + s0; 0 <= D <= 2^64 - 2^32 < p, and a borrow out of V - D is
// paid by adding p (subtracting 2^32 - 1 from the wrapped difference). The carries of that
// chain are __builtin_addc / __builtin_subc, so the compiler places their hazard waits.
//
// Value ranges: products and sums may be any 64-bit word congruent to the result (below 2^64,
// not necessarily below p); interpolation differences, folded entries and everything stored
// are canonical (below p), which the lazy additions x + d need of d.
// ---------------------------------------------------------------------------------------------
SC_HD uint32_t gl_adc(uint32_t a, uint32_t b, uint32_t ci, uint32_t& co) {
#if defined(__clang__)
  unsigned c;
  const uint32_t s = __builtin_addc(a, b, ci, &c);
  co = c;
  return s;
#else
  const uint64_t s = (uint64_t)a + b + ci;
  co = (uint32_t)(s >> 32);
  return (uint32_t)s;
#endif
}

SC_HD uint32_t gl_sbb(uint32_t a, uint32_t b, uint32_t bi, uint32_t& bo) {
#if defined(__clang__)
  unsigned c;
  const uint32_t d = __builtin_subc(a, b, bi, &c);
  bo = c;
  return d;
#else
  const uint64_t d = (uint64_t)a - b - bi;
  bo = (uint32_t)(d >> 63);
  return (uint32_t)d;
#endif
}

SC_HD uint64_t gl_mad(uint32_t a, uint32_t b, uint64_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t r, carry;  // the carry out cannot be set here; it goes to an SGPR nobody reads
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(carry) : "v"(a), "v"(b), "v"(c));
  return r;
#else
  return (uint64_t)a * b + c;
#endif
}

// 128-bit product of two 64-bit words: x0, x1 and the upper word V.
struct GlWide {
  uint32_t x0, x1;
  uint64_t hi;
};

SC_HD GlWide gl_wide_mul(uint64_t a, uint64_t b) {
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  const uint64_t l = gl_mad(a0, b0, 0);
  const uint64_t t = gl_mad((uint32_t)(l >> 32), 1u, gl_mad(a1, b0, 0));
  const uint64_t u = gl_mad((uint32_t)t, 1u, gl_mad(a0, b1, 0));
  const uint64_t v = gl_mad((uint32_t)(u >> 32), 1u, gl_mad((uint32_t)(t >> 32), 1u, gl_mad(a1, b1, 0)));
  GlWide x;
  x.x0 = (uint32_t)l;
  x.x1 = (uint32_t)u;
  x.hi = v;
  return x;
}

// x 2^-64 mod p as a word below 2^64.
SC_HD uint64_t gl_redc(const GlWide& x) {
  uint32_t s1, b1, b2, b3, b4, b5, z;
  const uint32_t s0 = gl_adc(x.x0, x.x1, 0u, s1);
  uint32_t dh = gl_adc(s0, 0u, s1, z);
  uint32_t dl = gl_sbb(0u, x.x1, 0u, b1);
  dh = gl_sbb(dh, 0u, b1, z);
  dl = gl_sbb(dl, 0u, s1, b2);
  dh = gl_sbb(dh, 0u, b2, z);
  uint32_t r0 = gl_sbb((uint32_t)x.hi, dl, 0u, b3);
  uint32_t r1 = gl_sbb((uint32_t)(x.hi >> 32), dh, b3, b4);
  r0 = gl_sbb(r0, 0u - b4, 0u, b5);
  r1 = gl_sbb(r1, 0u, b5, z);
  return (uint64_t)r0 | ((uint64_t)r1 << 32);
}

// x + d for any x < 2^64 and d < p: a carry out of bit 64 is worth 2^32 - 1, and adding it
// cannot carry again (x + d < 2^65 - (2^32 - 1)).
SC_HD uint64_t gl_add_fast(uint64_t x, uint64_t d) {
  uint32_t c0, c1, c2, z;
  uint32_t s0 = gl_adc((uint32_t)x, (uint32_t)d, 0u, c0);
  uint32_t s1 = gl_adc((uint32_t)(x >> 32), (uint32_t)(d >> 32), c0, c1);
  s0 = gl_adc(s0, 0u - c1, 0u, c2);
  s1 = gl_adc(s1, 0u, c2, z);
  return (uint64_t)s0 | ((uint64_t)s1 << 32);
}

// x - y for any x < 2^64 and y < p: a borrow is paid by adding p, i.e. subtracting 2^32 - 1
// from the wrapped difference (which is then at least 2^64 - p). Canonical when x < p too.
SC_HD uint64_t gl_sub_fast(uint64_t x, uint64_t y) {
  uint32_t b0, b1, b2, z;
  uint32_t d0 = gl_sbb((uint32_t)x, (uint32_t)y, 0u, b0);
  uint32_t d1 = gl_sbb((uint32_t)(x >> 32), (uint32_t)(y >> 32), b0, b1);
  d0 = gl_sbb(d0, 0u - b1, 0u, b2);
  d1 = gl_sbb(d1, 0u, b2, z);
  return (uint64_t)d0 | ((uint64_t)d1 << 32);
}

// x mod p for any x < 2^64: x + 2^32 - 1 carries out of bit 64 exactly when x >= p.
SC_HD uint64_t gl_canon_fast(uint64_t x) {
  uint32_t c0, c1;
  const uint32_t t0 = gl_adc((uint32_t)x, 0xFFFFFFFFu, 0u, c0);
  const uint32_t t1 = gl_adc((uint32_t)(x >> 32), 0u, c0, c1);
  const uint32_t y0 = c1 ? t0 : (uint32_t)x, y1 = c1 ? t1 : (uint32_t)(x >> 32);
  return (uint64_t)y0 | ((uint64_t)y1 << 32);
}

template <>
struct Work<FieldP64> {
  static constexpr int L = 2;
  static constexpr bool kMontIsStored = true;
  static constexpr bool kCheapMontInput = false;
  static constexpr bool kFoldTable = false;
  static constexpr bool kLazyStep = false;
  using V = Elem<2>;
  SC_HD static V load(const uint32_t* w) {
    V x;
    x.v[0] = w[0];
    x.v[1] = w[1];
    return x;
  }
  SC_HD static V from_mont(const uint32_t* w) { return load(w); }
  SC_HD static V from_stored(const uint32_t* w) { return load(w); }
  SC_HD static V coeff(const uint32_t* w) { return load(w); }
  SC_HD static void store(uint32_t* w, const V& x) {
    w[0] = x.v[0];
    w[1] = x.v[1];
  }
  SC_HD static V mul(const V& a, const V& b) { return gl_elem(gl_redc(gl_wide_mul(gl_u64(a), gl_u64(b)))); }
  // a1 b1 + a2 b2: the reduced products summed lazily (the first made canonical)
  SC_HD static V mul_sum(const V& a1, const V& b1, const V& a2, const V& b2) {
    const uint64_t p1 = gl_canon_fast(gl_redc(gl_wide_mul(gl_u64(a1), gl_u64(b1))));
    return gl_elem(gl_add_fast(gl_redc(gl_wide_mul(gl_u64(a2), gl_u64(b2))), p1));
  }
  // hi - lo for canonical hi, lo: canonical
  SC_HD static V diff(const V& hi, const V& lo) { return gl_elem(gl_sub_fast(gl_u64(hi), gl_u64(lo))); }
  // x + d for any x and canonical d
  SC_HD static V step(const V& x, const V& d) { return gl_elem(gl_add_fast(gl_u64(x), gl_u64(d))); }
  // lo + r (hi - lo) for canonical lo, hi and any r: canonical, so it can be stored and its
  // differences stay canonical
  SC_HD static V fold(const V& lo, const V& hi, const V& r) {
    const uint64_t l = gl_u64(lo);
    const uint64_t m = gl_redc(gl_wide_mul(gl_u64(r), gl_sub_fast(gl_u64(hi), l)));
    return gl_elem(gl_canon_fast(gl_add_fast(m, l)));
  }
  // lo + t d for a per-lane 0 <= t < 8 (d canonical)
  SC_HD static V at(const V& lo, const V& d, int t) {
    uint64_t x = gl_u64(lo);
    for (int s = 0; s < t; ++s) x = gl_add_fast(x, gl_u64(d));
    return gl_elem(x);
  }
  // Running sum W = c0 + c1 2^32 + c2 2^64 + c3 2^96 of 32-bit words per 64-bit column, whose
  // value is W 2^-64 mod p: a raw 128-bit product enters as its four words (one gl_mad each,
  // no carries), a value x as x 2^64. A column grows by less than 2^32 per addition.
  struct Acc {
    uint64_t c0, c1, c2, c3;
    SC_HD void clear() {
      c0 = 0;
      c1 = 0;
      c2 = 0;
      c3 = 0;
    }
    SC_HD void add_u64(uint64_t x) {
      c2 = gl_mad((uint32_t)x, 1u, c2);
      c3 = gl_mad((uint32_t)(x >> 32), 1u, c3);
    }
    SC_HD void add(const V& x) { add_u64(gl_u64(x)); }  // any x < 2^64
    SC_HD void sub(const V& x) {                          // adds p - canon(x) (0 stays 0)
      const uint64_t c = gl_canon_fast(gl_u64(x));
      add_u64(c == 0 ? 0 : GL_P - c);
    }
    // W += a b for any a, b < 2^64
    SC_HD void add_product(const V& a, const V& b) {
      const GlWide x = gl_wide_mul(gl_u64(a), gl_u64(b));
      c0 = gl_mad(x.x0, 1u, c0);
      c1 = gl_mad(x.x1, 1u, c1);
      c2 = gl_mad((uint32_t)x.hi, 1u, c2);
      c3 = gl_mad((uint32_t)(x.hi >> 32), 1u, c3);
    }
    // Columns carried into W = w3:w2:w1:w0 + top 2^128 (top < 2^34), then
    // W 2^-64 = redc(w) + top 2^64 with 2^64 = 2^32 - 1: top = t1 2^32 + t0 adds
    // t0 (2^32 - 1) + t1 (2^64 - 2^32) = t0 (2^32 - 1) - t1 (mod p).
    SC_HD Elem<2> value() const {
      uint64_t s = (c0 >> 32) + (uint32_t)c1;
      const uint32_t w0 = (uint32_t)c0, w1 = (uint32_t)s;
      s = (s >> 32) + (c1 >> 32) + (uint32_t)c2;
      const uint32_t w2 = (uint32_t)s;
      s = (s >> 32) + (c2 >> 32) + (uint32_t)c3;
      const uint32_t w3 = (uint32_t)s;
      const uint64_t top = (s >> 32) + (c3 >> 32);
      GlWide x;
      x.x0 = w0;
      x.x1 = w1;
      x.hi = (uint64_t)w2 | ((uint64_t)w3 << 32);
      uint64_t r = gl_canon(gl_redc(x));
      r = gl_add(r, gl_canon((uint64_t)(uint32_t)top * GL_EPS));
      return gl_elem(gl_sub(r, top >> 32));
    }
  };
};

// Host: Montgomery coefficient -> the word form Work<F>::coeff reads.
template <class F>
inline Elem<F::L> work_coeff(const Elem<F::L>& c) {
  return c;
}

// ---------------------------------------------------------------------------------------------
// R29: BLS12-381 scalar field in nine 29-bit limbs, Montgomery radix R' = 2^261.
//
// A residue x is held as any value congruent to x * 2^261 (mod p), limbs little-endian; the
// Montgomery image X = x * 2^256 of the caller is 32 X in this form, so conversion in is a
// shifted unpacking and conversion out divides by 32 mod p. Values are not kept in [0, p):
// the bounds each operation needs are stated with it and hold for every use in the kernels
// (x(t) < 29p for t <= 7, products < 13p, sums reduced before they can overflow).
//   mul:   a, b with limbs < 2^31, one of them < 2^29 (or both < 2^30): a column sum of 9
//          products of a*b, 8 of m*p, m_k and the carry stays below 2^63.5. Result limbs
//          0..7 < 2^29, value < a*b/2^261 + p.
//   reduce: any limbs < 2^30, top limb < 2^32 -> value < p + 2^241 (< 2^256, so it packs
//          into eight 32-bit words), normalized.
// ---------------------------------------------------------------------------------------------
struct R29 {
  uint32_t l[9];
};

namespace r29 {
constexpr uint32_t MASK = (1u << 29) - 1;

SC_HD constexpr uint32_t P(int i) {  // p in radix 2^29
  return i == 0   ? 0x00000001u
         : i == 1 ? 0x1FFFFFF8u
         : i == 2 ? 0x1F96FFBFu
         : i == 3 ? 0x1B4805FFu
         : i == 4 ? 0x1D80553Bu
         : i == 5 ? 0x0C0404D0u
         : i == 6 ? 0x1520CCE7u
         : i == 7 ? 0x0A6533AFu
                  : 0x0073EDA7u;
}

SC_HD constexpr uint32_t NP(int i) {  // 2^261 - p in radix 2^29
  return i == 0   ? 0x1FFFFFFFu
         : i == 1 ? 0x00000007u
         : i == 2 ? 0x00690040u
         : i == 3 ? 0x04B7FA00u
         : i == 4 ? 0x027FAAC4u
         : i == 5 ? 0x13FBFB2Fu
         : i == 6 ? 0x0ADF3318u
         : i == 7 ? 0x159ACC50u
                  : 0x1F8C1258u;
}

// limb i of k*p in radix 2^29 (k*p < 2^261 for k <= 70)
SC_HD constexpr uint32_t KP(uint32_t k, int i) {
  uint64_t c = 0;
  uint32_t out = 0;
  for (int j = 0; j <= i; ++j) {
    c += (uint64_t)k * P(j);
    out = (uint32_t)(c & MASK);
    c >>= 29;
  }
  return i == 8 ? out + (uint32_t)(c << 29) : out;
}

// k*p written with every limb in [2^29 - 1, 2^30): limb i lends 1 to limb i-1, so subtracting
// any normalized b with value(b) < (k-1) p limb by limb leaves no negative limb.
SC_HD constexpr uint32_t PAD(uint32_t k, int i) {
  return i == 0 ? KP(k, 0) + (1u << 29) : (i < 8 ? KP(k, i) + MASK : KP(k, 8) - 1u);
}

// q = floor(t / (p_top + 1)) for t < 2^32, p_top = p >> 232 = 0x73EDA7 (exact multiply-high).
SC_HD uint32_t top_quotient(uint32_t t) {
  return (uint32_t)(((uint64_t)t * 0x8D542469u) >> 54);
}
}  // namespace r29

SC_HD R29 r29_zero() {
  R29 z;
#pragma unroll
  for (int i = 0; i < 9; ++i) z.l[i] = 0u;
  return z;
}

// Carries of limbs 0..7 moved up; the top limb keeps the excess.
SC_HD R29 r29_normalize(R29 x) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    x.l[i + 1] += x.l[i] >> 29;
    x.l[i] &= r29::MASK;
  }
  return x;
}

// x - q p with q from the top limb, computed as x + q (2^261 - p) with the q * 2^261 dropped.
SC_HD R29 r29_reduce(const R29& x) {
  const uint32_t q = r29::top_quotient(x.l[8]);
  R29 r;
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    c += (uint64_t)q * r29::NP(i) + x.l[i];
    r.l[i] = (uint32_t)c & r29::MASK;
    c >>= 29;
  }
  return r;
}

// Montgomery product a * b * 2^-261 mod p by product scanning (column k of a*b and of m*p
// summed in 64 bits; m_k clears the column's low 29 bits). p = 1 (mod 2^29), so
// -p^-1 = -1 and m_k = -column mod 2^29.
//
// Each of the low nine columns is summed with 2^29 - 1 added (s = column + 2^29 - 1): the
// column plus m_k is a multiple of 2^29, so its carry (column + m_k) / 2^29 is the column
// over 2^29 rounded up, which is s >> 29; and -column = -(s + 1) = ~s (mod 2^29) gives m_k.
// This saves the negation's separate add of m_k (and the zero upper word it needs) per column.
SC_HD R29 r29_mul(const R29& a, const R29& b) {
  uint32_t m[9];
  R29 r;
  uint64_t acc = 0;
  // One multiply-add chain per column (independent products of the caller interleave with
  // it); m_k * p_0 = m_k since p_0 = 1.
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    uint64_t s = acc + r29::MASK;
#pragma unroll
    for (int i = 0; i <= k; ++i) s += (uint64_t)a.l[i] * b.l[k - i];
#pragma unroll
    for (int i = 0; i < k; ++i) s += (uint64_t)m[i] * r29::P(k - i);
    m[k] = ~(uint32_t)s & r29::MASK;
    acc = s >> 29;
  }
#pragma unroll
  for (int k = 9; k < 17; ++k) {
    uint64_t s = acc;
#pragma unroll
    for (int i = k - 8; i < 9; ++i) {
      s += (uint64_t)a.l[i] * b.l[k - i];
      s += (uint64_t)m[i] * r29::P(k - i);
    }
    r.l[k - 9] = (uint32_t)s & r29::MASK;
    acc = s >> 29;
  }
  r.l[8] = (uint32_t)acc;
  return r;
}

// (a1 b1 + a2 b2) * 2^-261 mod p with one reduction: both products summed per column (18
// products, 8 of m*p, m_k and the carry: below 2^62.7 for limbs < 2^29). Value below
// (a1 b1 + a2 b2)/2^261 + p.
SC_HD R29 r29_mul_sum(const R29& a1, const R29& b1, const R29& a2, const R29& b2) {
  uint32_t m[9];
  R29 r;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    uint64_t s = acc + r29::MASK;  // carry and m_k as in r29_mul
#pragma unroll
    for (int i = 0; i <= k; ++i) {
      s += (uint64_t)a1.l[i] * b1.l[k - i];
      s += (uint64_t)a2.l[i] * b2.l[k - i];
    }
#pragma unroll
    for (int i = 0; i < k; ++i) s += (uint64_t)m[i] * r29::P(k - i);
    m[k] = ~(uint32_t)s & r29::MASK;
    acc = s >> 29;
  }
#pragma unroll
  for (int k = 9; k < 17; ++k) {
    uint64_t s = acc;
#pragma unroll
    for (int i = k - 8; i < 9; ++i) {
      s += (uint64_t)a1.l[i] * b1.l[k - i];
      s += (uint64_t)a2.l[i] * b2.l[k - i];
      s += (uint64_t)m[i] * r29::P(k - i);
    }
    r.l[k - 9] = (uint32_t)s & r29::MASK;
    acc = s >> 29;
  }
  r.l[8] = (uint32_t)acc;
  return r;
}

// ---------------------------------------------------------------------------------------------
// Folding by the round's challenge without a Montgomery reduction. Every fold of a round
// multiplies by the same r, so for D = sum_k D_k 2^(29k)
//   r D = sum_k D_k (r 2^(29k) mod p)   (mod p):
// nine rows of nine limbs (the fold table, built once per block from r, fold_table_row) and 81
// multiply-adds into nine 64-bit columns, the base value (lo) added in. With D_k < 2^31 and the
// table canonical, each column stays below 2^63.2 and the sum V below (2^34.2 + 33) p. The
// quotient q = floor(V / p - 2^-16) comes from the top columns in double precision: the
// terms dropped (column 7's low word and below) and the rounding of three fused multiply-adds
// move the estimate by less than 2^-16.6 down and 2^-17.1 up, so 0 <= V - q p < p (1 + 2^-14.9).
// V - q p is formed as V + q (2^261 - p) with the bits from 2^261 up, which hold exactly q,
// dropped: 17 multiply-adds and one carry pass. The result has normalized limbs and a value
// below p + 2^240 (r29_reduce's output bound, p + 2^241, holds).
// ---------------------------------------------------------------------------------------------
namespace r29 {
constexpr int TAB_STRIDE = 12;  // words per fold-table row: rows start 16-byte aligned
constexpr double K264 = 0x1.1aa84a76ff6f2p+9;   // 2^264 / p, rounded to nearest
constexpr double K232 = 0x1.1aa84a76ff6f2p-23;  // 2^232 / p
constexpr double K235 = 0x1.1aa84a76ff6f2p-20;  // 2^235 / p
}  // namespace r29

// Columns c (value V = sum_i c_i 2^(29i), bounds above) -> V mod p as a normalized R29 below
// p + 2^240.
SC_HD R29 r29_fold_finish(uint64_t (&c)[9]) {
  const double h8 = (double)(uint32_t)(c[8] >> 32), l8 = (double)(uint32_t)c[8];
  const double h7 = (double)(uint32_t)(c[7] >> 32);
  double q = fma(h7, r29::K235, -0x1p-16);
  q = fma(l8, r29::K232, q);
  q = fma(h8, r29::K264, q);
  q = floor(fmax(q, 0.0));
  const double q1d = floor(q * 0x1p-29);  // q < 2^34.3: q = q1 2^29 + q0, both exact
  const uint32_t q1 = (uint32_t)q1d, q0 = (uint32_t)fma(q1d, -0x1p29, q);
#pragma unroll
  for (int i = 0; i < 9; ++i) c[i] += (uint64_t)q0 * r29::NP(i);
#pragma unroll
  for (int i = 1; i < 9; ++i) c[i] += (uint64_t)q1 * r29::NP(i - 1);
  R29 x;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    c[i + 1] += c[i] >> 29;
    x.l[i] = (uint32_t)c[i] & r29::MASK;
  }
  x.l[8] = (uint32_t)c[8] & r29::MASK;
  return x;
}

// 2 x + 2 y - z for products x, y, z (limbs 0..7 below 2^29, values below 2p; the round-0
// identity of the compiled a*b*c compositions): value below 7.5 p, normalized, so it can be the
// first factor of a product.
SC_HD R29 r29_two_x_plus_two_y_minus_z(const R29& x, const R29& y, const R29& z) {
  R29 s;
#pragma unroll
  for (int i = 0; i < 9; ++i) s.l[i] = 2u * x.l[i] + 2u * y.l[i] + r29::PAD(3, i) - z.l[i];
  return r29_normalize(s);
}

// Eight 32-bit words W -> limbs of W * 2^S (S = 0: fold buffers; S = 5: Montgomery input,
// turning X = x 2^256 into 32 X = x 2^261).
template <int S>
SC_HD R29 r29_unpack(const uint32_t* w) {
  R29 r;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int lo = 29 * k - S;
    uint32_t x;
    if (lo < 0) {
      x = w[0] << S;
    } else {
      const int q = lo >> 5, b = lo & 31;
      x = w[q] >> b;
      if (b > 3 && q + 1 < 8) x |= w[q + 1] << (32 - b);
    }
    r.l[k] = x & r29::MASK;
  }
  return r;
}

// Bits [S, S + 256) of a normalized value as eight 32-bit words.
template <int S>
SC_HD void r29_pack(uint32_t* w, const R29& x) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int lo = 32 * j + S;
    const int k = lo / 29, b = lo % 29;
    uint32_t y = x.l[k] >> b;
    if (k + 1 < 9) y |= x.l[k + 1] << (29 - b);
    if (k + 2 < 9 && 58 - b < 32) y |= x.l[k + 2] << (58 - b);
    w[j] = y;
  }
}

// x * 2^-5 mod p in [0, p) as Montgomery limbs (x < 2p, normalized): x + k p with
// k = -x mod 32 (p = 1 mod 32) is divisible by 32, and (x + k p) / 32 < 1.04 p.
SC_HD Elem<8> r29_to_mont(const R29& x) {
  const uint32_t k = (0u - x.l[0]) & 31u;
  R29 u;
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    c += (uint64_t)k * r29::P(i) + x.l[i];
    u.l[i] = (uint32_t)c & r29::MASK;
    c >>= 29;
  }
  Elem<8> y;
  r29_pack<5>(y.v, u);
  Elem<8> t;
  const uint32_t borrow = sub_p<FieldP256>(t, y);
#pragma unroll
  for (int i = 0; i < 8; ++i) y.v[i] = borrow ? y.v[i] : t.v[i];
  return y;
}

// Row k of the fold table of the challenge whose Montgomery words are rm: the limbs of
// r 2^(29k + S) mod p, canonical. S = 5 for folds straight from the caller's Montgomery input,
// whose working value is 32 X.
template <int S>
SC_HD void fold_table_row(uint32_t* row, const uint32_t* rm, int k) {
  Elem<8> x, c;
#pragma unroll
  for (int i = 0; i < 8; ++i) x.v[i] = rm[i];
  const int e = 29 * k + S;  // below 254, so 2^e < p
#pragma unroll
  for (int i = 0; i < 8; ++i) c.v[i] = i == (e >> 5) ? 1u << (e & 31) : 0u;  // no indexed register array
  const Elem<8> t = mont_mul<FieldP256>(x, c);  // X 2^e 2^-256 = r 2^e
  const R29 l = r29_unpack<0>(t.v);
#pragma unroll
  for (int j = 0; j < 9; ++j) row[j] = l.l[j];
}

template <>
struct Work<FieldP256> {
  static constexpr int L = 8;
  static constexpr bool kMontIsStored = false;
  static constexpr bool kCheapMontInput = true;
  static constexpr bool kFoldTable = true;
  static constexpr bool kLazyStep = true;
  using V = R29;
  // Words per fold table (9 rows of r29::TAB_STRIDE), and its rows for a challenge: row k of
  // the table the folds of stored entries use (S = 0) or of folds straight from the caller's
  // Montgomery input (S = 5).
  static constexpr int kFoldTableWords = 9 * r29::TAB_STRIDE;
  template <int S>
  SC_HD static void fold_row(uint32_t* tab, const uint32_t* rm, int k) {
    fold_table_row<S>(tab + k * r29::TAB_STRIDE, rm, k);
  }
  // One fold lo + r (hi - lo) of words lo_w, hi_w (as for fold2_tab), r given by its fold table.
  template <bool MONT>
  SC_HD static V fold1_tab(const uint32_t* lo_w, const uint32_t* hi_w, const uint32_t* tab) {
    uint32_t D[1][9];
    uint64_t c[1][9];
    const R29 lo = r29_unpack<0>(lo_w), hi = r29_unpack<0>(hi_w);
    const R29 base = MONT ? r29_unpack<5>(lo_w) : lo;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      D[0][i] = hi.l[i] + r29::PAD(3, i) - lo.l[i];
      c[0][i] = base.l[i];
    }
    tab_columns<1>(c, D, tab);
    return r29_fold_finish(c[0]);
  }
  // Both folded entries of one table of a pair, e = w0 + r (w1 - w0) and o = w2 + r (w3 - w2),
  // from its four entries' words w (stored working values, or with MONT the caller's
  // Montgomery words), r given by its fold table tab (16-byte aligned). One fold after the
  // other: folding both in one pass over the table rows reads the rows once, but its live
  // columns took the a*b*c + d*e pass kernel from 231 to 256 VGPRs (one wave per SIMD).
  template <bool MONT>
  SC_HD static void fold2_tab(const uint32_t* w, const uint32_t* tab, V& e, V& o) {
    e = fold1_tab<MONT>(w, w + 8, tab);
    o = fold1_tab<MONT>(w + 16, w + 24, tab);
  }
  // c[n][i] += sum_k D[n][k] row_k[i]. On the device the table sits in LDS; its address is
  // passed through an empty asm statement per call so the compiler re-reads the rows at each
  // fold instead of hoisting all 81 words into registers for the whole loop.
  template <int N>
  SC_HD static void tab_columns(uint64_t (&c)[N][9], const uint32_t (&D)[N][9], const uint32_t* tab) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t z = 0;
    asm volatile("" : "+v"(z));
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4* t4 = reinterpret_cast<const u32x4*>(tab) + z;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const u32x4 a = t4[3 * k], b = t4[3 * k + 1], cc = t4[3 * k + 2];
      const uint32_t row[9] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, cc.x};
#pragma unroll
      for (int n = 0; n < N; ++n) {
#pragma unroll
        for (int i = 0; i < 9; ++i) c[n][i] += (uint64_t)D[n][k] * row[i];
      }
    }
#else
    for (int k = 0; k < 9; ++k) {
      for (int n = 0; n < N; ++n) {
        for (int i = 0; i < 9; ++i) c[n][i] += (uint64_t)D[n][k] * tab[k * r29::TAB_STRIDE + i];
      }
    }
#endif
  }
  SC_HD static V from_mont(const uint32_t* w) { return r29_reduce(r29_unpack<5>(w)); }
  // The caller's Montgomery X = x 2^256 < p taken as the working value itself, which stands
  // for the residue x 2^-5 (no shift, no reduction; bounds as a stored value).
  SC_HD static V from_mont_scaled(const uint32_t* w) { return r29_unpack<0>(w); }
  // 32 r for a working r below p + 2^241: five doublings, normalized (value < 32.1 p).
  SC_HD static V times32(V x) {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
#pragma unroll
      for (int i = 0; i < 9; ++i) x.l[i] += x.l[i];
      x = r29_normalize(x);
    }
    return x;
  }
  // Fold of two Montgomery input entries lo_w, hi_w with challenge r, given r32 = times32(r):
  // 32 X_lo + (32 r)(X_hi - X_lo) 2^-261 is the working value of x_lo + r (x_hi - x_lo), so
  // the inputs need no reduction of their own. X_hi + 3p - X_lo < 4p; the product is below
  // 32.1p * 4p / 2^261 + p < 3p, and the sum below 35p (top limb < 2^29), which the final
  // reduction brings below p + 2^241.
  SC_HD static V fold_mont(const uint32_t* lo_w, const uint32_t* hi_w, const V& r32) {
    const R29 lo = r29_unpack<0>(lo_w), hi = r29_unpack<0>(hi_w);
    R29 d;
#pragma unroll
    for (int i = 0; i < 9; ++i) d.l[i] = hi.l[i] + r29::PAD(3, i) - lo.l[i];
    const R29 m = r29_mul(r32, d);
    const R29 lo32 = r29_unpack<5>(lo_w);
    R29 sum;
#pragma unroll
    for (int i = 0; i < 9; ++i) sum.l[i] = lo32.l[i] + m.l[i];
    return r29_reduce(sum);
  }
  SC_HD static V from_stored(const uint32_t* w) { return r29_unpack<0>(w); }
  SC_HD static V coeff(const uint32_t* w) { return r29_unpack<0>(w); }
  SC_HD static void store(uint32_t* w, const V& x) { r29_pack<0>(w, x); }
  SC_HD static V mul(const V& a, const V& b) { return r29_mul(a, b); }
  SC_HD static V mul_sum(const V& a1, const V& b1, const V& a2, const V& b2) { return r29_mul_sum(a1, b1, a2, b2); }
  // c x for a small integer c, |c| <= 7, and x with limbs 0..7 below 2^29 and top limb below
  // 2^29 (products, coefficients, interpolated values): |c| x limb by limb (below 2^32),
  // normalized and reduced below p + 2^241; for c < 0, 3p minus that, normalized (value <= 3p).
  // Costs 9 multiplies where a product by the Montgomery image of c costs 153 multiply-adds.
  SC_HD static V scale_small(const V& x, int c) {
    const uint32_t m = (uint32_t)(c < 0 ? -c : c);
    R29 y;
#pragma unroll
    for (int i = 0; i < 9; ++i) y.l[i] = x.l[i] * m;
    y = r29_reduce(r29_normalize(y));
    if (c < 0) {
#pragma unroll
      for (int i = 0; i < 9; ++i) y.l[i] = r29::PAD(3, i) - y.l[i];
      y = r29_normalize(y);
    }
    return y;
  }
  // hi + 3p - lo for hi, lo < p + 2^241: value < 4p + 2^241.
  SC_HD static V diff(const V& hi, const V& lo) {
    R29 d;
#pragma unroll
    for (int i = 0; i < 9; ++i) d.l[i] = hi.l[i] + r29::PAD(3, i) - lo.l[i];
    return r29_normalize(d);
  }
  SC_HD static V step(const V& x, const V& d) {
    R29 s;
#pragma unroll
    for (int i = 0; i < 9; ++i) s.l[i] = x.l[i] + d.l[i];
    return r29_normalize(s);
  }
  // x + d with the limbs left as summed (x, d normalized: limbs below 2^30), for a value that
  // only enters products as their second operand (static_lazy_step).
  SC_HD static V step_lazy(const V& x, const V& d) {
    R29 s;
#pragma unroll
    for (int i = 0; i < 9; ++i) s.l[i] = x.l[i] + d.l[i];
    return s;
  }
  // lo + t d for 0 <= t < 8 (limbs stay below 2^32; the value below 29p as for t steps)
  SC_HD static V at(const V& lo, const V& d, int t) {
    R29 x;
#pragma unroll
    for (int i = 0; i < 9; ++i) x.l[i] = lo.l[i] + (uint32_t)t * d.l[i];
    return r29_normalize(x);
  }
  // lo + r (hi + 3p - lo), reduced below p + 2^241. The difference is passed to the product
  // with limbs < 2^31 (r's are < 2^29), and the sum's carries are resolved by the reduction.
  SC_HD static V fold(const V& lo, const V& hi, const V& r) {
    R29 d;
#pragma unroll
    for (int i = 0; i < 9; ++i) d.l[i] = hi.l[i] + r29::PAD(3, i) - lo.l[i];
    const R29 m = r29_mul(r, d);
    R29 s;
#pragma unroll
    for (int i = 0; i < 9; ++i) s.l[i] = lo.l[i] + m.l[i];
    return r29_reduce(s);
  }
  // Running sum of terms: products (< 13p) and single factors, which are interpolated
  // values x(t) < 29p + 2^244 (t <= 7). Limbs 0..7 normalized, the top limb reduced before it
  // can reach 2^32.
  struct Acc {
    R29 v;
    SC_HD void clear() { v = r29_zero(); }
    SC_HD void add(const R29& x) {
#pragma unroll
      for (int i = 0; i < 9; ++i) v.l[i] += x.l[i];
      v = r29_normalize(v);
      if (v.l[8] >= (1u << 31)) v = r29_reduce(v);
    }
    // Adds 33p - x for any normalized x below 32p: with a fresh (zero) accumulator the top
    // limb of 33p - x must not borrow, and x(7) of a single-factor term comes to 29p.
    SC_HD void sub(const R29& x) {
#pragma unroll
      for (int i = 0; i < 9; ++i) v.l[i] += r29::PAD(33, i) - x.l[i];
      v = r29_normalize(v);
      if (v.l[8] >= (1u << 31)) v = r29_reduce(v);
    }
    SC_HD Elem<8> value() const { return r29_to_mont(r29_reduce(v)); }
  };
};

// ---------------------------------------------------------------------------------------------
// R26: 2^127 - 1 in five 26-bit limbs holding the plain residue (not a Montgomery image).
//
// The Montgomery image of x is X = 2x mod p (R = 2^128 = 2), so conversion is a one-bit
// rotation within 127 bits each way. Products fold at 2^130 = 8 (mod p) and every result
// has its bits at and above 2^127 folded back (2^127 = 1), so values stay below 2^127 + 2^5
// with limbs below 2^26 + 2^5: column sums of five products stay far below 2^64.
// ---------------------------------------------------------------------------------------------
struct R26 {
  uint32_t l[5];
};

namespace r26 {
constexpr uint32_t MASK = (1u << 26) - 1;
constexpr uint32_t PAD2 = 0x7FFFFFEu;   // 2p with limbs 0..3 in [2^26, 2^27)
constexpr uint32_t PAD2_TOP = 0xFFFFFEu;
}  // namespace r26

// Carries of limbs 0..3 moved up, then bits 127 and above of the top limb folded into limb 0.
SC_HD R26 r26_normalize(R26 x) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    x.l[i + 1] += x.l[i] >> 26;
    x.l[i] &= r26::MASK;
  }
  x.l[0] += x.l[4] >> 23;
  x.l[4] &= (1u << 23) - 1;
  return x;
}

// Four words of a value below 2^128 -> limbs.
SC_HD R26 r26_unpack(const uint32_t* w) {
  R26 r;
  r.l[0] = w[0] & r26::MASK;
  r.l[1] = ((w[0] >> 26) | (w[1] << 6)) & r26::MASK;
  r.l[2] = ((w[1] >> 20) | (w[2] << 12)) & r26::MASK;
  r.l[3] = ((w[2] >> 14) | (w[3] << 18)) & r26::MASK;
  r.l[4] = w[3] >> 8;
  return r;
}

// Limbs (each below 2^27, value below 2^128) -> four words, carries resolved on the way.
SC_HD void r26_pack(uint32_t* w, const R26& x) {
  uint64_t acc = x.l[0] + ((uint64_t)x.l[1] << 26);
  w[0] = (uint32_t)acc;
  acc = (acc >> 32) + ((uint64_t)x.l[2] << 20);
  w[1] = (uint32_t)acc;
  acc = (acc >> 32) + ((uint64_t)x.l[3] << 14);
  w[2] = (uint32_t)acc;
  acc = (acc >> 32) + ((uint64_t)x.l[4] << 8);
  w[3] = (uint32_t)acc;
}

// The nine column sums of a * b (each below 2^57 for limbs below 2^27).
SC_HD void r26_columns(uint64_t (&c)[9], const R26& a, const R26& b) {
#pragma unroll
  for (int k = 0; k < 9; ++k) {
#pragma unroll
    for (int i = (k < 5 ? 0 : k - 4); i <= (k < 5 ? k : 4); ++i) c[k] += (uint64_t)a.l[i] * b.l[k - i];
  }
}

// Column sums -> lazily reduced residue: the upper five columns folded down with 2^130 = 8.
SC_HD R26 r26_reduce_columns(uint64_t (&c)[9]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) c[k] += c[k + 5] << 3;
  R26 r;
  uint64_t carry = 0;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const uint64_t s = c[k] + carry;
    r.l[k] = (uint32_t)s & r26::MASK;
    carry = s >> 26;
  }
  // carry has weight 2^130 = 8; it is below 2^35, so 8 * carry fits after one more step
  const uint64_t s0 = (uint64_t)r.l[0] + (carry << 3);
  r.l[0] = (uint32_t)s0 & r26::MASK;
  r.l[1] += (uint32_t)(s0 >> 26);
  r.l[0] += r.l[4] >> 23;
  r.l[4] &= (1u << 23) - 1;
  return r;
}

// a * b mod p (lazily reduced).
SC_HD R26 r26_mul(const R26& a, const R26& b) {
  uint64_t c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  r26_columns(c, a, b);
  return r26_reduce_columns(c);
}

// a1 b1 + a2 b2 mod p with one reduction (column sums below 2^58).
SC_HD R26 r26_mul_sum(const R26& a1, const R26& b1, const R26& a2, const R26& b2) {
  uint64_t c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  r26_columns(c, a1, b1);
  r26_columns(c, a2, b2);
  return r26_reduce_columns(c);
}

// Plain residue (value below 2^127 + 2^5) -> Montgomery limbs 2x mod p in [0, p).
SC_HD Elem<4> r26_to_mont(const R26& x) {
  uint32_t w[4];
  r26_pack(w, r26_normalize(x));
  // value v <= 2^127 + small: v >= p  <=>  v + 1 >= 2^127
  uint64_t c = (uint64_t)w[0] + 1u;
  uint32_t y[4];
  y[0] = (uint32_t)c;
#pragma unroll
  for (int i = 1; i < 4; ++i) {
    c = (c >> 32) + w[i];
    y[i] = (uint32_t)c;
  }
  if (y[3] >> 31) {  // v - p = v + 1 - 2^127
    w[0] = y[0];
    w[1] = y[1];
    w[2] = y[2];
    w[3] = y[3] & 0x7FFFFFFFu;
  }
  Elem<4> out;  // rotate left by one within 127 bits
  out.v[0] = (w[0] << 1) | (w[3] >> 30);
  out.v[1] = (w[1] << 1) | (w[0] >> 31);
  out.v[2] = (w[2] << 1) | (w[1] >> 31);
  out.v[3] = ((w[3] << 1) | (w[2] >> 31)) & 0x7FFFFFFFu;
  return out;
}

template <>
struct Work<FieldP128> {
  static constexpr int L = 4;
  static constexpr bool kMontIsStored = false;
  static constexpr bool kCheapMontInput = false;
  static constexpr bool kFoldTable = false;
  static constexpr bool kLazyStep = false;
  using V = R26;
  // Montgomery X = 2x (X < p): x is X rotated right by one within 127 bits.
  SC_HD static V from_mont(const uint32_t* w) {
    uint32_t x[4];
    x[0] = (w[0] >> 1) | (w[1] << 31);
    x[1] = (w[1] >> 1) | (w[2] << 31);
    x[2] = (w[2] >> 1) | (w[3] << 31);
    x[3] = (w[3] >> 1) | ((w[0] & 1u) << 30);
    return r26_unpack(x);
  }
  SC_HD static V from_stored(const uint32_t* w) { return r26_unpack(w); }
  SC_HD static V coeff(const uint32_t* w) { return r26_unpack(w); }
  SC_HD static void store(uint32_t* w, const V& x) { r26_pack(w, x); }
  SC_HD static V mul(const V& a, const V& b) { return r26_mul(a, b); }
  SC_HD static V mul_sum(const V& a1, const V& b1, const V& a2, const V& b2) { return r26_mul_sum(a1, b1, a2, b2); }
  SC_HD static V diff(const V& hi, const V& lo) {  // hi + 2p - lo
    R26 d;
#pragma unroll
    for (int i = 0; i < 4; ++i) d.l[i] = hi.l[i] + r26::PAD2 - lo.l[i];
    d.l[4] = hi.l[4] + r26::PAD2_TOP - lo.l[4];
    return r26_normalize(d);
  }
  SC_HD static V step(const V& x, const V& d) {
    R26 s;
#pragma unroll
    for (int i = 0; i < 5; ++i) s.l[i] = x.l[i] + d.l[i];
    return r26_normalize(s);
  }
  SC_HD static V at(const V& lo, const V& d, int t) {  // 0 <= t < 8
    R26 x;
#pragma unroll
    for (int i = 0; i < 5; ++i) x.l[i] = lo.l[i] + (uint32_t)t * d.l[i];
    return r26_normalize(x);
  }
  SC_HD static V fold(const V& lo, const V& hi, const V& r) {
    const R26 m = r26_mul(r, diff(hi, lo));
    R26 s;
#pragma unroll
    for (int i = 0; i < 5; ++i) s.l[i] = lo.l[i] + m.l[i];
    return r26_normalize(s);
  }
  struct Acc {
    R26 v;
    SC_HD void clear() {
#pragma unroll
      for (int i = 0; i < 5; ++i) v.l[i] = 0u;
    }
    SC_HD void add(const R26& x) {
#pragma unroll
      for (int i = 0; i < 5; ++i) v.l[i] += x.l[i];
      v = r26_normalize(v);
    }
    SC_HD void sub(const R26& x) {
      R26 s;
#pragma unroll
      for (int i = 0; i < 4; ++i) s.l[i] = v.l[i] + r26::PAD2 - x.l[i];
      s.l[4] = v.l[4] + r26::PAD2_TOP - x.l[4];
      v = r26_normalize(s);
    }
    SC_HD Elem<4> value() const { return r26_to_mont(v); }
  };
};

template <>
inline Elem<4> work_coeff<FieldP128>(const Elem<4>& c) {
  // the plain residue c / 2 mod p, as stored words
  Elem<4> x;
  x.v[0] = (c.v[0] >> 1) | (c.v[1] << 31);
  x.v[1] = (c.v[1] >> 1) | (c.v[2] << 31);
  x.v[2] = (c.v[2] >> 1) | (c.v[3] << 31);
  x.v[3] = (c.v[3] >> 1) | ((c.v[0] & 1u) << 30);
  return x;
}

template <>
inline Elem<8> work_coeff<FieldP256>(const Elem<8>& c) {
  // 32 c mod p: the R29 image of the Montgomery value c, below p so it packs.
  Elem<8> x = c;
  for (int i = 0; i < 5; ++i) x = add<FieldP256>(x, x);
  return x;
}

}  // namespace sc
