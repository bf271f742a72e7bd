// Host build of the device field templates (sumcheck-triton_amd/csrc/field.hpp), so the
// arithmetic can be checked against Python integers without a GPU. Test-only.
#include <stdint.h>

#include "field.hpp"
#include "keccak.hpp"
#include "prng.hpp"
#include "work.hpp"

using namespace sc;

template <class F>
static void run(int op, uint64_t n, const uint32_t* a, const uint32_t* b, uint32_t* out) {
  constexpr int L = F::L;
  for (uint64_t i = 0; i < n; ++i) {
    Elem<L> x, y, z;
    for (int k = 0; k < L; ++k) {
      x.v[k] = a[i * L + k];
      y.v[k] = b[i * L + k];
    }
    if (op == 0) z = mont_mul<F>(x, y);
    else if (op == 1) z = add<F>(x, y);
    else if (op == 2) z = sub<F>(x, y);
    else if (op == 3) z = mont_mul_cios<F>(x, y);
    else z = mont_mul_fips<F>(x, y);
    for (int k = 0; k < L; ++k) out[i * L + k] = z.v[k];
  }
}

extern "C" void field_host_op(int bit_width, int op, uint64_t n, const uint32_t* a, const uint32_t* b,
                              uint32_t* out) {
  switch (bit_width) {
    case 32: run<FieldP32>(op, n, a, b, out); break;
    case 64: run<FieldP64>(op, n, a, b, out); break;
    case 128: run<FieldP128>(op, n, a, b, out); break;
    default: run<FieldP256>(op, n, a, b, out); break;
  }
}

// Working-form path of the prover kernels (csrc/work.hpp), on Montgomery inputs a, b, c:
//   op 0: a*b                         op 1: a + c (b - a), through fold, store and reload
//   op 2: -(c * x^3) with x = a + 7 (b - a) reached by seven interpolation steps
//   op 3: sum of 40 products a*b minus one, through the accumulator (39 a*b)
//   op 4: a*b - c*x(t) for every t in 0..7, summed (x(t) = a + t (b - a))
//   op 5: a x^2 + c a, x = a + 7 (b - a), the last two products sharing one reduction
//   op 6: -x(t) into a fresh accumulator, t = i mod 8, x(t) reached by interpolation steps
//   op 7: -x(t) into a fresh accumulator, t = i mod 8, x(t) = at(a, b - a, t)
//   op 8: a + c (b - a) folded straight from the Montgomery words of a and b (fold_mont),
//         through store and reload (fields with a cheaper Montgomery-input path; else op 1)
//   op 9: a read scaled (from_mont_scaled: the residue times 2^-5; else from_mont)
//   op 10: a x3 + (a x2 + c x2) with x2, x3 = x(2), x(3) left unnormalized (step_lazy), the
//          largest second operands the compiled compositions pass to mul and mul_sum
//   op 11: a(2) b(2) b with a(0) = a, a(1) = b, b(0) = c, b(1) = a; at 255 bits through the
//          round-0 identity 2 a1 b1 - a0 b0 + 2 (a1 - a0)(b1 - b0) (r29_two_x_plus_two_y_minus_z)
template <class F>
static void run_work(int op, uint64_t n, const uint32_t* a, const uint32_t* b, const uint32_t* c, uint32_t* out) {
  using W = Work<F>;
  constexpr int L = F::L;
  for (uint64_t i = 0; i < n; ++i) {
    const typename W::V x = W::from_mont(a + i * L), y = W::from_mont(b + i * L), z = W::from_mont(c + i * L);
    typename W::Acc acc;
    acc.clear();
    if (op == 0) {
      acc.add(W::mul(x, y));
    } else if (op == 1) {
      uint32_t w[L];
      W::store(w, W::fold(x, y, z));
      acc.add(W::from_stored(w));
    } else if (op == 2) {
      const typename W::V d = W::diff(y, x);
      typename W::V t = y;
      for (int k = 2; k <= 7; ++k) t = W::step(t, d);
      acc.sub(W::mul(W::mul(W::mul(t, t), t), z));
    } else if (op == 3) {
      const typename W::V pr = W::mul(x, y);
      for (int k = 0; k < 40; ++k) acc.add(pr);
      acc.sub(pr);
    } else if (op == 5) {
      const typename W::V d = W::diff(y, x);
      typename W::V t = y;
      for (int k = 2; k <= 7; ++k) t = W::step(t, d);
      acc.add(W::mul_sum(W::mul(x, t), t, z, x));  // a x7^2 + c a with x7 = a + 7 (b - a)
    } else if (op == 8) {
      uint32_t w[L];
      if constexpr (W::kCheapMontInput) {
        W::store(w, W::fold_mont(a + i * L, b + i * L, W::times32(z)));
      } else {
        W::store(w, W::fold(x, y, z));
      }
      acc.add(W::from_stored(w));
    } else if (op == 9) {
      if constexpr (W::kCheapMontInput) {
        acc.add(W::from_mont_scaled(a + i * L));
      } else {
        acc.add(x);
      }
    } else if (op == 10) {
      const typename W::V d = W::diff(y, x);
      typename W::V x2, x3;
      if constexpr (W::kLazyStep) {
        x2 = W::step_lazy(y, d);
        x3 = W::step_lazy(x2, d);
      } else {
        x2 = W::step(y, d);
        x3 = W::step(x2, d);
      }
      acc.add(W::mul(x, x3));
      acc.add(W::mul_sum(x, x2, z, x2));
    } else if (op == 11) {
      // a(2) b(2) c from a(0) = a, a(1) = b, b(0) = c, b(1) = a, c = b as round 0 of the
      // compiled a*b*c compositions forms it at 255 bits: 2 a1 b1 - a0 b0 + 2 (a1 - a0)(b1 - b0)
      if constexpr (W::kFoldTable) {
        const typename W::V ab0 = W::mul(x, z), ab1 = W::mul(y, x);
        const typename W::V dadb = W::mul(W::diff(y, x), W::diff(x, z));
        acc.add(W::mul(r29_two_x_plus_two_y_minus_z(ab1, dadb, ab0), y));
      } else {
        const typename W::V a2 = W::step(y, W::diff(y, x)), b2 = W::step(x, W::diff(x, z));
        acc.add(W::mul(W::mul(a2, b2), y));
      }
    } else if (op == 6 || op == 7) {
      // a single-factor term with coefficient -1 is the first term a thread adds: the
      // subtraction must hold for the largest interpolated value on its own
      const int tt = (int)(i % 8);
      const typename W::V d = W::diff(y, x);
      typename W::V t = x;
      if (op == 7) {
        t = W::at(x, d, tt);
      } else {
        for (int k = 1; k <= tt; ++k) t = k == 1 ? y : W::step(t, d);
      }
      acc.sub(t);
    } else {
      const typename W::V d = W::diff(y, x);
      typename W::V t = x;
      for (int k = 0; k <= 7; ++k) {
        if (k == 1) t = y;
        else if (k > 1) t = W::step(t, d);
        acc.add(W::mul(x, y));
        acc.sub(t);                // a single factor: the largest value a term can take
        acc.sub(W::mul(t, z));
        acc.add(t);
      }
    }
    const Elem<L> r = acc.value();
    for (int k = 0; k < L; ++k) out[i * L + k] = r.v[k];
  }
}

extern "C" void field_host_work(int bit_width, int op, uint64_t n, const uint32_t* a, const uint32_t* b,
                                const uint32_t* c, uint32_t* out) {
  switch (bit_width) {
    case 32: run_work<FieldP32>(op, n, a, b, c, out); break;
    case 64: run_work<FieldP64>(op, n, a, b, c, out); break;
    case 128: run_work<FieldP128>(op, n, a, b, c, out); break;
    default: run_work<FieldP256>(op, n, a, b, c, out); break;
  }
}

// Goldilocks raw words: out[i] = goldilocks_mont_mul_sum(a1, b1, a2, b2) for any 64-bit
// operands (the unreduced values the kernels pass it), as a 64-bit word.
extern "C" void gl_host_mul_sum(uint64_t n, const uint64_t* a1, const uint64_t* b1, const uint64_t* a2,
                                const uint64_t* b2, uint64_t* out) {
  for (uint64_t i = 0; i < n; ++i) {
    out[i] = gl_u64(goldilocks_mont_mul_sum(gl_elem(a1[i]), gl_elem(b1[i]), gl_elem(a2[i]), gl_elem(b2[i])));
  }
}

// Radix-2^29 helpers of the 255-bit field on raw limbs: op 0 unpack<0> then pack<0> (identity
// for values < 2^256); op 1 top-limb reduction of the nine limbs in a (limbs < 2^30, top < 2^32);
// op 2 Work<FieldP256>::scale_small on rows of ten words (nine limbs, then c as int32), the
// result's nine limbs in the row's first nine words.
extern "C" void r29_host(int op, uint64_t n, const uint32_t* a, uint32_t* out) {
  for (uint64_t i = 0; i < n; ++i) {
    if (op == 2) {
      R29 x;
      for (int k = 0; k < 9; ++k) x.l[k] = a[i * 10 + k];
      const R29 r = Work<FieldP256>::scale_small(x, (int32_t)a[i * 10 + 9]);
      for (int k = 0; k < 9; ++k) out[i * 10 + k] = r.l[k];
    } else if (op == 0) {
      r29_pack<0>(out + i * 8, r29_unpack<0>(a + i * 8));
    } else {
      R29 x;
      for (int k = 0; k < 9; ++k) x.l[k] = a[i * 9 + k];
      const R29 r = r29_reduce(x);
      for (int k = 0; k < 9; ++k) out[i * 9 + k] = r.l[k];
    }
  }
}

// Bit interleaving of the wave Keccak (csrc/keccak.hpp): the halves of each word and the join.
// 255-bit fold by the challenge's fold table (Work<FieldP256>::fold2_tab / fold1_tab) on raw
// words: lo, hi hold 8 words per entry (stored working values, or with mont != 0 the caller's
// Montgomery words), rm the challenge's Montgomery words; out gets the result's nine limbs.
// Entries are taken two at a time (fold2_tab); an odd last one goes through fold1_tab.
extern "C" void r29_fold_tab_host(int mont, uint64_t n, const uint32_t* lo, const uint32_t* hi, const uint32_t* rm,
                                  uint32_t* out) {
  using W = Work<FieldP256>;
  uint32_t tab[W::kFoldTableWords];
  for (int k = 0; k < 9; ++k) {
    if (mont) W::fold_row<5>(tab, rm, k);
    else W::fold_row<0>(tab, rm, k);
  }
  uint64_t i = 0;
  for (; i + 1 < n; i += 2) {
    uint32_t w[32];
    for (int k = 0; k < 8; ++k) {
      w[k] = lo[i * 8 + k];
      w[8 + k] = hi[i * 8 + k];
      w[16 + k] = lo[(i + 1) * 8 + k];
      w[24 + k] = hi[(i + 1) * 8 + k];
    }
    R29 e, o;
    if (mont) W::fold2_tab<true>(w, tab, e, o);
    else W::fold2_tab<false>(w, tab, e, o);
    for (int k = 0; k < 9; ++k) {
      out[i * 9 + k] = e.l[k];
      out[(i + 1) * 9 + k] = o.l[k];
    }
  }
  if (i < n) {
    const R29 e = mont ? W::fold1_tab<true>(lo + i * 8, hi + i * 8, tab) : W::fold1_tab<false>(lo + i * 8, hi + i * 8, tab);
    for (int k = 0; k < 9; ++k) out[i * 9 + k] = e.l[k];
  }
}

// V mod p from L word-column sums per value (reduce_word_columns).
template <class F>
static void word_columns(uint64_t n, const uint64_t* cols, uint32_t* out) {
  constexpr int L = F::L;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t c[L];
    for (int k = 0; k < L; ++k) c[k] = cols[i * L + k];
    const Elem<L> x = reduce_word_columns<F>(c);
    for (int k = 0; k < L; ++k) out[i * L + k] = x.v[k];
  }
}

extern "C" void word_columns_host(int bit_width, uint64_t n, const uint64_t* cols, uint32_t* out) {
  switch (bit_width) {
    case 32: word_columns<FieldP32>(n, cols, out); break;
    case 64: word_columns<FieldP64>(n, cols, out); break;
    case 128: word_columns<FieldP128>(n, cols, out); break;
    default: word_columns<FieldP256>(n, cols, out); break;
  }
}

extern "C" void interleave_host(uint64_t n, const uint64_t* w, uint32_t* even, uint32_t* odd, uint64_t* joined) {
  for (uint64_t i = 0; i < n; ++i) {
    even[i] = interleaved_half(w[i], 0);
    odd[i] = interleaved_half(w[i], 1);
    joined[i] = interleaved_join(even[i], odd[i]);
  }
}

// Goldilocks running sum (Accum<FieldP64>): op 0 adds a, 1 subtracts a, 2 adds the product a b.
extern "C" uint64_t gl_host_accum(uint64_t n, const int32_t* ops, const uint64_t* a, const uint64_t* b) {
  Accum<FieldP64> acc;
  acc.clear();
  for (uint64_t i = 0; i < n; ++i) {
    if (ops[i] == 0) acc.add(gl_elem(a[i]));
    else if (ops[i] == 1) acc.sub(gl_elem(a[i]));
    else acc.add_product(gl_elem(a[i]), gl_elem(b[i]));
  }
  return gl_u64(acc.value());
}

// The same operations on the Goldilocks working form's running sum (Work<FieldP64>::Acc:
// 64-bit column sums of 32-bit words, carried once in value()).
extern "C" uint64_t gl_host_work_accum(uint64_t n, const int32_t* ops, const uint64_t* a, const uint64_t* b) {
  Work<FieldP64>::Acc acc;
  acc.clear();
  for (uint64_t i = 0; i < n; ++i) {
    if (ops[i] == 0) acc.add(gl_elem(a[i]));
    else if (ops[i] == 1) acc.sub(gl_elem(a[i]));
    else acc.add_product(gl_elem(a[i]), gl_elem(b[i]));
  }
  return gl_u64(acc.value());
}

// Goldilocks working-form primitives on raw 64-bit words: op 0 gl_redc(gl_wide_mul(a, b)),
// op 1 gl_add_fast(a, b), op 2 gl_sub_fast(a, b), op 3 gl_canon_fast(a); op 4 the high word
// of gl_wide_mul(a, b) and op 5 its low word.
extern "C" void gl_host_fast(int op, uint64_t n, const uint64_t* a, const uint64_t* b, uint64_t* out) {
  for (uint64_t i = 0; i < n; ++i) {
    const GlWide x = gl_wide_mul(a[i], b[i]);
    if (op == 0) out[i] = gl_redc(x);
    else if (op == 1) out[i] = gl_add_fast(a[i], b[i]);
    else if (op == 2) out[i] = gl_sub_fast(a[i], b[i]);
    else if (op == 3) out[i] = gl_canon_fast(a[i]);
    else if (op == 4) out[i] = x.hi;
    else out[i] = (uint64_t)x.x0 | ((uint64_t)x.x1 << 32);
  }
}

// The input generator's uniform entries (prng.hpp) for indices start .. start + n - 1 of one
// stream key: L words each.
template <class F>
static void uniform_run(uint64_t key, uint64_t start, uint64_t n, uint32_t* out) {
  for (uint64_t i = 0; i < n; ++i) {
    const Elem<F::L> x = uniform_element<F>(key, start + i);
    for (int k = 0; k < F::L; ++k) out[i * F::L + k] = x.v[k];
  }
}

extern "C" void uniform_host(int bit_width, uint64_t key, uint64_t start, uint64_t n, uint32_t* out) {
  if (bit_width == 32) uniform_run<FieldP32>(key, start, n, out);
  else if (bit_width == 64) uniform_run<FieldP64>(key, start, n, out);
  else if (bit_width == 128) uniform_run<FieldP128>(key, start, n, out);
  else uniform_run<FieldP256>(key, start, n, out);
}
