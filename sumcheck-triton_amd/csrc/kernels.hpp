// HIP kernels of the SumCheck prover for one field policy F (instantiated once per field in
// inst_f*.hip). Reference semantics: cuda/src/field_sumcheck.cu:593-694 (fill, eval, reduce,
// fold) and :1012-1088 (round loop); src/zkduel/reference.py:78-108.
//
// One round j is one `pass_kernel` launch plus one single-block `finalize_kernel` launch (or,
// with fused finalize, the pass kernel's last block to finish does the finalize work):
//   pass_kernel     each thread owns pairs i of round j. For j >= 1 it reads entries
//                   4i..4i+3 of round j-1's tables, folds them with r_{j-1} into round j's
//                   entries 2i, 2i+1 (written back for round j+1), and evaluates the
//                   composition at the d+1 points t = 0..d of that pair, accumulating
//                   g_j(t) in registers. So each round streams the tables once (fused
//                   fold(j-1) + eval(j)); round 0 only reads. Block partial sums go to HBM.
//   finalize_kernel sums the block partials, stores g_j, and derives r_j on the device
//                   (SHA3 transcript, fixed r = 1, or an explicit list), so the host never
//                   waits between rounds and a whole proof can be captured in a hipGraph.
// Late rounds whose tables are small run in `tail_kernel`: one workgroup loops over the
// remaining rounds with the same per-pair code, reducing and hashing in-kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "field.hpp"
#include "keccak.hpp"
#include "layout.hpp"
#include "prng.hpp"
#include "work.hpp"

namespace sc {

// fn(integral_constant<int, I>) for I = 0..N-1, expanded at compile time. Used where a loop
// body is large enough that the compiler might keep the loop and index register arrays
// through scratch memory.
template <int N, int I = 0, class Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
  if constexpr (I < N) {
    fn(std::integral_constant<int, I>{});
    static_for<N, I + 1>(static_cast<Fn&&>(fn));
  }
}

// static_for when EXPAND, else an ordinary loop left to the compiler's unroller (for the
// interpreted kernels with up to 32 tables, whose full expansion is too large to compile).
template <int N, bool EXPAND, class Fn>
__device__ __forceinline__ void table_for(Fn&& fn) {
  if constexpr (EXPAND) {
    static_for<N>(static_cast<Fn&&>(fn));
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) fn(i);
  }
}

// ---------------------------------------------------------------------------------------------
// Vector memory helpers (element-major: NW consecutive words per thread)
// ---------------------------------------------------------------------------------------------
// The loads use clang's native vector types: with HIP's uint4 struct the backend kept a
// prefetched 16-byte load for its upper half and loaded the lower half a second time (8 bytes)
// where the kernel held the words across a loop iteration.
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));

template <int NW>
__device__ __forceinline__ void load_words(uint32_t (&w)[NW], const uint32_t* __restrict__ p) {
  if constexpr (NW % 4 == 0) {
#pragma unroll
    for (int i = 0; i < NW / 4; ++i) {
      const u32x4_t x = reinterpret_cast<const u32x4_t*>(p)[i];
      w[4 * i + 0] = x.x;
      w[4 * i + 1] = x.y;
      w[4 * i + 2] = x.z;
      w[4 * i + 3] = x.w;
    }
  } else {
    static_assert(NW == 2, "pairs of 1-limb entries");
    const u32x2_t x = *reinterpret_cast<const u32x2_t*>(p);
    w[0] = x.x;
    w[1] = x.y;
  }
}

template <int NW>
__device__ __forceinline__ void store_words(uint32_t* __restrict__ p, const uint32_t (&w)[NW]) {
  if constexpr (NW % 4 == 0) {
#pragma unroll
    for (int i = 0; i < NW / 4; ++i) {
      reinterpret_cast<uint4*>(p)[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
    }
  } else {
    static_assert(NW == 2, "pairs of 1-limb entries");
    *reinterpret_cast<uint2*>(p) = make_uint2(w[0], w[1]);
  }
}

template <int L>
__device__ __forceinline__ Elem<L> elem_at(const uint32_t* w, int k) {
  Elem<L> x;
#pragma unroll
  for (int i = 0; i < L; ++i) x.v[i] = w[k * L + i];
  return x;
}

template <int L>
__device__ __forceinline__ Elem<L> load_elem(const uint32_t* p) {
  Elem<L> x;
#pragma unroll
  for (int i = 0; i < L; ++i) x.v[i] = p[i];
  return x;
}

// ---------------------------------------------------------------------------------------------
// Table layouts. Entry e of table v of a buffer lies at entry v vstride + (e >> shift) qstride +
// (e mod 2^shift) of it. Inputs are table-major with a stride (shift 63: v stride + e). Fold
// buffers (the tables round j >= 1 writes and round j + 1 reads) keep each table in chunks of
// c = min(n, 128) entries with chunk q of every table side by side, table v's chunk q at
// (q T + v) c: the T tables' entries of a wave's pairs then form one run instead of T runs
// a table apart, and the folding rounds' 2:1 read/write stream moves faster that way (3-5%
// in a probe on MI355X). The chunked layout holds the same T n entries; the chunk size is a
// launch parameter (chunk_shift), table-major when it reaches the table size.
// ---------------------------------------------------------------------------------------------
struct TableMap {
  uint64_t vstride;
  uint64_t qstride;
  uint32_t shift;
  __device__ __forceinline__ uint64_t at(uint64_t v, uint64_t e) const {
    return v * vstride + (e >> shift) * (uint64_t)(uint32_t)qstride + (e & ((((uint64_t)1) << shift) - 1u));
  }
};

__device__ __forceinline__ TableMap strided_map(uint64_t stride) { return TableMap{stride, 0, 63u}; }

// n entries per table (a power of two), T tables, chunks of 2^cs entries, cs >= 2 (a chunk at
// least the table's size makes the buffer table-major; the 4 entries a folding thread reads and
// the 2 it writes are contiguous in any chunk of 4 or more)
__device__ __forceinline__ TableMap fold_map(uint64_t n, int T, int cs) {
  const uint32_t tz = (uint32_t)__builtin_ctzll(n);
  const uint32_t sh = tz < (uint32_t)cs ? tz : (uint32_t)cs;
  return TableMap{(uint64_t)1 << sh, (uint64_t)T << sh, sh};
}

// The map of round j's source (the input for j <= 1, else round j - 1's fold buffer of src_n
// entries per table).
__device__ __forceinline__ TableMap source_map(int j, uint64_t in_stride, uint64_t src_n, int T, int cs) {
  return j <= 1 ? strided_map(in_stride) : fold_map(src_n, T, cs);
}

// ---------------------------------------------------------------------------------------------
// Composition and per-pair evaluation
// ---------------------------------------------------------------------------------------------

// C(x) = sum_k c_k prod_v x_v^e_kv. Term structure is wave-uniform, so every branch below is
// a scalar branch; table slots are indexed statically (unrolled v), keeping x in VGPRs.
// Terms with fewer than min_len factors are left out (the point-parallel lane that evaluates
// the leading coefficient passes the degree).
template <class F, int TMAX, bool EXPAND = (TMAX <= 8), class A>
__device__ __forceinline__ void compose_add(A& acc, const typename Work<F>::V (&x)[TMAX],
                                            const CompDesc* __restrict__ cd, int K, int min_len = 0) {
  using W = Work<F>;
  for (int k = 0; k < K; ++k) {
    typename W::V prod = x[0];  // every term has a factor, so prod is set before it is used
    bool have = false;
    table_for<TMAX, EXPAND>([&](auto vi) {
      const int v = vi;
      uint32_t e = (cd->exps[k][v >> 3] >> ((v & 7) * 4)) & 15u;
      for (; e != 0u; --e) {
        if (have) {
          prod = W::mul(prod, x[v]);
        } else {
          prod = x[v];
          have = true;
        }
      }
    });
    const int kind = cd->kind[k];
    if (kind == COEFF_GENERAL) prod = W::mul(prod, W::coeff(cd->coeff[k]));
    if (cd->len[k] >= min_len) {
      if (kind == COEFF_MINUS_ONE) {
        acc.sub(prod);
      } else {
        acc.add(prod);
      }
    }
  }
}

// acc += a b. Goldilocks adds the raw 128-bit product to its wide accumulator (one reduction
// per thread for the whole sum instead of one per product).
template <class F, class A>
__device__ __forceinline__ void acc_add_mul(A& acc, const typename Work<F>::V& a,
                                            const typename Work<F>::V& b) {
  if constexpr (std::is_same_v<F, FieldP64>) {
    acc.add_product(a, b);
  } else {
    acc.add(Work<F>::mul(a, b));
  }
}

// Compiled composition STATIC_COMPS[SID]: every loop has constant bounds and indices, so it
// unrolls into straight-line field operations on fixed registers.
template <class F, int SID, int TMAX, class A>
__device__ __forceinline__ void compose_static(A& acc, const typename Work<F>::V (&x)[TMAX],
                                               const CompDesc* __restrict__ cd) {
  using W = Work<F>;
  static_assert(STATIC_COMPS[SID].num_tables <= TMAX, "table slots");
  constexpr StaticComp C2 = STATIC_COMPS[SID];
  if constexpr (C2.num_terms == 2 && C2.kinds[0] == COEFF_ONE && C2.kinds[1] == COEFF_ONE && C2.lens[0] >= 2 &&
                C2.lens[1] >= 2) {
    // Two products with unit coefficients: their last multiplications share one reduction.
    typename W::V pre[2];
    static_for<2>([&](auto ki) {
      constexpr int k = decltype(ki)::value;
      pre[k] = x[C2.vars[k][0]];
      static_for<C2.lens[k] - 2>([&](auto mi) {
        constexpr int m = decltype(mi)::value + 1;
        pre[k] = W::mul(pre[k], x[STATIC_COMPS[SID].vars[k][m]]);
      });
    });
    if constexpr (std::is_same_v<F, FieldP64>) {
      acc.add_product(pre[0], x[C2.vars[0][C2.lens[0] - 1]]);
      acc.add_product(pre[1], x[C2.vars[1][C2.lens[1] - 1]]);
    } else {
      acc.add(W::mul_sum(pre[0], x[C2.vars[0][C2.lens[0] - 1]], pre[1], x[C2.vars[1][C2.lens[1] - 1]]));
    }
    return;
  }
  static_for<STATIC_COMPS[SID].num_terms>([&](auto ki) {
    constexpr int k = decltype(ki)::value;
    constexpr StaticComp C = STATIC_COMPS[SID];
    typename W::V prod = x[C.vars[k][0]];
    // a unit-coefficient product keeps its last multiplication for the accumulator
    constexpr int kept = (C.kinds[k] == COEFF_ONE && C.lens[k] >= 2) ? 1 : 0;
    static_for<C.lens[k] - 1 - kept>([&](auto mi) {
      constexpr int m = decltype(mi)::value + 1;
      prod = W::mul(prod, x[STATIC_COMPS[SID].vars[k][m]]);
    });
    if constexpr (kept == 1) {
      acc_add_mul<F>(acc, prod, x[C.vars[k][C.lens[k] - 1]]);
    } else if constexpr (C.kinds[k] == COEFF_ONE) {
      acc.add(prod);
    } else if constexpr (C.kinds[k] == COEFF_MINUS_ONE) {
      acc.sub(prod);
    } else {
      acc_add_mul<F>(acc, prod, W::coeff(cd->coeff[k]));
    }
  });
}

template <class F, int SID, int TMAX, class A>
__device__ __forceinline__ void compose(A& acc, const typename Work<F>::V (&x)[TMAX],
                                        const CompDesc* __restrict__ cd, int K) {
  if constexpr (SID == 0) {
    compose_add<F, TMAX>(acc, x, cd, K);
  } else {
    compose_static<F, SID, TMAX>(acc, x, cd);
  }
}

// The leading coefficient of a pair's polynomial C(lo + t (hi - lo)) (degree d): the terms of
// d factors, each factor replaced by its slope s_v = hi_v - lo_v. acc += that coefficient.
template <class F, int TMAX, bool EXPAND = (TMAX <= 8), class A>
__device__ __forceinline__ void compose_add_lead(A& acc, const typename Work<F>::V (&s)[TMAX],
                                                 const CompDesc* __restrict__ cd, int K, int d) {
  using W = Work<F>;
  for (int k = 0; k < K; ++k) {
    if (cd->len[k] != d) continue;  // wave-uniform
    typename W::V prod = s[0];
    bool have = false;
    table_for<TMAX, EXPAND>([&](auto vi) {
      const int v = vi;
      uint32_t e = (cd->exps[k][v >> 3] >> ((v & 7) * 4)) & 15u;
      for (; e != 0u; --e) {
        if (have) {
          prod = W::mul(prod, s[v]);
        } else {
          prod = s[v];
          have = true;
        }
      }
    });
    const int kind = cd->kind[k];
    if (kind == COEFF_ONE) {
      acc.add(prod);
    } else if (kind == COEFF_MINUS_ONE) {
      acc.sub(prod);
    } else {
      acc.add(W::mul(prod, W::coeff(cd->coeff[k])));
    }
  }
}

template <class F, int SID, int TMAX, class A>
__device__ __forceinline__ void compose_static_lead(A& acc, const typename Work<F>::V (&s)[TMAX],
                                                    const CompDesc* __restrict__ cd) {
  using W = Work<F>;
  constexpr int D = static_degree(SID);
  static_for<STATIC_COMPS[SID].num_terms>([&](auto ki) {
    constexpr int k = decltype(ki)::value;
    constexpr StaticComp C = STATIC_COMPS[SID];
    if constexpr (C.lens[k] == D) {
      typename W::V prod = s[C.vars[k][0]];
      constexpr int kept = C.kinds[k] == COEFF_ONE ? 1 : 0;
      static_for<C.lens[k] - 1 - kept>([&](auto mi) {
        constexpr int m = decltype(mi)::value + 1;
        prod = W::mul(prod, s[STATIC_COMPS[SID].vars[k][m]]);
      });
      if constexpr (kept == 1) {
        acc_add_mul<F>(acc, prod, s[C.vars[k][C.lens[k] - 1]]);
      } else if constexpr (C.kinds[k] == COEFF_MINUS_ONE) {
        acc.sub(prod);
      } else {
        acc_add_mul<F>(acc, prod, W::coeff(cd->coeff[k]));
      }
    }
  });
}

template <class F, int SID, int TMAX, class A>
__device__ __forceinline__ void compose_lead(A& acc, const typename Work<F>::V (&s)[TMAX],
                                             const CompDesc* __restrict__ cd, int K, int d) {
  if constexpr (SID == 0) {
    compose_add_lead<F, TMAX>(acc, s, cd, K, d);
  } else {
    compose_static_lead<F, SID, TMAX>(acc, s, cd);
  }
}

// Shape of the composition a kernel runs: compile-time for compiled ones.
template <int SID>
struct Shape {
  __device__ static int tables(const CompDesc* cd) {
    if constexpr (SID == 0) return cd->num_tables; else return STATIC_COMPS[SID].num_tables;
  }
  __device__ static int terms(const CompDesc* cd) {
    if constexpr (SID == 0) return cd->num_terms; else return STATIC_COMPS[SID].num_terms;
  }
  __device__ static int npts(const CompDesc* cd) {
    if constexpr (SID == 0) return cd->npts; else return static_degree(SID) + 1;
  }
};

// Compiled compositions whose first term is a*b*c with a unit coefficient and whose degree is
// 3 (a*b*c, a*b*c + d*e), at 255 bits: round 0 (points 0, 1, 2 and the leading coefficient)
// by eval_round0_abc.
template <class F, int SID>
constexpr bool round0_abc_form() {
  if constexpr (!std::is_same_v<F, FieldP256> || SID == 0) {
    return false;
  } else {
    constexpr StaticComp C = STATIC_COMPS[SID];
    return static_degree(SID) == 3 && C.lens[0] == 3 && C.kinds[0] == COEFF_ONE &&
           (C.num_terms == 1 || (C.num_terms == 2 && C.kinds[1] == COEFF_ONE && C.lens[1] == 2));
  }
}

// Round 0 of round0_abc_form compositions. a(t) b(t) is quadratic in t, so
//   a(2) b(2) = 2 a(1) b(1) - a(0) b(0) + 2 (a(1) - a(0)) (b(1) - b(0)),
// and that last product is the one the leading coefficient (a b c over the differences) needs
// anyway: per pair one product and the interpolation of a and b at t = 2 fewer than the
// general order (round 0 issues VALU at the mad probe's rate, so instructions are its time).
template <class F, int TMAX, int PMAX, int SID, class A>
__device__ __forceinline__ void eval_round0_abc(typename Work<F>::V (&cur)[TMAX], typename Work<F>::V (&oth)[TMAX],
                                                A (&acc)[PMAX]) {
  using W = Work<F>;
  using V = typename W::V;
  constexpr StaticComp C = STATIC_COMPS[SID];
  constexpr int ia = C.vars[0][0], ib = C.vars[0][1], ic = C.vars[0][2];
  static_assert(PMAX >= 4, "points 0, 1, 2 and the leading coefficient");
  auto finish = [&](A& ac, const V& ab) {  // ac += ab c (+ d e)
    if constexpr (C.num_terms == 2) {
      ac.add(W::mul_sum(ab, cur[ic], cur[C.vars[1][0]], cur[C.vars[1][1]]));
    } else {
      acc_add_mul<F>(ac, ab, cur[ic]);
    }
  };
  const V ab0 = W::mul(cur[ia], cur[ib]);
  finish(acc[0], ab0);
  static_for<C.num_tables>([&](auto vi) {
    constexpr int v = decltype(vi)::value;
    const V o = oth[v];
    oth[v] = W::diff(o, cur[v]);
    cur[v] = o;
  });
  const V ab1 = W::mul(cur[ia], cur[ib]);
  finish(acc[1], ab1);
  const V dadb = W::mul(oth[ia], oth[ib]);
  acc_add_mul<F>(acc[3], dadb, oth[ic]);
  const V ab2 = r29_two_x_plus_two_y_minus_z(ab1, dadb, ab0);
  static_for<C.num_tables>([&](auto vi) {
    constexpr int v = decltype(vi)::value;
    if constexpr (v != ia && v != ib) {
      if constexpr (static_lazy_step(SID, v, 2)) {
        cur[v] = W::step_lazy(cur[v], oth[v]);
      } else {
        cur[v] = W::step(cur[v], oth[v]);
      }
    }
  });
  finish(acc[2], ab2);
}

// cur holds f[2i] and oth holds f[2i+1] of every table on entry. Accumulates
// C(f[2i] + t (f[2i+1] - f[2i])) into acc[t] for t < d = npts - 1, with t >= 2 reached by
// repeated addition of the difference (reference.py:89-99; those sums may stay unreduced
// where the field allows, step_add), and into acc[d] either the value at t = d (LEAD false)
// or, for d >= 2, the polynomial's leading coefficient (its value at infinity: the degree-d
// terms over the differences), from which the round's finish recovers g(d).
template <class F, int TMAX, int PMAX, int SID, bool LEAD = true, class A>
__device__ __forceinline__ void eval_points(typename Work<F>::V (&cur)[TMAX], typename Work<F>::V (&oth)[TMAX],
                                            A (&acc)[PMAX], const CompDesc* __restrict__ cd,
                                            int T, int K, int npts, bool skip1) {
  using W = Work<F>;
  if constexpr (LEAD && round0_abc_form<F, SID>()) {
    if (!skip1) {
      eval_round0_abc<F, TMAX, PMAX, SID>(cur, oth, acc);
      return;
    }
  }
  constexpr bool X = SID != 0 || TMAX <= 8;  // expand every loop unless the table count is large
  const int d = LEAD ? npts - 1 : npts;  // the points t < d are evaluated
  compose<F, SID, TMAX>(acc[0], cur, cd, K);
  table_for<TMAX, X>([&](auto vi) {
    const int v = vi;
    if (v < T) {
      const typename W::V o = oth[v];
      oth[v] = W::diff(o, cur[v]);
      cur[v] = o;
    }
  });
  if (!skip1) compose<F, SID, TMAX>(acc[1], cur, cd, K);  // else finish_round derives g(1)
  table_for<PMAX - 2, X>([&](auto ti) {
    const int t = (int)ti + 2;
    if (t == d) compose_lead<F, SID, TMAX>(acc[t], oth, cd, K, d);
    if (t < d) {
      table_for<TMAX, X>([&](auto vi) {
        const int v = vi;
        if constexpr (SID != 0 && W::kLazyStep) {
          if constexpr (static_lazy_step(SID, vi, (int)ti + 2)) {
            cur[v] = W::step_lazy(cur[v], oth[v]);
          } else {
            cur[v] = W::step(cur[v], oth[v]);
          }
        } else {
          if (v < T) cur[v] = W::step(cur[v], oth[v]);
        }
      });
      compose<F, SID, TMAX>(acc[t], cur, cd, K);
    }
  });
}

// Fields that fold by a table of the round's challenge (Work::kFoldTable: the 255-bit one)
// build it per block in LDS: row k = r 2^(29k + S) mod p, S = 5 for folds straight from the
// caller's Montgomery input. fold_table_words is the LDS size to declare (a placeholder for the
// other fields). build_fold_table is called by the whole block and ends with a barrier.
template <class F>
constexpr int fold_table_words() {
  if constexpr (Work<F>::kFoldTable) {
    return Work<F>::kFoldTableWords;
  } else {
    return 4;
  }
}

template <class F>
__device__ __forceinline__ void build_fold_table(uint32_t* s_tab, const uint32_t* rm, bool src_mont) {
  using W = Work<F>;
  if constexpr (W::kFoldTable) {
    if (threadIdx.x < 9) {
      if (src_mont) {
        W::template fold_row<5>(s_tab, rm, (int)threadIdx.x);
      } else {
        W::template fold_row<0>(s_tab, rm, (int)threadIdx.x);
      }
    }
    __syncthreads();
  }
}

// Loads pair i of the current round for every table, folding from the previous round's
// tables when fold != 0 (f' = e + r (o - e), field_sumcheck.cu:677-694), and writes the
// folded entries to dst. src holds the caller's Montgomery tables when src_mont (rounds 0
// and 1), else entries written by an earlier fold. sm and dm place the entries of src and dst.
// With USE_TAB (fields with Work::kFoldTable) the folds use the block's fold table tab
// (build_fold_table) instead of r.
template <class F, int TMAX, bool USE_TAB = false>
__device__ __forceinline__ void load_pair(typename Work<F>::V (&cur)[TMAX], typename Work<F>::V (&oth)[TMAX],
                                          const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                          const TableMap& sm, const TableMap& dm, uint64_t i, int T, bool fold,
                                          bool src_mont, const typename Work<F>::V& r,
                                          const uint32_t* tab = nullptr) {
  using W = Work<F>;
  constexpr int L = F::L;
  table_for<TMAX, (TMAX <= 8)>([&](auto vi) {
    const int v = vi;
    if (v < T) {
      if (fold) {
        uint32_t w[4 * L];
        load_words<4 * L>(w, src + sm.at(v, 4 * i) * L);
        typename W::V e, o;
        if constexpr (USE_TAB && W::kFoldTable) {
          if (src_mont) {
            W::template fold2_tab<true>(w, tab, e, o);
          } else {
            W::template fold2_tab<false>(w, tab, e, o);
          }
        } else {
          typename W::V f[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) f[k] = src_mont ? W::from_mont(w + k * L) : W::from_stored(w + k * L);
          e = W::fold(f[0], f[1], r);
          o = W::fold(f[2], f[3], r);
        }
        uint32_t y[2 * L];
        W::store(y, e);
        W::store(y + L, o);
        store_words<2 * L>(dst + dm.at(v, 2 * i) * L, y);
        cur[v] = e;
        oth[v] = o;
      } else {
        uint32_t w[2 * L];
        load_words<2 * L>(w, src + sm.at(v, 2 * i) * L);
        cur[v] = W::from_mont(w);
        oth[v] = W::from_mont(w + L);
      }
    }
  });
}

// load_pair with the round's kind fixed at compile time (FOLD: round >= 1; SRC_MONT: src is
// the caller's Montgomery input), every table's words loaded before any is used. With the
// kind a run-time value (load_pair) the compiler keeps one branch per table and issues each
// table's loads after the previous table was folded and stored: one table's loads in flight
// per thread at a time, but reads and writes interleaved.
template <class F, int TMAX, bool FOLD, bool SRC_MONT>
__device__ __forceinline__ void load_pair_fixed(typename Work<F>::V (&cur)[TMAX], typename Work<F>::V (&oth)[TMAX],
                                                const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                const TableMap& sm, const TableMap& dm, uint64_t i, int T,
                                                const typename Work<F>::V& r) {
  using W = Work<F>;
  constexpr int L = F::L;
  constexpr int NW = FOLD ? 4 * L : 2 * L;
  uint32_t w[TMAX][NW];
  table_for<TMAX, (TMAX <= 8)>([&](auto vi) {
    const int v = vi;
    if (v < T) load_words<NW>(w[v], src + sm.at(v, (FOLD ? 4 : 2) * i) * L);
  });
  table_for<TMAX, (TMAX <= 8)>([&](auto vi) {
    const int v = vi;
    if (v < T) {
      if constexpr (FOLD) {
        typename W::V f[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) f[k] = SRC_MONT ? W::from_mont(w[v] + k * L) : W::from_stored(w[v] + k * L);
        const typename W::V e = W::fold(f[0], f[1], r);
        const typename W::V o = W::fold(f[2], f[3], r);
        uint32_t y[2 * L];
        W::store(y, e);
        W::store(y + L, o);
        store_words<2 * L>(dst + dm.at(v, 2 * i) * L, y);
        cur[v] = e;
        oth[v] = o;
      } else {
        cur[v] = W::from_mont(w[v]);
        oth[v] = W::from_mont(w[v] + L);
      }
    }
  });
}

// The 255-bit field's per-pair loads: table v + 1's words are requested before table v is
// converted and folded (two tables' words in flight per thread; all of them would not fit the
// registers). With PRE the caller has already requested table 0 (the folding rounds of compiled
// compositions request the next pair's table 0 before this pair's evaluation, with their
// running sums moved to LDS to make room; with the sums in registers that prefetch took the
// kernel to 254 VGPRs and measured 4-10% slower). FOLD / SRC_MONT fix the round kind at compile
// time: round 0 (SRC_MONT, no fold), round 1 (folding the caller's Montgomery input) or a
// later round (folding stored entries). Round 0 of a compiled composition reads table v as
// from_mont_scaled unless it is one of the factors read exactly (static_round0_exact).
// r32 = Work::times32(r) for round 1. w is the caller's word buffer: with PRE, w[0] holds table
// 0's words of pair i on entry.
template <class F, int TMAX, bool FOLD, bool SRC_MONT, int SID, bool PRE = false>
__device__ __forceinline__ void load_pair_pipelined(typename Work<F>::V (&cur)[TMAX], typename Work<F>::V (&oth)[TMAX],
                                                    const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                    const TableMap& sm, const TableMap& dm, uint64_t i, int T,
                                                    const typename Work<F>::V& r, const typename Work<F>::V& r32,
                                                    const uint32_t* tab, uint32_t (&w)[2][FOLD ? 4 * F::L : 2 * F::L]) {
  using W = Work<F>;
  constexpr int L = F::L;
  constexpr int NW = FOLD ? 4 * L : 2 * L;
  if constexpr (!PRE) load_words<NW>(w[0], src + sm.at(0, (FOLD ? 4 : 2) * i) * L);
  table_for<TMAX, (TMAX <= 8)>([&](auto vi) {
    const int v = vi;
    if (v < T) {
      if (v + 1 < T && v + 1 < TMAX)
        load_words<NW>(w[(v + 1) & 1], src + sm.at(v + 1, (FOLD ? 4 : 2) * i) * L);
      const uint32_t* wv = w[v & 1];
      if constexpr (FOLD) {
        typename W::V e, o;
        if constexpr (W::kFoldTable) {
          W::template fold2_tab<SRC_MONT>(wv, tab, e, o);
        } else if constexpr (SRC_MONT && W::kCheapMontInput) {
          e = W::fold_mont(wv, wv + L, r32);
          o = W::fold_mont(wv + 2 * L, wv + 3 * L, r32);
        } else {
          typename W::V f[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) f[k] = SRC_MONT ? W::from_mont(wv + k * L) : W::from_stored(wv + k * L);
          e = W::fold(f[0], f[1], r);
          o = W::fold(f[2], f[3], r);
        }
        uint32_t y[2 * L];
        W::store(y, e);
        W::store(y + L, o);
        store_words<2 * L>(dst + dm.at(v, 2 * i) * L, y);
        cur[v] = e;
        oth[v] = o;
      } else if constexpr (SID != 0 && W::kCheapMontInput) {
        if (static_round0_exact(SID, v)) {
          cur[v] = W::from_mont(wv);
          oth[v] = W::from_mont(wv + L);
        } else {
          cur[v] = W::from_mont_scaled(wv);
          oth[v] = W::from_mont_scaled(wv + L);
        }
      } else {
        cur[v] = W::from_mont(wv);
        oth[v] = W::from_mont(wv + L);
      }
    }
  });
}

// ---------------------------------------------------------------------------------------------
// Reductions (exact modular sums, so any order gives the same bits)
// ---------------------------------------------------------------------------------------------

// Cross-lane moves as DPP operand modifiers (register to register, no LDS round trip). CTRL
// is the gfx9 DPP control: quad_perm, row_half_mirror (lane i <- 7 - i within 8),
// row_mirror (i <- 15 - i within 16), row_bcast15 / row_bcast31 (lane 15 of a row / lane 31
// to the following rows selected by ROW_MASK). Lanes in rows outside ROW_MASK read 0.
constexpr int DPP_QUAD_1032 = 0xB1;  // quad_perm [1, 0, 3, 2]
constexpr int DPP_QUAD_2301 = 0x4E;  // quad_perm [2, 3, 0, 1]
constexpr int DPP_ROW_HALF_MIRROR = 0x141;
constexpr int DPP_ROW_MIRROR = 0x140;
constexpr int DPP_ROW_BCAST15 = 0x142;
constexpr int DPP_ROW_BCAST31 = 0x143;

template <int CTRL, int ROW_MASK = 0xF, int L>
__device__ __forceinline__ Elem<L> dpp_move(const Elem<L>& x) {
  Elem<L> y;
#pragma unroll
  for (int k = 0; k < L; ++k) y.v[k] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x.v[k], CTRL, ROW_MASK, 0xF, false);
  return y;
}

// Sum over each group of 8 lanes (8k .. 8k+7), returned to all lanes of the group.
template <class F>
__device__ __forceinline__ Elem<F::L> group8_sum(Elem<F::L> x) {
  x = add<F>(x, dpp_move<DPP_QUAD_1032>(x));
  x = add<F>(x, dpp_move<DPP_QUAD_2301>(x));
  return add<F>(x, dpp_move<DPP_ROW_HALF_MIRROR>(x));
}

// Sum over the whole wave, valid in lane 63: group sums, row sums (row_mirror), then row r
// adds row r-1 (bcast15 into rows 1 and 3) and rows 2, 3 add rows 0 + 1 (bcast31).
template <class F>
__device__ __forceinline__ Elem<F::L> wave_sum_to_lane63(Elem<F::L> x) {
  x = group8_sum<F>(x);
  x = add<F>(x, dpp_move<DPP_ROW_MIRROR>(x));
  x = add<F>(x, dpp_move<DPP_ROW_BCAST15, 0xA>(x));
  return add<F>(x, dpp_move<DPP_ROW_BCAST31, 0xC>(x));
}

// Sums acc over the NT-thread block into out[t*L + limb] (LDS) for t < npts. red needs
// (NT/64)*PMAX*L words. Ends with a barrier.
template <class F, int PMAX, int NT>
__device__ __forceinline__ void block_reduce(Elem<F::L> (&acc)[PMAX], int npts, uint32_t* red,
                                             uint32_t* out) {
  constexpr int L = F::L;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < PMAX; ++t) {
    if (t < npts) {
      acc[t] = wave_sum_to_lane63<F>(acc[t]);
      if (lane == 63) {
#pragma unroll
        for (int k = 0; k < L; ++k) red[(wave * PMAX + t) * L + k] = acc[t].v[k];
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < npts) {
    const int t = threadIdx.x;
    Elem<L> s = load_elem<L>(red + t * L);
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) s = add<F>(s, load_elem<L>(red + (w * PMAX + t) * L));
#pragma unroll
    for (int k = 0; k < L; ++k) out[t * L + k] = s.v[k];
  }
  __syncthreads();
}

// Sum over the NT threads of the block of each thread's PMAX canonical values (PMAX <= 4),
// into out[t * L ..] for t < npts, as 64-bit word columns: each wave writes its lanes' words to
// LDS transposed (one padded row of 64 words per (point, word) column), one lane per column
// adds its row, the waves' column sums are added, and one thread per point reduces its L
// columns once (reduce_word_columns). At 256 bits the element-wise reduction chained six
// modular adds per point through the wave (5.9 us of a 30 us tail round).
template <int NT, int C>
constexpr int cols_lds_words() {
  return (NT / 64) * C * 65;
}

template <class F, int PMAX, int NT>
__device__ __forceinline__ void block_reduce_cols(const Elem<F::L> (&vals)[PMAX], int npts, uint32_t* s_tr,
                                                  uint64_t* s_c, uint32_t* out) {
  constexpr int L = F::L;
  constexpr int C = PMAX * L;
  static_assert(C <= 64, "one lane per column");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t* rows = s_tr + wave * C * 65;
#pragma unroll
  for (int t = 0; t < PMAX; ++t) {
#pragma unroll
    for (int k = 0; k < L; ++k) rows[(t * L + k) * 65 + lane] = vals[t].v[k];
  }
  __syncthreads();
  if (lane < C) {
    uint64_t s = 0;
#pragma unroll 16
    for (int i = 0; i < 64; ++i) s += rows[lane * 65 + i];
    s_c[wave * C + lane] = s;
  }
  __syncthreads();
  if ((int)threadIdx.x < C) {
    uint64_t tot = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) tot += s_c[w * C + threadIdx.x];
    s_c[threadIdx.x] = tot;  // each thread rewrites its own column's first slot only
  }
  __syncthreads();
  if ((int)threadIdx.x < npts) {
    uint64_t col[L];
#pragma unroll
    for (int w = 0; w < L; ++w) col[w] = s_c[threadIdx.x * L + w];
    const Elem<L> g = reduce_word_columns<F>(col);
#pragma unroll
    for (int w = 0; w < L; ++w) out[threadIdx.x * L + w] = g.v[w];
  }
  __syncthreads();
}

// SHA3-256 of a padded message held in LDS (msg[0 .. 34 * nblocks)), run by one whole wave
// (lanes h * 32 + k hold half h of state word k, interleaved form), exchanging state halves
// through msg[3 * 34 .. MSG_WORDS). Returns digest word (lane mod 8) to every lane.
__device__ __forceinline__ uint32_t sha3_wave_word(uint32_t* msg, int nblocks) {
  const int lane = threadIdx.x & 63;
  const int h = keccak_half(), p = keccak_word();
  uint32_t* xs = msg + 3 * 34;
  uint32_t a = 0;
  for (int b = 0; b < nblocks; ++b) {
    if (p < 17) a ^= interleaved_half((uint64_t)msg[34 * b + 2 * p] | ((uint64_t)msg[34 * b + 2 * p + 1] << 32), h);
    keccak_f1600_wave(a, xs);
  }
  const int i = (lane & 7) >> 1;  // digest words 2i, 2i + 1 are state word i
  const uint32_t e = (uint32_t)__shfl((int)a, i, 64), o = (uint32_t)__shfl((int)a, 32 + i, 64);
  const uint64_t w = interleaved_join(e, o);
  return (lane & 1) ? (uint32_t)(w >> 32) : (uint32_t)w;
}

// Publishes g (LDS, npts*L words) as round `round`'s polynomial and, unless it is the last
// round (or last_challenge asks for the final-point challenge too), derives r_round into
// r_out (LDS, L words) and st->r[round]. All threads call it.
// The claim of round `round` >= 1: g_{round-1}(r_{round-1}) by barycentric interpolation
// from the previous polynomial's values at 0..npts-1, one point per lane of the calling wave,
// in the field's working form. Returns the canonical Montgomery value to lanes 0..7.
template <class F>
__device__ __forceinline__ Elem<F::L> previous_claim(const ProofState* st, int round, int npts,
                                                     const FieldConsts* __restrict__ fc) {
  using W = Work<F>;
  constexpr int L = F::L;
  const int t = threadIdx.x & 63;
  Elem<L> term = zero<F>();
  if (t < npts) {
    const typename W::V r = W::from_mont(st->r[round - 1]);
    typename W::V prod = W::mul(W::from_mont(st->polys[round - 1][t]), W::from_mont(fc->bary[npts - 1][t]));
    for (int s = 0; s < npts; ++s) {
      if (s != t) prod = W::mul(prod, W::diff(r, W::from_mont(fc->small[s])));
    }
    typename W::Acc acc;
    acc.clear();
    acc.add(prod);
    term = acc.value();
  }
  static_assert(MAX_POINTS == 8, "one point per lane of an 8-lane group");
  return group8_sum<F>(term);
}

// g(d) of a degree-d >= 2 polynomial from g(0..d-1) and its leading coefficient in g[d]
// (FieldConsts::lead), one term per lane of the calling wave; the value goes to lanes 0..7.
// With use_g1, lane 1 takes g1 (its own register) instead of g[1].
template <class F>
__device__ __forceinline__ Elem<F::L> value_from_lead(const uint32_t* g, int npts, const FieldConsts* __restrict__ fc,
                                                      bool use_g1, const Elem<F::L>& g1) {
  using W = Work<F>;
  constexpr int L = F::L;
  const int t = threadIdx.x & 63;
  const int d = npts - 1;
  Elem<L> term = zero<F>();
  if (t <= d) {
    const Elem<L> gt = (use_g1 && t == 1) ? g1 : load_elem<L>(g + t * L);
    typename W::Acc acc;
    acc.clear();
    acc.add(W::mul(W::from_mont(gt.v), W::from_mont(fc->lead[d][t])));
    term = acc.value();
  }
  return group8_sum<F>(term);
}

// With from_sums (every prover path), g holds the point sums of the round's kernels: from
// round 1 on g(1) is missing (the per-pair and point-parallel kernels skip point 1) and is
// set to claim - g(0), the protocol's own identity g_j(0) + g_j(1) = g_{j-1}(r_{j-1}). With
// lead as well (the pass, point-parallel and batched stage kernels; not the tail kernel), for
// a degree d >= 2 g[d] holds the sum of the leading coefficients (the kernels evaluate the
// pairs' polynomials at infinity instead of at t = d), from which g(d) follows by the finite
// difference identity. Both values are the ones evaluation at the point would give.
// claim_in (LDS, L words), when given, is the round's claim already computed (by a wave of the
// caller while the others summed the partials); else the first wave computes it here.
template <class F, int NT>
__device__ __forceinline__ void finish_round(uint32_t* g, int npts, int round, int num_rounds, int mode,
                             const FieldConsts* __restrict__ fc, ProofState* st,
                             const uint32_t* __restrict__ explicit_r, uint32_t* msg, uint32_t* r_out,
                             bool last_challenge = false, bool from_sums = false, bool lead = false,
                             const uint32_t* claim_in = nullptr) {
  constexpr int L = F::L;
  const int tid = threadIdx.x;
  const bool fill1 = from_sums && round > 0, filld = from_sums && lead && npts >= 3;
  if (fill1 || filld) {
    // one wave: g(1) = claim - g(0) is kept in lanes 0..7 for the g(d) terms and written by lane 0
    if (tid < 64) {
      Elem<L> g1 = zero<F>();
      if (fill1) {
        const Elem<L> claim = claim_in != nullptr ? load_elem<L>(claim_in) : previous_claim<F>(st, round, npts, fc);
        g1 = sub<F>(claim, load_elem<L>(g));
      }
      Elem<L> gd = zero<F>();
      if (filld) gd = value_from_lead<F>(g, npts, fc, fill1, g1);
      if (tid == 0) {
#pragma unroll
        for (int k = 0; k < L; ++k) {
          if (fill1) g[L + k] = g1.v[k];
          if (filld) g[(npts - 1) * L + k] = gd.v[k];
        }
      }
    }
    __syncthreads();
  }
  if (tid < npts * L) st->polys[round][tid / L][tid % L] = g[tid];
  if (round + 1 >= num_rounds && !last_challenge) return;
  if (mode == CHALLENGE_SHA3) {
    // Digest of the transcript, read as a little-endian integer, reduced mod p and
    // Montgomery-encoded (field_sumcheck.cu:857-907): sum_i digest_i * 2^(32i) * R mod p.
    const int nvals = npts * L;
    const int nb = transcript_blocks(nvals);
    for (int k = tid; k < 34 * nb; k += NT) msg[k] = transcript_word(k, (uint32_t)round, g, nvals, nb);
    __syncthreads();
    if (tid < 64) {
      // lane i < 8 multiplies digest word i by its power, then the eight terms are summed
      const uint32_t word = sha3_wave_word(msg, nb);
      Elem<L> r = zero<F>();
      if (tid < 8) {
        Elem<L> w = zero<F>();
        w.v[0] = word;
        r = mont_mul<F>(w, load_elem<L>(fc->word_pow[tid]));
      }
      r = group8_sum<F>(r);
      if (tid == 0) {
#pragma unroll
        for (int k = 0; k < L; ++k) r_out[k] = r.v[k];
      }
    }
  } else if (mode == CHALLENGE_FIXED) {
    if (tid < L) r_out[tid] = fc->one[tid];  // r = 1 (field_sumcheck.cu:1058)
  } else {
    if (tid == 0) {
      const Elem<L> x = load_elem<L>(explicit_r + (uint64_t)round * L);
      const Elem<L> r = mont_mul<F>(x, load_elem<L>(fc->r2));
#pragma unroll
      for (int k = 0; k < L; ++k) r_out[k] = r.v[k];
    }
  }
  __syncthreads();
  if (tid < L) st->r[round][tid] = r_out[tid];
}

// ---------------------------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------------------------

// The round's claim g_{j-1}(r_{j-1}) into st->claim, for round j >= 1 of a plain proof. It
// needs only the previous round's polynomial and challenge, so the round's pass (or
// point-parallel evaluation) kernel computes it, off the path of the finalize, which reads it
// (a later launch on the same stream): wave 0 of one extra block at the end of the grid
// (claim_block) that evaluates nothing, so no evaluating block is held up by it (as block 0
// computing it first would hold up a kernel of equal blocks by its length). The kernels branch
// to it first and the claim block never rejoins the evaluating blocks' path: a store to the
// proof state on (or merging into) that path keeps the compiler from reading the composition
// descriptor through the scalar cache (the constraint workload's evaluation then issued 50%
// more vector loads).
template <class F>
__device__ __forceinline__ void store_claim(ProofState* st, int round, int npts, const FieldConsts* __restrict__ fc) {
  if (round > 0 && threadIdx.x < 64) {
    const Elem<F::L> c = previous_claim<F>(st, round, npts, fc);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < F::L; ++k) st->claim[k] = c.v[k];
    }
  }
}

// Blocks of a pass or point-parallel evaluation grid that evaluate (all but the claim block).
__device__ __forceinline__ uint32_t evaluating_blocks(int claim_block) {
  return claim_block >= 0 ? (uint32_t)claim_block : gridDim.x;
}

// Sums over the blocks of a pass's partials (canonical values): each partial row (the first 4 or
// 8 point slots, L words each, of one block) is read by RW / 4 threads, 16 bytes each, so the
// block's threads cover S = 4 BLOCK / RW rows at a time with every load of a thread in flight together;
// each thread adds its four words across its rows in 64-bit columns (no carries; point slots
// past npts hold stale words and are summed but never read). The per-thread column sums are
// added in LDS, four per thread and then one column per thread, and one thread per point
// reduces its L column sums once (reduce_word_columns). At 256 bits the element-wise form
// took the finalize's partial sums and block reduction 5.8-14 us (each add is a carry chain and
// a conditional subtraction, and the wave reduction chained six of them per point).
template <class F, int PMAX>
__device__ __forceinline__ void sum_partials_cols(const uint32_t* __restrict__ partials, uint32_t nblocks, int npts,
                                                  uint64_t* s_cols, uint32_t* s_g) {
  constexpr int L = F::L;
  constexpr int RW = (PMAX <= 4 ? 4 : 8) * L;  // words of a row read (point slots 0..3 or 0..7)
  constexpr int TPR = RW / 4;                  // threads per row
  constexpr int S = BLOCK / TPR;
  static_assert(RW % 4 == 0 && BLOCK % TPR == 0, "16-byte loads over whole rows");
  const int q = (int)threadIdx.x % TPR, seg = (int)threadIdx.x / TPR;
  uint64_t sum[4] = {0, 0, 0, 0};
  // 8 rows per thread per batch (16 measured slower: the Goldilocks pass kernel, which inlines
  // this for its last-block finish, went from 95 registers and 5 waves per SIMD to 142 and 3;
  // 2.27 -> 2.47 ms per proof, same box)
  for (uint32_t b0 = (uint32_t)seg; b0 < nblocks; b0 += 8u * S) {
    u32x4_t x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t b = b0 + (uint32_t)u * S;
      x[u] = b < nblocks ? reinterpret_cast<const u32x4_t*>(partials + (uint64_t)b * MAX_POINTS * L)[q]
                         : u32x4_t{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      sum[0] += x[u].x;
      sum[1] += x[u].y;
      sum[2] += x[u].z;
      sum[3] += x[u].w;
    }
  }
  // S segments of RW columns
#pragma unroll
  for (int w = 0; w < 4; ++w) s_cols[seg * RW + 4 * q + w] = sum[w];
  __syncthreads();
  // every thread adds four segments of one column (S * RW = 4 BLOCK values in all)
  constexpr int P2 = BLOCK / RW;  // partial sums per column after this step
  const int c = (int)threadIdx.x % RW, part = (int)threadIdx.x / RW;
  uint64_t tot = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) tot += s_cols[(part * 4 + k) * RW + c];
  __syncthreads();
  s_cols[part * RW + c] = tot;
  __syncthreads();
  if ((int)threadIdx.x < RW) {
    uint64_t t2 = 0;
    for (int k = 0; k < P2; ++k) t2 += s_cols[k * RW + threadIdx.x];
    s_cols[threadIdx.x] = t2;  // each thread rewrites its own column's first slot only
  }
  __syncthreads();
  if ((int)threadIdx.x < npts) {
    uint64_t col[L];
#pragma unroll
    for (int w = 0; w < L; ++w) col[w] = s_cols[threadIdx.x * L + w];
    const Elem<L> g = reduce_word_columns<F>(col);
#pragma unroll
    for (int w = 0; w < L; ++w) s_g[threadIdx.x * L + w] = g.v[w];
  }
  __syncthreads();
}

// Sums the per-block partials of a pass and finishes the round (g_j, r_j). Used by the
// finalize kernel and by the last block of a pass kernel. From round 1 on the round's claim
// is read from st->claim (store_claim).
template <class F, int PMAX>
__device__ __forceinline__ void reduce_partials_and_finish(const uint32_t* partials, uint32_t nblocks,
                                                           int npts, int round, int num_rounds, int mode,
                                                           const FieldConsts* fc, ProofState* st,
                                                           const uint32_t* explicit_r, uint32_t* s_red,
                                                           uint32_t* s_g, uint32_t* s_msg, uint32_t* s_r,
                                                           uint32_t* s_claim, int scale_pow) {
  constexpr int L = F::L;
  const bool early = round > 0;
  if (early && threadIdx.x < L) s_claim[threadIdx.x] = st->claim[threadIdx.x];
  __shared__ uint64_t s_cols[4 * BLOCK];
  sum_partials_cols<F, PMAX>(partials, nblocks, npts, s_cols, s_g);  // its barriers publish s_claim
  if (scale_pow != 0) {  // the pass summed g / 32^scale_pow
    if (threadIdx.x < npts) {
      const Elem<L> g = mont_mul<F>(load_elem<L>(s_g + threadIdx.x * L), load_elem<L>(fc->pow32[scale_pow]));
#pragma unroll
      for (int k = 0; k < L; ++k) s_g[threadIdx.x * L + k] = g.v[k];
    }
    __syncthreads();
  }
  finish_round<F, BLOCK>(s_g, npts, round, num_rounds, mode, fc, st, explicit_r, s_msg, s_r, false, true, true,
                         early ? s_claim : nullptr);
}

// The 255-bit running sum of one point, kept in LDS between updates: limb i of this thread's
// sum at p[i * BLOCK] (one word array per limb, so a wave's accesses hit distinct banks).
struct AccLds256 {
  uint32_t* p;
  using Acc = Work<FieldP256>::Acc;
  __device__ __forceinline__ Acc load() const {
    Acc a;
#pragma unroll
    for (int i = 0; i < 9; ++i) a.v.l[i] = p[i * BLOCK];
    return a;
  }
  __device__ __forceinline__ void save(const Acc& a) {
#pragma unroll
    for (int i = 0; i < 9; ++i) p[i * BLOCK] = a.v.l[i];
  }
  __device__ __forceinline__ void clear() {
    Acc a;
    a.clear();
    save(a);
  }
  __device__ __forceinline__ void add(const R29& x) {
    Acc a = load();
    a.add(x);
    save(a);
  }
  __device__ __forceinline__ void sub(const R29& x) {
    Acc a = load();
    a.sub(x);
    save(a);
  }
};

// A pass kernel's own finish of the round (fused_finalize): the block that increments the
// counter last sees every partial and the claim (each block's thread 0 released its stores
// with the fence before its increment; the claim block's thread 0 stored the claim) and sums
// the partials and derives the challenge.
template <class F, int PMAX, int SID>
__device__ __forceinline__ void pass_fused_finish(const PassArgs& a, int npts, uint32_t* s_red, uint32_t* s_g,
                                                  uint32_t* s_msg, uint32_t* s_r, uint32_t* s_claim, int* s_last) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned prev = atomicAdd(&a.st->counter, 1u);
    *s_last = prev == gridDim.x - 1;
  }
  __syncthreads();
  if (!*s_last) return;
  __threadfence();  // acquire the other blocks' partials
  constexpr int kScale = (F::L == 8 && SID != 0) ? static_round0_scale(SID) : 0;
  reduce_partials_and_finish<F, PMAX>(a.partials, evaluating_blocks(a.claim_block), npts, a.round, a.num_rounds,
                                      a.challenge_mode, a.fc, a.st, a.explicit_r, s_red, s_g, s_msg, s_r, s_claim,
                                      a.round == 0 ? kScale : 0);
  if (threadIdx.x == 0) a.st->counter = 0u;  // ready for the next pass
}

// Waves per SIMD the compiler must leave room for: two for the compiled 255-bit compositions
// (the a*b*c + d*e kernel holds 227 registers per lane; left free, the compiler may spend more
// and fall to one wave per SIMD).
template <class F, int SID>
constexpr int pass_min_waves() {
  return (F::L == 8 && SID != 0) ? 2 : 1;
}

template <class F, int TMAX, int PMAX, int SID>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(pass_min_waves<F, SID>())))
pass_kernel(PassArgs a) {
  constexpr int L = F::L;
  __shared__ uint32_t s_red[(BLOCK / 64) * PMAX * L];
  __shared__ uint32_t s_g[PMAX * L];
  __shared__ __align__(16) uint32_t s_msg[MSG_WORDS];
  __shared__ uint32_t s_r[L];
  __shared__ uint32_t s_claim[L];
  __shared__ int s_last;
  const CompDesc* __restrict__ cd = a.comp;
  const int T = Shape<SID>::tables(cd);
  const int K = Shape<SID>::terms(cd);
  const int npts = Shape<SID>::npts(cd);
  const bool fold = a.fold != 0;
  if ((int)blockIdx.x == a.claim_block) {
    store_claim<F>(a.st, a.round, npts, a.fc);
    if (a.fused_finalize) {
      pass_fused_finish<F, PMAX, SID>(a, npts, s_red, s_g, s_msg, s_r, s_claim, &s_last);
    }
    return;
  }
  using W = Work<F>;
  typename W::Acc acc[PMAX];
#pragma unroll
  for (int t = 0; t < PMAX; ++t) acc[t].clear();
  typename W::V r = W::from_mont(a.st->r[fold ? a.round - 1 : 0]);  // unused in round 0
  const bool src_mont = a.round <= 1;
  const TableMap sm = source_map(a.round, a.src_stride, a.src_n, T, a.chunk_shift);
  const TableMap dm = fold_map(a.src_n >> 1, T, a.chunk_shift);

  const uint64_t stride = (uint64_t)evaluating_blocks(a.claim_block) * BLOCK;
  // The 2^127 - 1 field loads all tables of a pair first, in one loop per round kind
  // (load_pair_fixed; a*b*c + d*e num_vars 24 on MI355X: 1.69 -> 1.63 ms, same box). The
  // 255-bit field loads one table ahead of the one it converts (load_pair_pipelined): a whole
  // pair of its tables does not fit its registers. The other fields keep the per-table order:
  // grouped loads measured no faster at 32 bits and slower at 64 (Goldilocks round 1 798 ->
  // 825 us).
  if (L <= 4 && !fold) {
    // Round 0 only reads: one pair per iteration, loads then evaluation, the other waves of the
    // SIMD (four to five) covering each wave's memory round trip. Same-box timings on MI355X
    // (Goldilocks num_vars 26 round 0): 488-493 us this way, 502-514 us with two pairs per
    // iteration (the second pair's loads in flight during the first pair's evaluation, kept in
    // that order by scheduling barriers); a one-ahead prefetch held across iterations made the
    // compiler's load waits stall the evaluation. Round 0 reads the input (table-major).
    const TableMap im = strided_map(a.src_stride);
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < a.pairs; i += stride) {
      typename W::V cur[TMAX], oth[TMAX];
      load_pair_fixed<F, TMAX, false, true>(cur, oth, a.src, a.dst, im, dm, i, T, r);
      eval_points<F, TMAX, PMAX, SID>(cur, oth, acc, cd, T, K, npts, false);
    }
  } else if constexpr (L == 4) {
    auto sweep = [&](auto fold_c, auto mont_c) {
      constexpr bool FOLD = decltype(fold_c)::value, MONT = decltype(mont_c)::value;
      for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < a.pairs; i += stride) {
        typename W::V cur[TMAX], oth[TMAX];
        load_pair_fixed<F, TMAX, FOLD, MONT>(cur, oth, a.src, a.dst, sm, dm, i, T, r);
        eval_points<F, TMAX, PMAX, SID>(cur, oth, acc, cd, T, K, npts, FOLD);
      }
    };
    if (W::kMontIsStored || !src_mont) {
      sweep(std::true_type{}, std::false_type{});
    } else {
      sweep(std::true_type{}, std::true_type{});
    }
  } else if constexpr (L == 8) {
    // The 255-bit field: one loop per round kind, loads one table ahead (round 0 measured 7%
    // faster, rounds 1-3 5-6%, a*b*c + d*e num_vars 24 on MI355X).
    // Folds multiply by the round's challenge through the block's fold table (no Montgomery
    // reduction; Work<FieldP256>::fold2_tab).
    __shared__ __align__(16) uint32_t s_tab[fold_table_words<F>()];
    if (fold) build_fold_table<F>(s_tab, a.st->r[a.round - 1], src_mont);
    const typename W::V r32 = W::kCheapMontInput && fold && src_mont && !W::kFoldTable ? W::times32(r) : r;
    // Folding rounds of compiled compositions: the running sums live in LDS between pairs
    // (AccLds256), which frees their registers (227 instead of 255 + 2 spilled) for table 0 of
    // the thread's next pair, requested before this pair's evaluation so the wave does not
    // start each pair on a whole memory round trip (a*b*c + d*e num_vars 24 on MI355X, same
    // box, two A/B pairs: pass kernel 3773-3823 -> 3863-3987 GB/s). Interpreted compositions
    // keep their sums in registers (eight points would need 72 KiB of LDS per block).
    constexpr bool kPipe = SID != 0;
    __shared__ uint32_t s_acc[kPipe ? PMAX * 9 * BLOCK : 1];
    auto sweep = [&](auto fold_c, auto mont_c) {
      constexpr bool FOLD = decltype(fold_c)::value, MONT = decltype(mont_c)::value;
      constexpr int NW = FOLD ? 4 * L : 2 * L;
      uint32_t w[2][NW];
      if constexpr (FOLD && kPipe) {
        static_assert(std::is_same_v<F, FieldP256>, "LDS running sums of the 255-bit form");
        AccLds256 lacc[PMAX];
#pragma unroll
        for (int t = 0; t < PMAX; ++t) {
          lacc[t].p = s_acc + t * 9 * BLOCK + threadIdx.x;
          lacc[t].clear();
        }
        uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
        if (i < a.pairs) load_words<NW>(w[0], a.src + sm.at(0, 4 * i) * L);
        for (; i < a.pairs; i += stride) {
          typename W::V cur[TMAX], oth[TMAX];
          load_pair_pipelined<F, TMAX, FOLD, MONT, SID, true>(cur, oth, a.src, a.dst, sm, dm, i, T, r, r32, s_tab, w);
          if (i + stride < a.pairs) load_words<NW>(w[0], a.src + sm.at(0, 4 * (i + stride)) * L);
          uint32_t z = 0;
          asm volatile("" : "+v"(z));  // the sums are re-read from LDS at every pair, not hoisted
          AccLds256 pacc[PMAX];
#pragma unroll
          for (int t = 0; t < PMAX; ++t) pacc[t].p = lacc[t].p + z;
          eval_points<F, TMAX, PMAX, SID>(cur, oth, pacc, cd, T, K, npts, FOLD);
        }
#pragma unroll
        for (int t = 0; t < PMAX; ++t) acc[t] = lacc[t].load();
      } else {
        for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < a.pairs; i += stride) {
          typename W::V cur[TMAX], oth[TMAX];
          load_pair_pipelined<F, TMAX, FOLD, MONT, SID>(cur, oth, a.src, a.dst, sm, dm, i, T, r, r32, s_tab, w);
          eval_points<F, TMAX, PMAX, SID>(cur, oth, acc, cd, T, K, npts, FOLD);
        }
      }
    };
    if (!fold) {
      sweep(std::false_type{}, std::true_type{});
    } else if (src_mont) {
      sweep(std::true_type{}, std::true_type{});
    } else {
      sweep(std::true_type{}, std::false_type{});
    }
  } else {
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < a.pairs; i += stride) {
      typename W::V cur[TMAX], oth[TMAX];
      load_pair<F, TMAX>(cur, oth, a.src, a.dst, sm, dm, i, T, fold, src_mont, r);
      eval_points<F, TMAX, PMAX, SID>(cur, oth, acc, cd, T, K, npts, fold);
    }
  }
  Elem<L> sums[PMAX];
#pragma unroll
  for (int t = 0; t < PMAX; ++t) sums[t] = acc[t].value();
  block_reduce<F, PMAX, BLOCK>(sums, npts, s_red, s_g);
  if (threadIdx.x < npts * L) {
    a.partials[(uint64_t)blockIdx.x * (MAX_POINTS * L) + threadIdx.x] = s_g[threadIdx.x];
    if (a.fused_finalize) __threadfence();  // release this block's partial device-wide
  }
  if (a.fused_finalize) pass_fused_finish<F, PMAX, SID>(a, npts, s_red, s_g, s_msg, s_r, s_claim, &s_last);
}

template <class F, int PMAX>
__global__ void __launch_bounds__(BLOCK) finalize_kernel(FinalizeArgs a) {
  constexpr int L = F::L;
  __shared__ uint32_t s_red[(BLOCK / 64) * PMAX * L];
  __shared__ uint32_t s_g[PMAX * L];
  __shared__ __align__(16) uint32_t s_msg[MSG_WORDS];
  __shared__ uint32_t s_r[L];
  __shared__ uint32_t s_claim[L];
  reduce_partials_and_finish<F, PMAX>(a.partials, a.nblocks, a.comp->npts, a.round, a.num_rounds,
                                      a.challenge_mode, a.fc, a.st, a.explicit_r, s_red, s_g, s_msg, s_r,
                                      s_claim, a.scale_pow);
}

// One workgroup of TB threads (more than a pass block: the tail's rounds have no other
// workgroups to overlap with, so it uses a whole CU).
template <class F, int TMAX, int PMAX, int SID, int TB>
__device__ __forceinline__ void tail_rounds(const TailArgs& a) {
  constexpr int L = F::L;
  __shared__ uint32_t s_red[(TB / 64) * PMAX * L];
  __shared__ uint32_t s_g[PMAX * L];
  __shared__ __align__(16) uint32_t s_msg[MSG_WORDS];
  __shared__ uint32_t s_r[L];
  __shared__ uint32_t s_claim[L];
  // The 255-bit field sums a round's points with block_reduce_cols (tail round 30 -> 27 us at
  // num_vars 24; at 64 bits the element-wise reduction measured faster, 10.7 against 11.6 us):
  // its transposed rows and column sums (a word each where unused).
  constexpr bool kColReduce = L == 8 && PMAX <= 4;
  __shared__ uint32_t s_tr[kColReduce ? cols_lds_words<TB, PMAX * L>() : 1];
  __shared__ uint64_t s_c[kColReduce ? (TB / 64) * PMAX * L : 1];
  const CompDesc* __restrict__ cd = a.comp;
  const int T = Shape<SID>::tables(cd);
  const int K = Shape<SID>::terms(cd);
  const int npts = Shape<SID>::npts(cd);

  if (threadIdx.x < L) s_r[threadIdx.x] = a.first_round > 0 ? a.st->r[a.first_round - 1][threadIdx.x] : 0u;
  __syncthreads();
  for (int j = a.first_round; j < a.num_rounds; ++j) {
    const bool fold = j > 0;
    const uint32_t* src = j <= 1 ? a.input : (((j - 1) & 1) ? a.buf_a : a.buf_b);
    uint32_t* dst = (j & 1) ? a.buf_a : a.buf_b;
    const uint64_t src_n = fold ? (a.n0 >> (j - 1)) : a.n0;
    const TableMap sm = source_map(j, a.in_stride, src_n, T, a.chunk_shift);
    const TableMap dm = fold_map(src_n >> 1, T, a.chunk_shift);
    const uint64_t pairs = a.n0 >> (j + 1);
    using W = Work<F>;
    const typename W::V r = W::from_mont(s_r);
    typename W::Acc acc[PMAX];
#pragma unroll
    for (int t = 0; t < PMAX; ++t) acc[t].clear();
    // a round whose pairs leave the last wave idle has it compute the claim meanwhile (from the
    // previous round's polynomial and challenge, written before the last barrier)
    const bool early = fold && pairs <= (uint64_t)(TB - 64);
    if (early && (threadIdx.x >> 6) == TB / 64 - 1) {
      const Elem<L> c = previous_claim<F>(a.st, j, npts, a.fc);
      if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < L; ++k) s_claim[k] = c.v[k];
      }
    } else {
      for (uint64_t i = threadIdx.x; i < pairs; i += TB) {
        typename W::V cur[TMAX], oth[TMAX];
        load_pair<F, TMAX>(cur, oth, src, dst, sm, dm, i, T, fold, j <= 1, r);
        eval_points<F, TMAX, PMAX, SID, false>(cur, oth, acc, cd, T, K, npts, fold);
      }
    }
    Elem<L> sums[PMAX];
#pragma unroll
    for (int t = 0; t < PMAX; ++t) sums[t] = acc[t].value();
    if constexpr (kColReduce) {
      block_reduce_cols<F, PMAX, TB>(sums, npts, s_tr, s_c, s_g);  // its barriers publish s_claim too
    } else {
      block_reduce<F, PMAX, TB>(sums, npts, s_red, s_g);
    }
    finish_round<F, TB>(s_g, npts, j, a.num_rounds, a.challenge_mode, a.fc, a.st, a.explicit_r, s_msg, s_r,
                        a.want_final != 0, true, false, early ? s_claim : nullptr);
    __syncthreads();
  }
  if (!a.want_final) return;
  // Final-point step: fold the last round's pair of every table with the last challenge.
  // The last round's entries are in its output buffer, or the input when there is one round.
  const int jl = a.num_rounds - 1;
  if (threadIdx.x < T) {
    using W = Work<F>;
    const int v = threadIdx.x;
    typename W::V lo, hi;
    if (jl == 0) {
      lo = W::from_mont(a.input + (uint64_t)v * a.in_stride * L);
      hi = W::from_mont(a.input + ((uint64_t)v * a.in_stride + 1) * L);
    } else {
      const uint32_t* last = (jl & 1) ? a.buf_a : a.buf_b;
      const TableMap lm = fold_map(2, T, a.chunk_shift);
      lo = W::from_stored(last + lm.at(v, 0) * L);
      hi = W::from_stored(last + lm.at(v, 1) * L);
    }
    typename W::Acc f;
    f.clear();
    f.add(W::fold(lo, hi, W::from_mont(s_r)));
    const Elem<L> m = f.value();
#pragma unroll
    for (int k = 0; k < L; ++k) a.st->final_vals[v][k] = m.v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) a.st->final_ready = 1u;
}

template <class F, int TMAX, int PMAX, int SID, int TB>
__global__ void __launch_bounds__(TB) tail_kernel(TailArgs a) {
  tail_rounds<F, TMAX, PMAX, SID, TB>(a);
}

// ---------------------------------------------------------------------------------------------
// Point-parallel rounds. The per-pair kernels above hold 2 x tables working values and PMAX
// accumulators per thread: with more than 8 tables they would spill, and at 255 bits they run
// one wave per SIMD. Such rounds run as two launches instead: a fold launch writes round j's
// tables (fold of round j-1 with r_{j-1}); the evaluation gives each pair G = PMAX lanes, lane
// t evaluating the composition at point t only, so a lane holds one value per table and one
// accumulator. Same partials layout as pass_kernel (batch_pass_kernel for the batched forms).
// ---------------------------------------------------------------------------------------------
// Elements idx = begin, begin + stride, ... < end of round j's tables (T tables of n_new
// entries, idx = table * n_new + entry), each folded from two entries of round j-1's src.
template <class F>
__device__ __forceinline__ void wide_fold_range(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                uint64_t src_n, const TableMap& sm, const TableMap& dm,
                                                bool src_mont, const typename Work<F>::V& r, uint64_t begin,
                                                uint64_t end, uint64_t stride, const uint32_t* tab) {
  using W = Work<F>;
  constexpr int L = F::L;
  const uint64_t n_new = src_n >> 1;  // a power of two
  const int shift = __builtin_ctzll(n_new);
  for (uint64_t idx = begin; idx < end; idx += stride) {
    const uint64_t v = idx >> shift, i = idx & (n_new - 1);
    uint32_t w[2 * L];
    load_words<2 * L>(w, src + sm.at(v, 2 * i) * L);
    uint32_t y[L];
    if constexpr (W::kFoldTable) {
      W::store(y, src_mont ? W::template fold1_tab<true>(w, w + L, tab) : W::template fold1_tab<false>(w, w + L, tab));
    } else {
      const typename W::V lo = src_mont ? W::from_mont(w) : W::from_stored(w);
      const typename W::V hi = src_mont ? W::from_mont(w + L) : W::from_stored(w + L);
      W::store(y, W::fold(lo, hi, r));
    }
    uint32_t* d = dst + dm.at(v, i) * L;
#pragma unroll
    for (int k = 0; k < L; ++k) d[k] = y[k];
  }
}

// x[v] for a wave-uniform table slot v, with x in registers: a switch, so a scalar branch to
// the slot's copy (an indexed read of a register array would go through scratch memory; an
// if-chain over the slots is folded back into such a read by the compiler, a switch is not).
#define SC_SLOT_CASE(I, STMT) \
  case I:                     \
    if constexpr (I < N) {    \
      constexpr int kSlot = I; \
      STMT;                   \
    }                         \
    break;
#define SC_SLOT_SWITCH(V, STMT)                                                                 \
  switch (V) {                                                                                  \
    SC_SLOT_CASE(0, STMT) SC_SLOT_CASE(1, STMT) SC_SLOT_CASE(2, STMT) SC_SLOT_CASE(3, STMT)       \
    SC_SLOT_CASE(4, STMT) SC_SLOT_CASE(5, STMT) SC_SLOT_CASE(6, STMT) SC_SLOT_CASE(7, STMT)       \
    SC_SLOT_CASE(8, STMT) SC_SLOT_CASE(9, STMT) SC_SLOT_CASE(10, STMT) SC_SLOT_CASE(11, STMT)     \
    SC_SLOT_CASE(12, STMT) SC_SLOT_CASE(13, STMT) SC_SLOT_CASE(14, STMT) SC_SLOT_CASE(15, STMT)   \
    SC_SLOT_CASE(16, STMT) SC_SLOT_CASE(17, STMT) SC_SLOT_CASE(18, STMT) SC_SLOT_CASE(19, STMT)   \
    SC_SLOT_CASE(20, STMT) SC_SLOT_CASE(21, STMT) SC_SLOT_CASE(22, STMT) SC_SLOT_CASE(23, STMT)   \
    SC_SLOT_CASE(24, STMT) SC_SLOT_CASE(25, STMT) SC_SLOT_CASE(26, STMT) SC_SLOT_CASE(27, STMT)   \
    SC_SLOT_CASE(28, STMT) SC_SLOT_CASE(29, STMT) SC_SLOT_CASE(30, STMT) SC_SLOT_CASE(31, STMT)   \
    default:                                                                                    \
      break;                                                                                    \
  }
static_assert(MAX_TABLES == 32, "one switch case per table slot");

template <class F, int N>
__device__ __forceinline__ void mul_slot(typename Work<F>::V& prod, const typename Work<F>::V (&x)[N], int v) {
  SC_SLOT_SWITCH(v, prod = Work<F>::mul(prod, x[kSlot]))
}

template <class F, int N>
__device__ __forceinline__ typename Work<F>::V pick_slot(const typename Work<F>::V (&x)[N], int v) {
  typename Work<F>::V out = x[0];
  SC_SLOT_SWITCH(v, out = x[kSlot])
  return out;
}

// C(x) by the factor lists of the descriptor: every table index a term names is wave-uniform,
// so each factor is a scalar branch to a multiplication by a fixed register slot. A general
// coefficient starts the product (its multiplication is one of the term's len); a +-1 term
// starts from its first factor. At 255 bits, the first factor of a term from a table in sel01
// (round-0 entries all 0 or 1) after the product's start is applied last as a multiply by the
// lane's small integer: selc holds it for table v in bits 4v..4v+3, plus 8.
template <class F, int N>
__device__ __forceinline__ void compose_slots(typename Work<F>::Acc& acc, const typename Work<F>::V (&x)[N],
                                              const CompDesc* __restrict__ cd, int K, int min_len,
                                              uint32_t sel01 = 0u, uint64_t selc = 0u) {
  using W = Work<F>;
  for (int k = 0; k < K; ++k) {
    const int len = cd->len[k];
    const int kind = cd->kind[k];
    typename W::V prod;
    int m = 0;
    if (kind == COEFF_GENERAL) {
      prod = W::coeff(cd->coeff[k]);
    } else {
      prod = pick_slot<F, N>(x, cd->fac[k][0]);
      m = 1;
    }
    int small = -1;  // table of the factor applied as a small integer (wave-uniform)
    for (; m < len; ++m) {
      const int v = cd->fac[k][m];
      if constexpr (F::L == 8) {
        if (small < 0 && ((sel01 >> v) & 1u)) {
          small = v;
          continue;
        }
      }
      mul_slot<F, N>(prod, x, v);
    }
    if constexpr (F::L == 8) {
      if (small >= 0) prod = W::scale_small(prod, (int)((selc >> (4 * small)) & 15u) - 8);
    }
    if (len >= min_len) {
      if (kind == COEFF_MINUS_ONE) acc.sub(prod); else acc.add(prod);
    }
  }
}

// Point-parallel evaluation of pairs [begin, end) of round `round`: lane q of a group of
// P = wide_points(npts, round) lanes evaluates the composition at the q-th evaluated point of
// pair begin + threadIdx.x / P, then every stride pairs (lanes past the last whole group
// idle). Writes the block's sums for t < npts to part[t * L + k] (zero for t = 1 from round
// 1 on, which the round's finish derives from the claim). s_red: BLOCK * L words.
template <class F, int TMAX>
__device__ __forceinline__ void wide_eval_block(const uint32_t* __restrict__ src, const TableMap& sm, bool src_mont,
                                                int round, uint64_t begin, uint64_t end, uint64_t stride,
                                                const CompDesc* __restrict__ cd, uint32_t* __restrict__ part,
                                                uint32_t* s_red, uint32_t sel01 = 0u) {
  using W = Work<F>;
  constexpr int L = F::L;
  const int T = cd->num_tables;
  const int K = cd->num_terms;
  const int npts = cd->npts;
  const int P = wide_points(npts, round);
  const int pairs_here = BLOCK / P;
  const int q = threadIdx.x % P;
  const int group = threadIdx.x / P;
  const int t = (round > 0 && q >= 1) ? q + 1 : q;
  // the lane of t = d >= 2 evaluates the leading coefficient (the round's finish turns it
  // into g(d)): the differences as the values, terms of d factors only
  const bool lead = npts >= 3 && t == npts - 1;
  const int min_len = lead ? npts - 1 : 0;
  typename W::Acc acc;
  acc.clear();
  // Register slots while TMAX values fit in 160 VGPRs; 32 values of the 255-bit field do not,
  // and keep the term loop (x then lives in scratch memory).
  constexpr bool kSlots = TMAX > 8 && TMAX * (int)(sizeof(typename W::V) / 4) <= 160;
  // 0/1 tables as small-integer factors: 255 bits, register slots, at most 16 tables
  constexpr bool kSmall = F::L == 8 && kSlots && TMAX <= 16;
  if constexpr (!kSmall) sel01 = 0u;
  if (group < pairs_here) {
    for (uint64_t i = begin + group; i < end; i += stride) {
      typename W::V x[TMAX];
      uint64_t selc = 0;  // table v's value at t in bits 4v..4v+3, plus 8 (tables in sel01)
      static_for<TMAX>([&](auto vi) {
        constexpr int v = decltype(vi)::value;
        if (v < T) {
          uint32_t w[2 * L];
          load_words<2 * L>(w, src + sm.at(v, 2 * i) * L);
          if constexpr (kSmall) {
            if ((sel01 >> v) & 1u) {  // entries 0 or 1: any nonzero word means 1
              uint32_t o0 = 0, o1 = 0;
#pragma unroll
              for (int k = 0; k < L; ++k) {
                o0 |= w[k];
                o1 |= w[L + k];
              }
              const int s0 = o0 != 0u, s1 = o1 != 0u;
              const int c = lead ? s1 - s0 : s0 + t * (s1 - s0);  // in [-5, 6]: t <= 6
              selc |= (uint64_t)(uint32_t)(c + 8) << (4 * v);
            }
          }
          const typename W::V lo = src_mont ? W::from_mont(w) : W::from_stored(w);
          const typename W::V hi = src_mont ? W::from_mont(w + L) : W::from_stored(w + L);
          const typename W::V dv = W::diff(hi, lo);
          x[v] = lead ? dv : W::at(lo, dv, t);
        } else {
          x[v] = x[0];  // defined (an undefined slot measured 50 more VGPRs), never read
        }
      });
      if constexpr (kSlots) {
        compose_slots<F, TMAX>(acc, x, cd, K, min_len, sel01, selc);
      } else {
        compose_add<F, TMAX>(acc, x, cd, K, min_len);
      }
    }
  }
  // Sum each point's lanes: lane values to LDS, then per point a tree over its groups.
  const Elem<L> mine = acc.value();
#pragma unroll
  for (int k = 0; k < L; ++k) s_red[threadIdx.x * L + k] = group < pairs_here ? mine.v[k] : 0u;
  __syncthreads();
  for (int half = 128; half >= 1; half >>= 1) {
    // lane (g, q) with g < half adds group g + half of the same point
    if (group < half && group + half < pairs_here) {
      const Elem<L> a = load_elem<L>(s_red + threadIdx.x * L);
      const Elem<L> b = load_elem<L>(s_red + (threadIdx.x + half * P) * L);
      const Elem<L> c = add<F>(a, b);
#pragma unroll
      for (int k = 0; k < L; ++k) s_red[threadIdx.x * L + k] = c.v[k];
    }
    __syncthreads();
  }
  if (threadIdx.x < npts) {
    const int tt = threadIdx.x;
    const int qq = (round > 0 && tt >= 1) ? tt - 1 : tt;  // lane of point tt (tt = 1 skipped)
    const bool skipped = round > 0 && tt == 1;
#pragma unroll
    for (int k = 0; k < L; ++k) part[tt * L + k] = skipped ? 0u : s_red[qq * L + k];
  }
}

template <class F>
__global__ void __launch_bounds__(BLOCK) wide_fold_kernel(PassArgs a) {
  __shared__ __align__(16) uint32_t s_tab[fold_table_words<F>()];
  build_fold_table<F>(s_tab, a.st->r[a.round - 1], a.round <= 1);
  const int T = a.comp->num_tables;
  const uint64_t total = (a.src_n >> 1) * (uint64_t)T;
  wide_fold_range<F>(a.src, a.dst, a.src_n, source_map(a.round, a.src_stride, a.src_n, T, a.chunk_shift),
                     fold_map(a.src_n >> 1, T, a.chunk_shift),
                     a.round <= 1, Work<F>::from_mont(a.st->r[a.round - 1]),
                     (uint64_t)blockIdx.x * BLOCK + threadIdx.x, total, (uint64_t)gridDim.x * BLOCK, s_tab);
}

template <class F, int TMAX, int G>
__global__ void __launch_bounds__(BLOCK) wide_eval_kernel(PassArgs a) {
  __shared__ uint32_t s_red[BLOCK * F::L];
  if ((int)blockIdx.x == a.claim_block) {
    store_claim<F>(a.st, a.round, a.comp->npts, a.fc);
    return;
  }
  const uint64_t pairs_here = wide_pairs_per_block(a.comp->npts, a.round);
  const TableMap sm = a.round == 0 ? strided_map(a.src_stride) : fold_map(a.src_n, a.comp->num_tables, a.chunk_shift);
  wide_eval_block<F, TMAX>(a.src, sm, a.round == 0, a.round, (uint64_t)blockIdx.x * pairs_here, a.pairs,
                           (uint64_t)evaluating_blocks(a.claim_block) * pairs_here, a.comp,
                           a.partials + (uint64_t)blockIdx.x * MAX_POINTS * F::L, s_red,
                           a.round == 0 ? a.sel01 : 0u);
}

// ---------------------------------------------------------------------------------------------
// Batched stages
// ---------------------------------------------------------------------------------------------

// Entry of block b: the last entry whose tile_begin <= b (entries sorted by tile_begin).
__device__ __forceinline__ uint32_t find_entry(const StageEntry* __restrict__ e, uint32_t n, uint32_t b) {
  uint32_t lo = 0, hi = n;  // invariant: e[lo].tile_begin <= b < e[hi].tile_begin
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (e[mid].tile_begin <= b) lo = mid; else hi = mid;
  }
  return lo;
}

// One stage: block b runs tile (b - tile_begin) of its entry's round, exactly like a
// pass_kernel block (fold with r_{round-1}, evaluate, reduce), writing one partial per tile.
template <class F, int TMAX, int PMAX>
__global__ void __launch_bounds__(BLOCK) batch_pass_kernel(BatchArgs a) {
  constexpr int L = F::L;
  __shared__ uint32_t s_red[(BLOCK / 64) * PMAX * L];
  __shared__ uint32_t s_g[PMAX * L];
  const uint32_t k = find_entry(a.entries, a.n_entries, blockIdx.x);
  const StageEntry ent = a.entries[k];
  const InstDesc& in = a.insts[ent.inst];
  const CompDesc* __restrict__ cd = in.comp;
  const int T = cd->num_tables;
  const int K = cd->num_terms;
  const int npts = cd->npts;
  const int j = (int)ent.round;
  const bool fold = j > 0;
  const uint32_t* src = j <= 1 ? in.input : (((j - 1) & 1) ? in.buf_a : in.buf_b);
  uint32_t* dst = (j & 1) ? in.buf_a : in.buf_b;
  const uint64_t src_n = fold ? (in.n0 >> (j - 1)) : in.n0;
  const uint64_t pairs = in.n0 >> (j + 1);
  const TableMap sm = source_map(j, in.n0, src_n, T, a.chunk_shift);  // inputs are packed
  const TableMap dm = fold_map(src_n >> 1, T, a.chunk_shift);
  using W = Work<F>;
  const typename W::V r = W::from_mont(in.st->r[fold ? j - 1 : 0]);  // unused in round 0
  __shared__ __align__(16) uint32_t s_tab[fold_table_words<F>()];
  if (fold) build_fold_table<F>(s_tab, in.st->r[j - 1], j <= 1);

  typename W::Acc acc[PMAX];
#pragma unroll
  for (int t = 0; t < PMAX; ++t) acc[t].clear();
  const uint64_t tile = blockIdx.x - ent.tile_begin;
  const uint64_t base = tile * (uint64_t)BLOCK * ent.items;
  for (uint32_t it = 0; it < ent.items; ++it) {
    const uint64_t i = base + (uint64_t)it * BLOCK + threadIdx.x;
    if (i < pairs) {
      typename W::V cur[TMAX], oth[TMAX];
      load_pair<F, TMAX, true>(cur, oth, src, dst, sm, dm, i, T, fold, j <= 1, r, s_tab);
      eval_points<F, TMAX, PMAX, 0>(cur, oth, acc, cd, T, K, npts, fold);
    }
  }
  Elem<L> sums[PMAX];
#pragma unroll
  for (int t = 0; t < PMAX; ++t) sums[t] = acc[t].value();
  block_reduce<F, PMAX, BLOCK>(sums, npts, s_red, s_g);
  if (threadIdx.x < npts * L) {
    a.partials[(uint64_t)blockIdx.x * (MAX_POINTS * MAX_LIMBS) + threadIdx.x] = s_g[threadIdx.x];
  }
}

// Point-parallel form of a stage (entries whose round runs point-parallel). Fold: block b
// folds elements of tile (b - tile_begin) of its entry's round (BLOCK * items elements of
// the T * n_round table entries per tile); entries have round >= 1.
template <class F>
__global__ void __launch_bounds__(BLOCK) batch_wide_fold_kernel(BatchArgs a) {
  const StageEntry ent = a.entries[find_entry(a.entries, a.n_entries, blockIdx.x)];
  const InstDesc& in = a.insts[ent.inst];
  const int j = (int)ent.round;
  const uint32_t* src = j <= 1 ? in.input : (((j - 1) & 1) ? in.buf_a : in.buf_b);
  uint32_t* dst = (j & 1) ? in.buf_a : in.buf_b;
  const uint64_t src_n = in.n0 >> (j - 1);
  const int T = in.comp->num_tables;
  const uint64_t total = (src_n >> 1) * (uint64_t)T;
  const uint64_t span = (uint64_t)BLOCK * ent.items;
  const uint64_t begin = (uint64_t)(blockIdx.x - ent.tile_begin) * span;
  __shared__ __align__(16) uint32_t s_tab[fold_table_words<F>()];
  build_fold_table<F>(s_tab, in.st->r[j - 1], j <= 1);
  wide_fold_range<F>(src, dst, src_n, source_map(j, in.n0, src_n, T, a.chunk_shift),
                     fold_map(src_n >> 1, T, a.chunk_shift), j <= 1,
                     Work<F>::from_mont(in.st->r[j - 1]), begin + threadIdx.x,
                     begin + span < total ? begin + span : total, BLOCK, s_tab);
}

// Evaluation: block b (tile index b + tile_base in the stage's partials) evaluates
// wide_pairs_per_block(npts, round) * items pairs of its entry's round and writes the tile's partial.
template <class F, int TMAX, int G>
__global__ void __launch_bounds__(BLOCK) batch_wide_eval_kernel(BatchArgs a) {
  __shared__ uint32_t s_red[BLOCK * F::L];
  const uint32_t b = blockIdx.x + a.tile_base;
  const StageEntry ent = a.entries[find_entry(a.entries, a.n_entries, b)];
  const InstDesc& in = a.insts[ent.inst];
  const int j = (int)ent.round;
  const uint32_t* src = j == 0 ? in.input : ((j & 1) ? in.buf_a : in.buf_b);
  const uint64_t pairs = in.n0 >> (j + 1);
  const uint64_t pairs_here = wide_pairs_per_block(in.comp->npts, j);
  const uint64_t span = pairs_here * ent.items;
  const uint64_t begin = (uint64_t)(b - ent.tile_begin) * span;
  const TableMap sm = j == 0 ? strided_map(in.n0) : fold_map(in.n0 >> j, in.comp->num_tables, a.chunk_shift);
  wide_eval_block<F, TMAX>(src, sm, j == 0, j, begin, begin + span < pairs ? begin + span : pairs,
                           pairs_here, in.comp, a.partials + (uint64_t)b * (MAX_POINTS * MAX_LIMBS), s_red);
}

// One block per entry: the entry's instance runs all its remaining rounds from entry.round in
// this block, exactly as tail_kernel does for a single proof (its tables are small by then).
template <class F, int TMAX, int PMAX>
__global__ void __launch_bounds__(BLOCK) batch_tail_kernel(BatchArgs a) {
  const StageEntry ent = a.entries[blockIdx.x];
  const InstDesc& in = a.insts[ent.inst];
  TailArgs t;
  t.input = in.input;
  t.buf_a = in.buf_a;
  t.buf_b = in.buf_b;
  t.n0 = in.n0;
  t.in_stride = in.n0;  // batch and server inputs are packed
  t.chunk_shift = a.chunk_shift;
  t.pad0 = 0;
  t.first_round = (int32_t)ent.round;
  t.num_rounds = in.num_vars;
  t.challenge_mode = in.challenge_mode;
  t.want_final = 0;
  t.comp = in.comp;
  t.fc = a.fc;
  t.st = in.st;
  t.explicit_r = nullptr;
  tail_rounds<F, TMAX, PMAX, 0, BLOCK>(t);
}

// One block per entry: sum the entry's tile partials, store g_round, derive r_round.
template <class F, int PMAX>
__global__ void __launch_bounds__(BLOCK) batch_finalize_kernel(BatchArgs a) {
  constexpr int L = F::L;
  __shared__ uint32_t s_red[(BLOCK / 64) * PMAX * L];
  __shared__ uint32_t s_g[PMAX * L];
  __shared__ __align__(16) uint32_t s_msg[MSG_WORDS];
  __shared__ uint32_t s_r[L];
  const StageEntry ent = a.entries[blockIdx.x];
  const InstDesc& in = a.insts[ent.inst];
  __shared__ uint32_t s_claim[L];
  const int npts = in.comp->npts;
  Elem<L> acc[PMAX];
#pragma unroll
  for (int t = 0; t < PMAX; ++t) acc[t] = zero<F>();
  const bool early = ent.round > 0;  // the last wave computes the claim meanwhile
  if (early && (threadIdx.x >> 6) == BLOCK / 64 - 1) {
    const Elem<L> c = previous_claim<F>(in.st, (int)ent.round, npts, a.fc);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
      for (int k = 0; k < L; ++k) s_claim[k] = c.v[k];
    }
  } else {
    const uint32_t step = early ? BLOCK - 64 : BLOCK;
    for (uint32_t b = threadIdx.x; b < ent.tiles; b += step) {
      const uint32_t* part = a.partials + (uint64_t)(ent.tile_begin + b) * (MAX_POINTS * MAX_LIMBS);
#pragma unroll
      for (int t = 0; t < PMAX; ++t) {
        if (t < npts) acc[t] = add<F>(acc[t], load_elem<L>(part + t * L));
      }
    }
  }
  block_reduce<F, PMAX, BLOCK>(acc, npts, s_red, s_g);
  finish_round<F, BLOCK>(s_g, npts, (int)ent.round, in.num_vars, in.challenge_mode, a.fc, in.st, nullptr,
                         s_msg, s_r, false, true, true, early ? s_claim : nullptr);
}

// Input generation and Montgomery encoding (cuda/src/field_sumcheck.cu:593-611).

template <class F>
__global__ void __launch_bounds__(BLOCK) fill_kernel(FillArgs a) {
  constexpr int L = F::L;
  const uint64_t total = a.n * (uint64_t)a.num_tables;
  const Elem<L> r2 = load_elem<L>(a.fc->r2);
  for (uint64_t idx = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; idx < total;
       idx += (uint64_t)gridDim.x * BLOCK) {
    const uint32_t v = (uint32_t)(idx / a.n);
    const uint64_t o = idx - (uint64_t)v * a.n;
    const uint64_t at = (uint64_t)v * a.stride + o;  // entry o of table v
    Elem<L> x;
    if (a.kind == FILL_AFFINE) {
      x = affine_element<F>(a.seed, v, o);
    } else if ((a.selector_mask >> v) & 1u) {
      x = selector_element<F>(stream_key(a.seed, a.instance, v), o);
    } else {
      x = uniform_element<F>(stream_key(a.seed, a.instance, v), o);
    }
    const Elem<L> m = mont_mul<F>(x, r2);
#pragma unroll
    for (int k = 0; k < L; ++k) a.tables[at * L + k] = m.v[k];
  }
}

// Limb-major tables [table][limb][entry] -> element-major [table][entry][limb] (input format
// of the reference's --layout limb); n entries per table, a power of two.
template <class F>
__global__ void __launch_bounds__(BLOCK) limb_to_element_kernel(const uint32_t* __restrict__ in,
                                                                uint32_t* __restrict__ out, uint64_t n,
                                                                uint64_t out_stride, int num_tables) {
  constexpr int L = F::L;
  const int shift = __builtin_ctzll(n);
  const uint64_t total = n * (uint64_t)num_tables;
  for (uint64_t idx = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; idx < total;
       idx += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t v = idx >> shift, i = idx & (n - 1);
    uint32_t w[L];
#pragma unroll
    for (int k = 0; k < L; ++k) w[k] = in[(v * L + k) * n + i];
#pragma unroll
    for (int k = 0; k < L; ++k) out[(v * out_stride + i) * L + k] = w[k];
  }
}

// Elementwise field operations, for the field-layer tests and host data conversion.
// OP_WORK_MUL and OP_WORK_FOLD run the kernels' working-form path end to end: a*b, and
// a + a (b - a) through fold, store and reload, each read back through the accumulator.
enum FieldOp : int32_t {
  OP_MUL = 0, OP_ADD = 1, OP_SUB = 2, OP_TO_MONT = 3, OP_FROM_MONT = 4, OP_WORK_MUL = 5, OP_WORK_FOLD = 6
};

template <class F>
__global__ void __launch_bounds__(BLOCK) field_op_kernel(int op, uint64_t n, const uint32_t* a,
                                                        const uint32_t* b, uint32_t* out,
                                                        const FieldConsts* fc) {
  constexpr int L = F::L;
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
    const Elem<L> x = load_elem<L>(a + i * L);
    Elem<L> y;
    if (op == OP_MUL) {
      y = mont_mul<F>(x, load_elem<L>(b + i * L));
    } else if (op == OP_ADD) {
      y = add<F>(x, load_elem<L>(b + i * L));
    } else if (op == OP_SUB) {
      y = sub<F>(x, load_elem<L>(b + i * L));
    } else if (op == OP_TO_MONT) {
      y = mont_mul<F>(x, load_elem<L>(fc->r2));
    } else if (op == OP_WORK_MUL || op == OP_WORK_FOLD) {
      using W = Work<F>;
      const typename W::V wa = W::from_mont(a + i * L), wb = W::from_mont(b + i * L);
      typename W::Acc acc;
      acc.clear();
      if (op == OP_WORK_MUL) {
        acc.add(W::mul(wa, wb));
      } else {
        uint32_t w[L];
        W::store(w, W::fold(wa, wb, wa));
        acc.add(W::from_stored(w));
      }
      y = acc.value();
    } else {
      Elem<L> one = zero<F>();
      one.v[0] = 1u;
      y = mont_mul<F>(x, one);
    }
#pragma unroll
    for (int k = 0; k < L; ++k) out[i * L + k] = y.v[k];
  }
}

// Derives one transcript challenge from g (Montgomery limbs) on the device.
template <class F>
__global__ void __launch_bounds__(BLOCK) transcript_kernel(const uint32_t* g, int npts, int round,
                                                          const FieldConsts* fc, ProofState* st) {
  constexpr int L = F::L;
  __shared__ uint32_t s_g[MAX_POINTS * L];
  __shared__ __align__(16) uint32_t s_msg[MSG_WORDS];
  __shared__ uint32_t s_r[L];
  if (threadIdx.x < npts * L) s_g[threadIdx.x] = g[threadIdx.x];
  __syncthreads();
  finish_round<F, BLOCK>(s_g, npts, round, round + 2, CHALLENGE_SHA3, fc, st, nullptr, s_msg, s_r);
}

}  // namespace sc
