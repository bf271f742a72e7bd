// Shared host/device data structures of the SumCheck library.
//
// HBM layout of a proof (element-major, like the reference's default layout,
// cuda/src/field_sumcheck.cu:143-163): table v of a round with n entries per table occupies
// words [v*n*L, (v+1)*n*L), entry i at words [i*L, (i+1)*L), little-endian limbs, Montgomery
// form. Round j of a proof with num_vars variables has n_j = 2^(num_vars - j) entries per
// table. Round 0 reads the input tables; round j >= 1 reads round j-1's tables, folds them
// with r_{j-1} and writes round j's tables into one of two scratch buffers (n/2 and n/4
// entries per table, used alternately), so a proof needs 1.75x its input in HBM.
#pragma once

#include <stdint.h>

namespace sc {

constexpr int MAX_TABLES = 32;
constexpr int MAX_TERMS = 32;
constexpr int MAX_POINTS = 8;  // degree <= 7, as cuda/src/field_sumcheck.cu:29
constexpr int MAX_ROUNDS = 40;
constexpr int MAX_LIMBS = 8;
// LDS words for a round's transcript: the padded message (up to 3 SHA3 blocks of 34 words)
// and the 25 64-bit words the wave Keccak exchanges its state through.
constexpr int MSG_WORDS = 3 * 34 + 2 * 25;
constexpr int BLOCK = 256;
constexpr int MAX_PASS_BLOCKS = 4096;

enum CoeffKind : int32_t { COEFF_ONE = 0, COEFF_MINUS_ONE = 1, COEFF_GENERAL = 2 };

// Compositions with compiled kernels: the round kernels unroll these at compile time
// (static table slots, no term interpretation). Any other composition runs through the
// interpreted kernels. The host selects a compiled one when the composition's tables,
// terms (in order), factors and coefficient kinds match exactly.
struct StaticComp {
  int num_tables;
  int num_terms;
  int lens[4];
  int vars[4][8];
  int kinds[4];
};
constexpr int NUM_STATIC_COMPS = 4;  // id 0 = interpreted
constexpr StaticComp STATIC_COMPS[NUM_STATIC_COMPS] = {
    {0, 0, {0}, {{0}}, {0}},
    {2, 1, {2}, {{0, 1}}, {COEFF_ONE}},                              // a*b (zk_spartan_2 shape)
    {3, 1, {3}, {{0, 1, 2}}, {COEFF_ONE}},                           // a*b*c (poly_abc)
    {5, 2, {3, 2}, {{0, 1, 2}, {3, 4}}, {COEFF_ONE, COEFF_ONE}},     // a*b*c + d*e (poly_abc_plus_de)
};
constexpr int static_degree(int id) {
  int d = 0;
  for (int k = 0; k < STATIC_COMPS[id].num_terms; ++k)
    if (STATIC_COMPS[id].lens[k] > d) d = STATIC_COMPS[id].lens[k];
  return d;
}

// Round 0 of a compiled composition in the 255-bit field reads the caller's Montgomery words
// X = x 2^256 directly as working values, which stand for x 2^-5 (the working form holds
// x 2^261), saving each input its reduction. A product of len such factors then carries
// 32^-len. Each term's first (len - S) factors are read exactly instead, S the shortest
// term's length, so every term carries 32^-S; the round's sum is multiplied by 32^S once,
// when its block partials are reduced. (The compiled compositions' terms share no tables.)
constexpr int static_round0_scale(int id) {
  int s = 8;
  for (int k = 0; k < STATIC_COMPS[id].num_terms; ++k)
    if (STATIC_COMPS[id].lens[k] < s) s = STATIC_COMPS[id].lens[k];
  return s;
}
constexpr bool static_round0_exact(int id, int v) {
  const int s = static_round0_scale(id);
  for (int k = 0; k < STATIC_COMPS[id].num_terms; ++k)
    for (int m = 0; m < STATIC_COMPS[id].lens[k] - s; ++m)
      if (STATIC_COMPS[id].vars[k][m] == v) return true;
  return false;
}
constexpr bool static_terms_disjoint(int id) {
  for (int k = 0; k < STATIC_COMPS[id].num_terms; ++k)
    for (int m = 0; m < STATIC_COMPS[id].lens[k]; ++m)
      for (int k2 = k + 1; k2 < STATIC_COMPS[id].num_terms; ++k2)
        for (int m2 = 0; m2 < STATIC_COMPS[id].lens[k2]; ++m2)
          if (STATIC_COMPS[id].vars[k][m] == STATIC_COMPS[id].vars[k2][m2]) return false;
  return true;
}
// Whether a compiled composition's interpolated value of table v at point t (t <= 3) may skip
// the limb normalization in the 255-bit form: only when the table is a later factor of its
// term (the second operand of the products it enters, which takes limbs below 2^31 when the
// first is normalized) and, where it enters the two-product sum of the last products (whose
// 18-product columns need limbs below 2^30), only at t = 2 (x(2) = x(1) + d stays below 2^30;
// x(3) is normalized from it).
constexpr bool static_lazy_step(int id, int v, int t) {
  const StaticComp& c = STATIC_COMPS[id];
  if (static_degree(id) > 3 || t > 3) return false;
  const bool shared_sum = c.num_terms == 2 && c.kinds[0] == COEFF_ONE && c.kinds[1] == COEFF_ONE && c.lens[0] >= 2 &&
                          c.lens[1] >= 2;
  for (int k = 0; k < c.num_terms; ++k) {
    if (c.vars[k][0] == v) return false;
    for (int m = 1; m < c.lens[k]; ++m) {
      if (c.vars[k][m] != v) continue;
      const bool in_sum = shared_sum && m == c.lens[k] - 1;
      return in_sum ? t == 2 : true;
    }
  }
  return false;
}

static_assert(static_terms_disjoint(1) && static_terms_disjoint(2) && static_terms_disjoint(3),
              "the round-0 scaling reads each table exactly or scaled for all its uses");
enum ChallengeMode : int32_t { CHALLENGE_FIXED = 0, CHALLENGE_SHA3 = 1, CHALLENGE_EXPLICIT = 2 };
enum FillKind : int32_t { FILL_AFFINE = 0, FILL_UNIFORM = 1 };

// Composition sum_k c_k prod_v x_v^e_kv, read by the kernels with wave-uniform loads.
struct CompDesc {
  int32_t num_tables;
  int32_t num_terms;
  int32_t npts;  // degree + 1
  int32_t pad0;
  uint32_t exps[MAX_TERMS][MAX_TABLES / 8];  // 4-bit exponent of table v in term k
  int32_t kind[MAX_TERMS];
  uint32_t coeff[MAX_TERMS][MAX_LIMBS];  // Montgomery form, used when kind == COEFF_GENERAL
  // The same terms as factor lists: term k is the product of tables fac[k][0 .. len[k]).
  int32_t len[MAX_TERMS];
  uint8_t fac[MAX_TERMS][MAX_POINTS];
};

// Point-parallel evaluation: one lane per (pair, evaluated point). Round 0 evaluates all
// npts points; later rounds skip t = 1, whose value the round's finish derives from the claim.
constexpr int wide_points(int npts, int round) { return round > 0 ? npts - 1 : npts; }
constexpr uint32_t wide_pairs_per_block(int npts, int round) { return BLOCK / wide_points(npts, round); }

struct FieldConsts {
  uint32_t one[MAX_LIMBS];               // R mod p, the Montgomery image of 1
  uint32_t r2[MAX_LIMBS];                // R^2 mod p
  uint32_t word_pow[8][MAX_LIMBS];       // 2^(32i) * R^2 mod p: digest word i -> Montgomery
  uint32_t small[MAX_POINTS][MAX_LIMBS];  // Montgomery images of 0..7 (interpolation points)
  // Barycentric weights: bary[n-1][t] = 1 / prod_{s != t, s < n} (t - s), Montgomery, so
  // g(r) = sum_t g(t) bary[n-1][t] prod_{s != t} (r - s) for a polynomial known at 0..n-1.
  uint32_t bary[MAX_POINTS][MAX_POINTS][MAX_LIMBS];
  uint32_t pow32[MAX_POINTS][MAX_LIMBS];  // Montgomery images of 32^k, k < 8
  // g(d) of a degree-d polynomial from g(0..d-1) and its leading coefficient a (d >= 2):
  // sum_{t <= d} (-1)^(d-t) C(d,t) g(t) = d! a, so g(d) = d! a + sum_{t<d} lead[d][t] g(t) with
  // lead[d][t] = (-1)^(d-t+1) C(d,t) and lead[d][d] = d!, Montgomery.
  uint32_t lead[MAX_POINTS][MAX_POINTS][MAX_LIMBS];
};

// Per-proof device state. polys and r are written by the kernel that finishes a round.
struct ProofState {
  uint32_t counter;  // blocks of the current pass that have finished
  uint32_t pad[63];
  uint32_t r[MAX_ROUNDS][MAX_LIMBS];                     // challenge after round j, Montgomery
  uint32_t polys[MAX_ROUNDS][MAX_POINTS][MAX_LIMBS];     // g_j(t), Montgomery
  // Final-point step (after the last round): r[num_rounds-1] is the last challenge and
  // final_vals[v] = f_v(r_0, ..., r_{num_rounds-1}), Montgomery; final_ready = 1 once written.
  uint32_t final_ready;
  uint32_t pad2[63];
  uint32_t final_vals[MAX_TABLES][MAX_LIMBS];
  // Claim of the current round j >= 1 of a plain proof, g_{j-1}(r_{j-1}) (Montgomery), written
  // by the round's pass or point-parallel evaluation kernel for its finalize.
  uint32_t claim[MAX_LIMBS];
};

// Entries from one input table to the next in an input buffer the library allocates. Tables of
// 512 MiB or more (a multiple of 2^29 bytes apart) put the same-index entries of every table
// in the same HBM channel and bank: on MI355X a read of 5 such tables at one index streams at
// 5.0-5.5 TB/s against 6.1-6.2 TB/s with a 12 KiB gap after each table (a 128 MiB table is
// not helped, so smaller tables stay packed). Fold buffers are written and read in the 2:1
// read/write mix, which the gap does not change, and stay packed.
// Fold buffers' chunk (kernels.hpp, TableMap), log2 of its entries: 2^7 (a wave's 64 pairs of
// round j + 1 read two chunks of each table) for fields of up to 4 limbs, table-major for the
// 255-bit field. Same-box timings on MI355X, a*b*c + d*e pass kernels, chunked vs table-major:
// +1-2% at 32, 64 and 128 bits, -1.5% at 256 bits (whose folding rounds are issue-bound).
constexpr int32_t default_fold_chunk_shift(int limbs) { return limbs == 8 ? 63 : 7; }
constexpr uint64_t INPUT_PAD_BYTES = 12u << 10;
constexpr uint64_t INPUT_PAD_MIN_TABLE_BYTES = (uint64_t)1 << 29;
constexpr uint64_t input_stride(uint64_t n, int limbs) {
  return n * 4u * (uint64_t)limbs >= INPUT_PAD_MIN_TABLE_BYTES ? n + INPUT_PAD_BYTES / (4u * (uint64_t)limbs) : n;
}

struct PassArgs {
  const uint32_t* src;
  uint32_t* dst;
  uint64_t src_n;       // entries per table in src
  uint64_t src_stride;  // entries from one table of src to the next (src_n, or the input's stride)
  uint64_t pairs;  // pairs of round `round`
  int32_t round;
  int32_t num_rounds;
  int32_t fold;  // 0 for round 0 (pairs read straight from src), 1 otherwise
  int32_t challenge_mode;
  int32_t fused_finalize;  // the last block to finish reduces the partials and derives r
  int32_t chunk_shift;     // fold buffers' chunk, log2 of its entries (63: table-major)
  const CompDesc* comp;
  const FieldConsts* fc;
  ProofState* st;
  uint32_t* partials;         // [gridDim][MAX_POINTS][L]
  const uint32_t* explicit_r; // [num_rounds-1][L] canonical, CHALLENGE_EXPLICIT only
  // Block that only computes the round's claim into st->claim (the grid's last, one past the
  // blocks that evaluate; from round 1 on), or -1.
  int32_t claim_block = -1;
  // Tables whose round-0 entries are all 0 or 1 (generated selector tables): the 255-bit
  // point-parallel evaluation applies such a factor as a multiply by a small integer.
  uint32_t sel01 = 0;
};

// Rounds first_round..num_rounds-1 in one workgroup. Round j >= 1 reads buf(j-1) and writes
// buf(j), where buf(0) = input and buf(k) = (k odd ? buf_a : buf_b); n_j = n0 >> j.
struct TailArgs {
  const uint32_t* input;
  uint32_t* buf_a;
  uint32_t* buf_b;
  uint64_t n0;         // entries per table of the input (2^num_vars)
  uint64_t in_stride;  // entries from one input table to the next (n0 or input_stride(n0, L))
  int32_t first_round;
  int32_t num_rounds;
  int32_t challenge_mode;
  int32_t want_final;  // also derive the last challenge and fold the final point
  int32_t chunk_shift;  // fold buffers' chunk, log2 of its entries (63: table-major)
  int32_t pad0;
  const CompDesc* comp;
  const FieldConsts* fc;
  ProofState* st;
  const uint32_t* explicit_r;
};

struct FinalizeArgs {
  const uint32_t* partials;  // [nblocks][MAX_POINTS][L]
  uint32_t nblocks;
  int32_t round;
  int32_t num_rounds;
  int32_t challenge_mode;
  int32_t scale_pow;  // the partials sum to g / 32^scale_pow (static_round0_scale), else 0
  int32_t pad0;
  const CompDesc* comp;
  const FieldConsts* fc;
  ProofState* st;
  const uint32_t* explicit_r;
};

struct FillArgs {
  uint32_t* tables;
  uint64_t n;       // entries per table
  uint64_t stride;  // entries from one table to the next
  int32_t num_tables;
  int32_t kind;
  uint64_t seed;
  uint64_t selector_mask;
  uint32_t instance;
  uint32_t pad0;
  const FieldConsts* fc;
};

// ---------------------------------------------------------------------------------------------
// Batched proving: many independent instances advance round by round. One stage launch covers
// a list of entries (instance, its current round, first tile); each block works on one tile of
// one entry, so the composition it interprets is uniform across the block.
// ---------------------------------------------------------------------------------------------
struct InstDesc {
  const uint32_t* input;  // round-0 tables, element-major Montgomery
  uint32_t* buf_a;        // odd rounds' tables (n0/2 per table)
  uint32_t* buf_b;        // even rounds' tables (n0/4 per table)
  uint64_t n0;
  int32_t num_vars;
  int32_t challenge_mode;
  const CompDesc* comp;
  ProofState* st;
};

struct StageEntry {
  uint32_t inst;
  uint32_t round;
  uint32_t tile_begin;  // first tile (block) of this entry in the stage
  uint32_t tiles;
  uint32_t items;       // pairs per thread of this entry's tiles (per-pair: BLOCK * items pairs)
};

struct BatchArgs {
  const InstDesc* insts;
  const StageEntry* entries;  // sorted by tile_begin
  uint32_t n_entries;
  int32_t chunk_shift;        // fold buffers' chunk, log2 of its entries (63: table-major)
  uint32_t* partials;         // [tiles][MAX_POINTS][MAX_LIMBS]
  const FieldConsts* fc;
  uint32_t tile_base;         // point-parallel evaluation: block b is tile b + tile_base
  uint32_t pad0;
};

}  // namespace sc
