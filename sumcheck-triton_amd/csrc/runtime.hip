// Host runtime and C ABI of libsumcheck_hip.so (declared in include/sumcheck_hip.h).
//
// A context owns one HIP stream and the per-field constant blocks; a proof owns its device
// buffers (input tables unless borrowed, two fold scratch buffers, block partials, the
// composition descriptor and the round state). A run enqueues the whole round loop without
// host synchronization: the reference's run_once (cuda/src/field_sumcheck.cu:1012-1088)
// with the per-point D2H copies and host hashing replaced by device finalize kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <map>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/sumcheck_hip.h"
#include "calib.hpp"
#include "field.hpp"
#include "launch.hpp"
#include "layout.hpp"
#include "work.hpp"

using namespace sc;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

#define SC_HIP(expr)                                                                          \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) {                                                                   \
      return fail(_e == hipErrorOutOfMemory ? SC_ERR_MEMORY : SC_ERR_HIP,                     \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                         \
    }                                                                                         \
  } while (0)

int field_index(int bit_width) {
  switch (bit_width) {
    case 32: return 0;
    case 64: return 1;
    case 128: return 2;
    case 256: return 3;
    default: return -1;
  }
}

const Launchers& launchers_for(int idx) {
  switch (idx) {
    case 0: return launchers_p32();
    case 1: return launchers_p64();
    case 2: return launchers_p128();
    default: return launchers_p256();
  }
}

// Host-side constants and conversions, using the same templates as the device.
template <class F>
struct HostField {
  static constexpr int L = F::L;
  static Elem<L> from(const uint32_t* w) {
    Elem<L> x;
    for (int i = 0; i < L; ++i) x.v[i] = w[i];
    return x;
  }
  static bool less_than_p(const Elem<L>& x) {
    Elem<L> t;
    return sub_p<F>(t, x) != 0u;
  }
  static Elem<L> dbl_times(Elem<L> x, int k) {
    for (int i = 0; i < k; ++i) x = add<F>(x, x);
    return x;
  }
  static void consts(FieldConsts& fc) {
    std::memset(&fc, 0, sizeof(fc));
    Elem<L> one = zero<F>();
    one.v[0] = 1u;
    const Elem<L> rmod = dbl_times(one, 32 * L);   // R mod p
    const Elem<L> r2 = dbl_times(rmod, 32 * L);    // R^2 mod p
    for (int i = 0; i < L; ++i) {
      fc.one[i] = rmod.v[i];
      fc.r2[i] = r2.v[i];
    }
    for (int w = 0; w < 8; ++w) {
      const Elem<L> pw = dbl_times(r2, 32 * w);
      for (int i = 0; i < L; ++i) fc.word_pow[w][i] = pw.v[i];
    }
    // small[s] = s R mod p; bary[n-1][t] = (-1)^(n-1-t) / (t! (n-1-t)!) in Montgomery form
    Elem<L> mont_s[MAX_POINTS];
    mont_s[0] = zero<F>();
    for (int k = 1; k < MAX_POINTS; ++k) mont_s[k] = add<F>(mont_s[k - 1], rmod);
    for (int k = 0; k < MAX_POINTS; ++k)
      for (int i = 0; i < L; ++i) fc.small[k][i] = mont_s[k].v[i];
    Elem<L> fact[MAX_POINTS];
    fact[0] = rmod;
    for (int k = 1; k < MAX_POINTS; ++k) fact[k] = mont_mul<F>(fact[k - 1], mont_s[k]);
    for (int n = 1; n <= MAX_POINTS; ++n) {
      for (int t = 0; t < n; ++t) {
        Elem<L> w = mont_inverse(mont_mul<F>(fact[t], fact[n - 1 - t]), rmod);
        if ((n - 1 - t) & 1) w = sub<F>(zero<F>(), w);
        for (int i = 0; i < L; ++i) fc.bary[n - 1][t][i] = w.v[i];
      }
    }
    Elem<L> pw = rmod;
    for (int k = 0; k < MAX_POINTS; ++k) {
      for (int i = 0; i < L; ++i) fc.pow32[k][i] = pw.v[i];
      pw = dbl_times(pw, 5);
    }
    for (int d = 2; d < MAX_POINTS; ++d) {
      int64_t binom = 1;  // C(d, t)
      for (int t = 0; t <= d; ++t) {
        Elem<L> w = fact[d];  // t = d: d!
        if (t < d) {
          w = zero<F>();
          for (int64_t k = 0; k < binom; ++k) w = add<F>(w, rmod);
          if (((d - t + 1) & 1) != 0) w = sub<F>(zero<F>(), w);
        }
        for (int i = 0; i < L; ++i) fc.lead[d][t][i] = w.v[i];
        binom = binom * (d - t) / (t + 1);
      }
    }
  }
  // x^(p-2) in the Montgomery domain (the Montgomery image of x^-1 for x != 0).
  static Elem<L> mont_inverse(const Elem<L>& x, const Elem<L>& one_mont) {
    Elem<L> e;  // p - 2
    uint32_t borrow = 2;
    for (int i = 0; i < L; ++i) {
      const uint64_t d = (uint64_t)F::p(i) - borrow;
      e.v[i] = (uint32_t)d;
      borrow = (uint32_t)(d >> 63);
    }
    Elem<L> r = one_mont;
    for (int i = L - 1; i >= 0; --i) {
      for (int b = 31; b >= 0; --b) {
        r = mont_mul<F>(r, r);
        if ((e.v[i] >> b) & 1u) r = mont_mul<F>(r, x);
      }
    }
    return r;
  }
  static Elem<L> to_mont(const Elem<L>& x, const FieldConsts& fc) {
    return mont_mul<F>(x, from(fc.r2));
  }
  static bool is_minus_one(const Elem<L>& x) {
    Elem<L> one = zero<F>();
    one.v[0] = 1u;
    const Elem<L> s = add<F>(x, one);
    return is_zero<F>(s);
  }
};

template <class F>
int build_comp(const sc_composition* c, const FieldConsts& fc, CompDesc& d, int& degree) {
  using H = HostField<F>;
  constexpr int L = F::L;
  if (c == nullptr || c->term_lens == nullptr || c->term_offsets == nullptr || c->term_vars == nullptr)
    return fail(SC_ERR_ARG, "composition pointers must not be NULL");
  if (c->num_tables < 1 || c->num_tables > MAX_TABLES)
    return fail(SC_ERR_UNSUPPORTED, "num_tables must be in [1, 32]");
  if (c->num_terms < 1 || c->num_terms > MAX_TERMS)
    return fail(SC_ERR_UNSUPPORTED, "num_terms must be in [1, 32]");
  std::memset(&d, 0, sizeof(d));
  d.num_tables = c->num_tables;
  d.num_terms = c->num_terms;
  degree = 0;
  for (int k = 0; k < c->num_terms; ++k) {
    const int len = c->term_lens[k];
    if (len < 1) return fail(SC_ERR_ARG, "every term needs at least one factor");
    if (len > MAX_POINTS - 1) return fail(SC_ERR_UNSUPPORTED, "term degree must be <= 7");
    if (len > degree) degree = len;
    int counts[MAX_TABLES] = {0};
    for (int m = 0; m < len; ++m) {
      const int v = c->term_vars[c->term_offsets[k] + m];
      if (v < 0 || v >= c->num_tables) return fail(SC_ERR_ARG, "term_vars entry out of range");
      ++counts[v];
    }
    // factor list in table order (the product does not depend on the order)
    d.len[k] = len;
    int m = 0;
    for (int v = 0; v < c->num_tables; ++v)
      for (int e = 0; e < counts[v]; ++e) d.fac[k][m++] = (uint8_t)v;
    for (int v = 0; v < MAX_TABLES; ++v) d.exps[k][v >> 3] |= (uint32_t)counts[v] << ((v & 7) * 4);
    Elem<L> coeff = zero<F>();
    if (c->coeffs != nullptr) coeff = H::from(c->coeffs + (size_t)k * L);
    else coeff.v[0] = 1u;
    if (!H::less_than_p(coeff)) return fail(SC_ERR_ARG, "coefficient must be canonical (< p)");
    if (coeff.v[0] == 1u && [&] { for (int i = 1; i < L; ++i) if (coeff.v[i]) return false; return true; }()) {
      d.kind[k] = COEFF_ONE;
    } else if (H::is_minus_one(coeff)) {
      d.kind[k] = COEFF_MINUS_ONE;
    } else {
      d.kind[k] = COEFF_GENERAL;
    }
    const Elem<L> cm = work_coeff<F>(H::to_mont(coeff, fc));
    for (int i = 0; i < L; ++i) d.coeff[k][i] = cm.v[i];
  }
  d.npts = degree + 1;
  return SC_OK;
}

int build_comp_any(int fidx, const sc_composition* c, const FieldConsts& fc, CompDesc& d, int& degree) {
  switch (fidx) {
    case 0: return build_comp<FieldP32>(c, fc, d, degree);
    case 1: return build_comp<FieldP64>(c, fc, d, degree);
    case 2: return build_comp<FieldP128>(c, fc, d, degree);
    default: return build_comp<FieldP256>(c, fc, d, degree);
  }
}

void consts_any(int fidx, FieldConsts& fc) {
  switch (fidx) {
    case 0: HostField<FieldP32>::consts(fc); break;
    case 1: HostField<FieldP64>::consts(fc); break;
    case 2: HostField<FieldP128>::consts(fc); break;
    default: HostField<FieldP256>::consts(fc); break;
  }
}

// Compiled composition matching d exactly (same tables, terms in order, factor counts and
// coefficient kinds), or 0.
int match_static(const CompDesc& d) {
  for (int id = 1; id < NUM_STATIC_COMPS; ++id) {
    const StaticComp& c = STATIC_COMPS[id];
    if (c.num_tables != d.num_tables || c.num_terms != d.num_terms) continue;
    bool ok = true;
    for (int k = 0; k < c.num_terms && ok; ++k) {
      uint32_t exps[MAX_TABLES / 8] = {0};
      for (int m = 0; m < c.lens[k]; ++m) {
        const int v = c.vars[k][m];
        exps[v >> 3] += 1u << ((v & 7) * 4);
      }
      for (int w = 0; w < MAX_TABLES / 8; ++w) ok = ok && exps[w] == d.exps[k][w];
      ok = ok && c.kinds[k] == d.kind[k];
    }
    if (ok) return id;
  }
  return 0;
}

int pick_tmax(int T) {
  for (int t : TMAX_CHOICES)
    if (T <= t) return t;
  return -1;
}

}  // namespace

struct sc_context {
  int device = 0;
  hipStream_t stream = nullptr;
  FieldConsts host_fc[4];
  FieldConsts* d_fc[4] = {nullptr, nullptr, nullptr, nullptr};
};

struct sc_proof {
  sc_context* ctx = nullptr;
  int fidx = 0, L = 0, T = 0, npts = 0, num_vars = 0, tmax = 0, pmax = 0;
  int sid = 0;             // compiled composition in use (0 = interpreted)
  int matched_sid = 0;     // compiled composition that matches, if any
  const Launchers* ln = nullptr;
  uint64_t n0 = 0;
  uint32_t* d_tables = nullptr;
  uint64_t in_stride = 0;  // entries from one input table to the next (input_stride, or n0 for caller tables)
  uint32_t sel01 = 0;      // generated selector tables (entries 0 or 1), PassArgs::sel01 of round 0
  bool owns_tables = false;
  uint32_t* d_a = nullptr;
  uint32_t* d_b = nullptr;
  uint32_t* d_partials = nullptr;
  uint32_t* d_explicit = nullptr;
  CompDesc* d_comp = nullptr;
  ProofState* d_state = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int max_pass_blocks = 2048;
  int64_t tail_pairs = 0;
  int32_t fold_chunk_shift = 7;  // fold buffers' chunk (log2 entries; init_shape sets the default)
  int64_t wide_pairs = 0;      // rounds with at most this many pairs run point-parallel
  bool default_wide = true;    // wide_pairs follows the kernel choice (no caller-set value)
  bool fused_finalize = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_mode = -1;
  int last_mode = -1;  // challenge mode of the last run (the final point needs fixed or sha3)
  uint64_t device_bytes = 0;
  bool profiling = false;
  bool profiled_last = false;
  hipEvent_t prof_ev[2 * (MAX_ROUNDS + 1)] = {};
  hipEvent_t prof_fin[MAX_ROUNDS] = {};  // after round j's finalize
  // sc_proof_run_async: two pinned host slots for the challenges and round polynomials of a
  // run (ProofState from r up to the proof's last polynomial row), each with its copy's event
  uint32_t* h_slot[2] = {nullptr, nullptr};
  hipEvent_t ev_slot[2] = {nullptr, nullptr};
  uint64_t async_runs = 0;
};

namespace {

// Pass-kernel grid caps from sweeps on MI355X (a*b*c + d*e at the benchmark sizes; fewer
// blocks also means fewer partials for the finalize kernel to sum). The point-parallel kernel
// holds few registers per lane and always takes the largest grid (BLS12-381 constraint
// workload: 22.9 ms at 512 blocks, 19.4 ms at 4096).
int default_pass_blocks(int L) {
  switch (L) {
    case 1: return 768;   // num_vars 26 sweep: 1.245 ms at 768, 1.27 at 512, 1.28 at 1024
    case 2: return 512;   // num_vars 26: 2.21 ms at 512, 2.25 at 768, 2.29 at 1024, 3.10 at 256
    case 4: return 512;   // num_vars 24: 1.38 ms at 512 (1.37 at 256), 1.40 at 768, 1.42 at 1024
    default: return 512;
  }
}

// Pairs below which the remaining rounds run in the single-workgroup tail kernel (batch and
// server instances). Compositions of more than 8 tables skip it: their per-pair work is large
// enough that one workgroup is slower than the fold + point-parallel launches even for a
// handful of pairs.
int64_t default_tail_pairs(int L, bool over8) {
  if (over8) return 1;
  return L == 8 ? 64 : 512;
}

// The same for a single proof: at 255 bits the tail from 256 pairs measured best (a*b*c + d*e,
// MI355X: num_vars 24 4.16 -> 4.10 ms, num_vars 20 1.04 -> 1.01 ms against 64 pairs), and at
// 127 bits too (the autotuner's pick at num_vars 24).
int64_t default_proof_tail_pairs(int L, bool over8) {
  if (over8) return 1;
  return L >= 4 ? 256 : 512;
}

// Rounds with at most this many pairs run as a fold launch plus the point-parallel evaluation
// instead of the per-pair pass kernel. Interpreted compositions of more than 8 tables use it
// in every round. Interpreted 255-bit compositions use it below 2^20 pairs: their per-pair
// kernel runs one wave per SIMD (2 x 8 working values and 8 accumulators of 9 limbs), which
// the point-parallel split beats by 8-30% at num_vars 16-20 (MI355X); above that the per-pair
// kernel's fused fold saves a write and a read of every table and wins for most compositions.
// Compiled compositions keep the per-pair kernels for their large rounds (a*b*c + d*e, 255-bit,
// num_vars 24: 4.3 ms against 6.2 ms all point-parallel); at 255 bits their rounds below 2^18
// pairs run point-parallel (num_vars 20: 1.01 -> 0.97 ms; num_vars 24: 4.10 -> 4.08 ms).
int64_t default_wide_pairs(const sc_proof* p) {
  if (p->sid != 0) return p->L == 8 ? (int64_t(1) << 17) : 0;
  if (p->tmax > 8) return INT64_MAX;
  if (p->L != 8) return 0;
  // 255-bit interpreted compositions of degree >= 3: every round (num_vars 24, MI355X: 5.94
  // against 6.10 ms for a*b*c + d*e forced interpreted, 15.1 against 18.2 ms for a degree-5
  // product sum); degree 2 keeps the per-pair kernel for rounds of 2^20 pairs and more
  // (8 tables: 5.49 ms against 6.23 all point-parallel).
  return p->npts >= 4 ? INT64_MAX : (int64_t(1) << 19);
}

// Batches and servers, per instance: every round point-parallel where the per-pair kernels
// would spill (the launch's table capacity above 8), and at 255 bits for compositions of
// degree >= 3, whose per-pair kernels run one wave per SIMD (256 instances of num_vars 10-22,
// MI355X: 81.9 ms all point-parallel against 100.8 ms per-pair); otherwise per-pair (64 bits:
// 12.8 ms per-pair against 15.5 ms all point-parallel).
int64_t default_instance_wide_pairs(int L, int tmax, int npts) {
  return (tmax > 8 || (L == 8 && npts >= 4)) ? INT64_MAX : 0;
}

// Work lists of one batched stage (one round of each listed instance). Per-pair entries own
// tiles [0, pass_tiles) of the stage's partials and point-parallel entries the next
// eval_tiles; fold entries (point-parallel rounds >= 1) number their tiles apart. On the
// device the lists sit as [pass][eval][fold], so one finalize launch covers [pass][eval].
struct StagePlan {
  std::vector<StageEntry> pass, eval, fold;
  uint64_t pass_tiles = 0, eval_tiles = 0, fold_tiles = 0;

  // Partial tiles the round adds: pass tiles of BLOCK * items pairs, or evaluation tiles of
  // wide_pairs_per_block(npts, round) * items pairs.
  static uint64_t tiles(uint64_t pairs, bool wide, uint32_t items, int npts, uint32_t round) {
    const uint64_t span = (wide ? (uint64_t)wide_pairs_per_block(npts, (int)round) : (uint64_t)BLOCK) * items;
    return (pairs + span - 1) / span;
  }
  void add(uint32_t inst, uint32_t round, uint64_t pairs, int num_tables, bool wide, uint32_t items, int npts) {
    const uint32_t t = (uint32_t)tiles(pairs, wide, items, npts, round);
    if (!wide) {
      pass.push_back(StageEntry{inst, round, (uint32_t)pass_tiles, t, items});
      pass_tiles += t;
      return;
    }
    eval.push_back(StageEntry{inst, round, (uint32_t)eval_tiles, t, items});  // rebased in layout()
    eval_tiles += t;
    if (round > 0) {
      const uint64_t elems = 2 * pairs * (uint64_t)num_tables, span = (uint64_t)BLOCK * items;
      const uint32_t f = (uint32_t)((elems + span - 1) / span);
      fold.push_back(StageEntry{inst, round, (uint32_t)fold_tiles, f, items});
      fold_tiles += f;
    }
  }
  uint64_t partial_tiles() const { return pass_tiles + eval_tiles; }
  // Appends [pass][eval][fold] to out, evaluation tiles after the pass tiles.
  void layout(std::vector<StageEntry>& out) const {
    out.insert(out.end(), pass.begin(), pass.end());
    for (StageEntry e : eval) {
      e.tile_begin += (uint32_t)pass_tiles;
      out.push_back(e);
    }
    out.insert(out.end(), fold.begin(), fold.end());
  }
  size_t size() const { return pass.size() + eval.size() + fold.size(); }
};

// Launches one stage laid out by StagePlan::layout at d (device), then its finalize.
int enqueue_stage(const Launchers& ln, hipStream_t s, BatchArgs a, int tmax, int pmax, const StageEntry* d,
                  uint32_t n_pass, uint32_t n_eval, uint32_t n_fold, uint64_t pass_tiles, uint64_t eval_tiles,
                  uint64_t fold_tiles) {
  a.tile_base = 0;
  a.pad0 = 0;
  if (n_pass) {
    a.entries = d;
    a.n_entries = n_pass;
    SC_HIP(ln.batch_pass(s, tmax, pmax, a, (unsigned)pass_tiles));
  }
  if (n_fold) {
    a.entries = d + n_pass + n_eval;
    a.n_entries = n_fold;
    SC_HIP(ln.batch_wide_fold(s, a, (unsigned)fold_tiles));
  }
  if (n_eval) {
    a.entries = d + n_pass;
    a.n_entries = n_eval;
    a.tile_base = (uint32_t)pass_tiles;
    SC_HIP(ln.batch_wide_eval(s, tmax, pmax, a, (unsigned)eval_tiles));
    a.tile_base = 0;
  }
  if (n_pass + n_eval) {
    a.entries = d;
    a.n_entries = n_pass + n_eval;
    SC_HIP(ln.batch_finalize(s, pmax, a, n_pass + n_eval));
  }
  return SC_OK;
}

// A captured proof graph bakes in the schedule; changing it discards the graph.
void drop_graph(sc_proof* p) {
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->graph_exec = nullptr;
  p->graph = nullptr;
  p->graph_mode = -1;
}

// First round that runs in the tail kernel.
int tail_first_round(const sc_proof* p) {
  int j = 0;
  while (j < p->num_vars && (int64_t)(p->n0 >> (j + 1)) > p->tail_pairs) ++j;
  return j;
}

uint32_t* buf_for_round(const sc_proof* p, int j) {
  if (j == 0) return p->d_tables;
  return (j & 1) ? p->d_a : p->d_b;
}

// The pass kernel's round 0 of a compiled 255-bit composition leaves its sum scaled by
// 32^-static_round0_scale (layout.hpp); the finalize restores it.
int round0_scale_pow(const sc_proof* p, int j, bool wide) {
  return (j == 0 && !wide && p->L == 8 && p->sid != 0) ? static_round0_scale(p->sid) : 0;
}

int enqueue_rounds(sc_proof* p, int mode) {
  hipStream_t s = p->ctx->stream;
  const FieldConsts* fc = p->ctx->d_fc[p->fidx];
  const int jt = tail_first_round(p);
  for (int j = 0; j < jt; ++j) {
    PassArgs a;
    a.src = buf_for_round(p, j == 0 ? 0 : j - 1);
    a.dst = j == 0 ? nullptr : buf_for_round(p, j);
    a.src_n = j == 0 ? p->n0 : (p->n0 >> (j - 1));
    a.src_stride = j <= 1 ? p->in_stride : a.src_n;  // rounds 0 and 1 read the input
    a.pairs = p->n0 >> (j + 1);
    a.round = j;
    a.num_rounds = p->num_vars;
    a.fold = j > 0;
    a.challenge_mode = mode;
    a.comp = p->d_comp;
    a.fc = fc;
    a.st = p->d_state;
    a.partials = p->d_partials;
    a.explicit_r = p->d_explicit;
    a.fused_finalize = p->fused_finalize ? 1 : 0;
    a.chunk_shift = p->fold_chunk_shift;
    uint64_t g = (a.pairs + BLOCK - 1) / BLOCK;
    if (g > (uint64_t)p->max_pass_blocks) g = (uint64_t)p->max_pass_blocks;
    const bool wide = (int64_t)a.pairs <= p->wide_pairs;
    if (p->profiling) SC_HIP(hipEventRecord(p->prof_ev[2 * j], s));
    if (wide) {
      if (j > 0) {
        uint64_t gf = ((a.src_n >> 1) * (uint64_t)p->T + BLOCK - 1) / BLOCK;
        SC_HIP(p->ln->wide_fold(s, a, (unsigned)std::min<uint64_t>(gf, 8192)));
      }
      PassArgs e = a;
      e.src = buf_for_round(p, j);
      e.dst = nullptr;
      e.src_n = p->n0 >> j;
      e.src_stride = j == 0 ? p->in_stride : e.src_n;
      e.fold = 0;
      e.fused_finalize = 0;
      e.sel01 = j == 0 ? p->sel01 : 0u;
      const uint64_t pairs_per_block = wide_pairs_per_block(p->npts, j);
      g = std::min<uint64_t>((a.pairs + pairs_per_block - 1) / pairs_per_block, (uint64_t)MAX_PASS_BLOCKS);
      // from round 1 one more block computes the round's claim for the finalize
      e.claim_block = j > 0 ? (int32_t)g : -1;
      SC_HIP(p->ln->wide_eval(s, p->tmax, p->pmax, e, (unsigned)(g + (j > 0 ? 1 : 0))));
    } else {
      a.claim_block = j > 0 ? (int32_t)g : -1;
      SC_HIP(p->ln->pass(s, p->sid, p->tmax, p->pmax, a, (unsigned)(g + (a.claim_block >= 0 ? 1 : 0))));
    }
    if (p->profiling) SC_HIP(hipEventRecord(p->prof_ev[2 * j + 1], s));
    if (p->fused_finalize && !wide) {
      if (p->profiling) SC_HIP(hipEventRecord(p->prof_fin[j], s));
      continue;
    }
    FinalizeArgs f;
    f.partials = p->d_partials;
    f.nblocks = (uint32_t)g;
    f.round = j;
    f.num_rounds = p->num_vars;
    f.challenge_mode = mode;
    f.comp = p->d_comp;
    f.fc = fc;
    f.st = p->d_state;
    f.explicit_r = p->d_explicit;
    f.scale_pow = round0_scale_pow(p, j, wide);
    f.pad0 = 0;
    SC_HIP(p->ln->finalize(s, p->pmax, f));
    if (p->profiling) SC_HIP(hipEventRecord(p->prof_fin[j], s));
  }
  if (jt < p->num_vars) {
    TailArgs t;
    t.input = p->d_tables;
    t.buf_a = p->d_a;
    t.buf_b = p->d_b;
    t.n0 = p->n0;
    t.in_stride = p->in_stride;
    t.chunk_shift = p->fold_chunk_shift;
    t.pad0 = 0;
    t.first_round = jt;
    t.num_rounds = p->num_vars;
    t.challenge_mode = mode;
    t.want_final = mode != SC_CHALLENGE_EXPLICIT ? 1 : 0;
    t.comp = p->d_comp;
    t.fc = fc;
    t.st = p->d_state;
    t.explicit_r = p->d_explicit;
    if (p->profiling) SC_HIP(hipEventRecord(p->prof_ev[2 * MAX_ROUNDS], s));
    SC_HIP(p->ln->tail(s, p->sid, p->tmax, p->pmax, t));
    if (p->profiling) SC_HIP(hipEventRecord(p->prof_ev[2 * MAX_ROUNDS + 1], s));
  }
  return SC_OK;
}

int check_mode(const sc_proof* p, int mode, const uint32_t* explicit_r) {
  if (mode != SC_CHALLENGE_FIXED && mode != SC_CHALLENGE_SHA3 && mode != SC_CHALLENGE_EXPLICIT)
    return fail(SC_ERR_ARG, "challenge_mode must be 0 (fixed), 1 (sha3) or 2 (explicit)");
  if (mode == SC_CHALLENGE_EXPLICIT && p->num_vars > 1 && explicit_r == nullptr)
    return fail(SC_ERR_ARG, "explicit challenges required");
  return SC_OK;
}

int upload_explicit(sc_proof* p, int mode, const uint32_t* explicit_r) {
  if (mode != SC_CHALLENGE_EXPLICIT || p->num_vars < 2) return SC_OK;
  // Validate canonical range on the host.
  for (int j = 0; j < p->num_vars - 1; ++j) {
    const uint32_t* w = explicit_r + (size_t)j * p->L;
    bool ok;
    switch (p->fidx) {
      case 0: ok = HostField<FieldP32>::less_than_p(HostField<FieldP32>::from(w)); break;
      case 1: ok = HostField<FieldP64>::less_than_p(HostField<FieldP64>::from(w)); break;
      case 2: ok = HostField<FieldP128>::less_than_p(HostField<FieldP128>::from(w)); break;
      default: ok = HostField<FieldP256>::less_than_p(HostField<FieldP256>::from(w)); break;
    }
    if (!ok) return fail(SC_ERR_ARG, "explicit challenge must be canonical (< p)");
  }
  SC_HIP(hipMemcpyAsync(p->d_explicit, explicit_r, sizeof(uint32_t) * p->L * (p->num_vars - 1),
                        hipMemcpyHostToDevice, p->ctx->stream));
  SC_HIP(hipStreamSynchronize(p->ctx->stream));
  return SC_OK;
}

// Shape of a proof and its default schedule (library defaults, as measured above).
void init_shape(sc_proof* p, sc_context* ctx, int fidx, const CompDesc& desc, int num_vars) {
  p->ctx = ctx;
  p->fidx = fidx;
  p->ln = &launchers_for(fidx);
  p->L = p->ln->limbs;
  p->T = desc.num_tables;
  p->npts = desc.npts;
  p->num_vars = num_vars;
  p->tmax = pick_tmax(p->T);
  p->pmax = p->npts <= 4 ? 4 : 8;
  p->n0 = (uint64_t)1 << num_vars;
  p->tail_pairs = default_proof_tail_pairs(p->L, p->tmax > 8);
  p->max_pass_blocks = default_pass_blocks(p->L);
  p->fold_chunk_shift = default_fold_chunk_shift(p->L);
  p->matched_sid = match_static(desc);
  p->sid = p->matched_sid;
  p->wide_pairs = default_wide_pairs(p);
}

void free_proof(sc_proof* p) {
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->owns_tables && p->d_tables) (void)hipFree(p->d_tables);
  if (p->d_a) (void)hipFree(p->d_a);
  if (p->d_b) (void)hipFree(p->d_b);
  if (p->d_partials) (void)hipFree(p->d_partials);
  if (p->d_explicit) (void)hipFree(p->d_explicit);
  if (p->d_comp) (void)hipFree(p->d_comp);
  if (p->d_state) (void)hipFree(p->d_state);
  if (p->ev0) (void)hipEventDestroy(p->ev0);
  if (p->ev1) (void)hipEventDestroy(p->ev1);
  for (hipEvent_t e : p->prof_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : p->prof_fin)
    if (e) (void)hipEventDestroy(e);
  for (int k = 0; k < 2; ++k) {
    if (p->h_slot[k]) (void)hipHostFree(p->h_slot[k]);
    if (p->ev_slot[k]) (void)hipEventDestroy(p->ev_slot[k]);
  }
  delete p;
}

// ---------------------------------------------------------------------------------------------
// Device memory of sc_prove_noalloc, the entry point for callers that cannot allocate: round
// 1's and round 2's tables (0.75x the input at most), then the block partials, the
// composition descriptor and the round state, in one fixed array per GPU.
// ---------------------------------------------------------------------------------------------
constexpr uint64_t NOALLOC_FOLD_WORDS = (uint64_t)SC_NOALLOC_FOLD_BYTES / 4;
constexpr uint64_t NOALLOC_PARTIAL_WORDS = (uint64_t)MAX_PASS_BLOCKS * MAX_POINTS * MAX_LIMBS;
constexpr uint64_t round_up256(uint64_t b) { return (b + 255) & ~uint64_t(255); }
constexpr uint64_t NOALLOC_OFF_PARTIALS = NOALLOC_FOLD_WORDS * 4;
constexpr uint64_t NOALLOC_OFF_COMP = NOALLOC_OFF_PARTIALS + NOALLOC_PARTIAL_WORDS * 4;
constexpr uint64_t NOALLOC_OFF_STATE = NOALLOC_OFF_COMP + round_up256(sizeof(CompDesc));
constexpr uint64_t NOALLOC_BYTES = NOALLOC_OFF_STATE + round_up256(sizeof(ProofState));

}  // namespace

// In host code `scratch` is a host-side stand-in, not device memory: host code uses only
// the per-device address that hipGetSymbolAddress returns (g_scratch), never this name.
static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch[MAX_DEVICES] = {};
std::string g_scratch_error[MAX_DEVICES];

// Called by sc_context_create with `device` current: resolves this GPU's copy of scratch once.
// A failure does not fail the context; sc_prove_noalloc then reports it.
void resolve_scratch(int device) {
  if (device < 0 || device >= MAX_DEVICES) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device] != nullptr) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) {
    g_scratch_error[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return;
  }
  g_scratch[device] = static_cast<uint8_t*>(ptr);
}

// Writes the composition descriptor (passed by value as the kernel argument, so nothing is
// copied from host memory) and clears the round state with vector stores.
__global__ void __launch_bounds__(BLOCK) noalloc_setup_kernel(CompDesc desc, CompDesc* comp, ProofState* st) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&desc);
  uint32_t* dc = reinterpret_cast<uint32_t*>(comp);
  for (uint32_t k = threadIdx.x; k < sizeof(CompDesc) / 4; k += BLOCK) dc[k] = src[k];
  uint32_t* ds = reinterpret_cast<uint32_t*>(st);
  for (uint32_t k = threadIdx.x; k < sizeof(ProofState) / 4; k += BLOCK) ds[k] = 0u;
}

// The proof's outputs from its round state: round_polys [num_vars][npts][L], challenges
// [num_vars-1][L] (may be NULL) and checksum sum_j sum_t sum_i (i + 1) limb_i (may be NULL),
// as sc_proof_result forms them (cuda/src/field_sumcheck.cu:1043-1049).
__global__ void __launch_bounds__(BLOCK) noalloc_export_kernel(const ProofState* st, int num_vars, int npts, int L,
                                                              uint32_t* polys, uint32_t* chal, uint64_t* checksum) {
  __shared__ uint64_t s_sum[BLOCK];
  const int per_round = npts * L;
  uint64_t sum = 0;
  for (int k = threadIdx.x; k < num_vars * per_round; k += BLOCK) {
    const int j = k / per_round, t = (k % per_round) / L, i = k % L;
    const uint32_t w = st->polys[j][t][i];
    polys[k] = w;
    sum += (uint64_t)(i + 1) * w;
  }
  if (chal != nullptr) {
    for (int k = threadIdx.x; k < (num_vars - 1) * L; k += BLOCK) chal[k] = st->r[k / L][k % L];
  }
  s_sum[threadIdx.x] = sum;
  __syncthreads();
  for (int half = BLOCK / 2; half > 0; half >>= 1) {
    if (threadIdx.x < half) s_sum[threadIdx.x] += s_sum[threadIdx.x + half];
    __syncthreads();
  }
  if (checksum != nullptr && threadIdx.x == 0) *checksum = s_sum[0];
}

}  // namespace

extern "C" {

const char* sc_last_error(void) { return g_error.c_str(); }

const char* sc_version(void) { return "sumcheck_hip 0.1 gfx950"; }

int sc_context_create(int device, sc_context** out) {
  if (out == nullptr) return fail(SC_ERR_ARG, "out must not be NULL");
  *out = nullptr;
  int count = 0;
  SC_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(SC_ERR_ARG, "no such device");
  SC_HIP(hipSetDevice(device));
  sc_context* c = new (std::nothrow) sc_context();
  if (c == nullptr) return fail(SC_ERR_MEMORY, "host allocation failed");
  c->device = device;
  resolve_scratch(device);
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(SC_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  for (int f = 0; f < 4; ++f) {
    consts_any(f, c->host_fc[f]);
    e = hipMalloc(&c->d_fc[f], sizeof(FieldConsts));
    if (e == hipSuccess) e = hipMemcpy(c->d_fc[f], &c->host_fc[f], sizeof(FieldConsts), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      sc_context_destroy(c);
      return fail(SC_ERR_HIP, std::string("field constants: ") + hipGetErrorString(e));
    }
  }
  *out = c;
  return SC_OK;
}

int sc_context_destroy(sc_context* ctx) {
  if (ctx == nullptr) return SC_OK;
  (void)hipSetDevice(ctx->device);
  for (int f = 0; f < 4; ++f)
    if (ctx->d_fc[f]) (void)hipFree(ctx->d_fc[f]);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SC_OK;
}

int sc_context_sync(sc_context* ctx) {
  if (ctx == nullptr) return fail(SC_ERR_ARG, "ctx must not be NULL");
  SC_HIP(hipStreamSynchronize(ctx->stream));
  return SC_OK;
}

int sc_proof_create(sc_context* ctx, int32_t bit_width, const sc_composition* comp, int32_t num_vars,
                    const sc_input* input, sc_proof** out) {
  if (ctx == nullptr || out == nullptr || input == nullptr)
    return fail(SC_ERR_ARG, "ctx, input and out must not be NULL");
  *out = nullptr;
  const int fidx = field_index(bit_width);
  if (fidx < 0) return fail(SC_ERR_ARG, "bit_width must be 32, 64, 128 or 256");
  if (num_vars < 1 || num_vars > MAX_ROUNDS) return fail(SC_ERR_UNSUPPORTED, "num_vars must be in [1, 40]");
  CompDesc desc;
  int degree = 0;
  int rc = build_comp_any(fidx, comp, ctx->host_fc[fidx], desc, degree);
  if (rc != SC_OK) return rc;
  if (input->kind < SC_INPUT_HOST_CANONICAL || input->kind > SC_INPUT_DEVICE_MONT_LIMB)
    return fail(SC_ERR_ARG, "unknown input kind");
  const bool limb_major = input->kind == SC_INPUT_HOST_MONT_LIMB || input->kind == SC_INPUT_DEVICE_MONT_LIMB;
  if ((input->kind <= SC_INPUT_DEVICE_MONT || limb_major) && input->data == nullptr)
    return fail(SC_ERR_ARG, "input data must not be NULL");

  SC_HIP(hipSetDevice(ctx->device));
  sc_proof* p = new (std::nothrow) sc_proof();
  if (p == nullptr) return fail(SC_ERR_MEMORY, "host allocation failed");
  init_shape(p, ctx, fidx, desc, num_vars);

  const uint64_t words = (uint64_t)p->T * p->n0 * p->L;
  auto cleanup = [&](int code) {
    free_proof(p);
    return code;
  };
  auto alloc = [&](void** ptr, uint64_t bytes) -> hipError_t {
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
    if (e == hipSuccess) p->device_bytes += bytes;
    return e;
  };
#define SC_HIP_P(expr)                                                                        \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) {                                                                   \
      fail(_e == hipErrorOutOfMemory ? SC_ERR_MEMORY : SC_ERR_HIP,                            \
           std::string(#expr) + ": " + hipGetErrorString(_e));                                \
      return cleanup(_e == hipErrorOutOfMemory ? SC_ERR_MEMORY : SC_ERR_HIP);                 \
    }                                                                                         \
  } while (0)

  if (input->kind == SC_INPUT_DEVICE_MONT) {
    p->d_tables = (uint32_t*)input->data;
    p->in_stride = p->n0;
    p->owns_tables = false;
  } else {
    p->in_stride = input_stride(p->n0, p->L);
    SC_HIP_P(alloc((void**)&p->d_tables, ((uint64_t)(p->T - 1) * p->in_stride + p->n0) * p->L * 4));
    p->owns_tables = true;
  }
  const uint64_t table_bytes = p->n0 * p->L * 4, stride_bytes = p->in_stride * p->L * 4;
  SC_HIP_P(alloc((void**)&p->d_a, (words / 2) * 4));
  SC_HIP_P(alloc((void**)&p->d_b, (words / 4) * 4));
  SC_HIP_P(alloc((void**)&p->d_partials, (uint64_t)MAX_PASS_BLOCKS * MAX_POINTS * p->L * 4));
  SC_HIP_P(alloc((void**)&p->d_explicit, (uint64_t)MAX_ROUNDS * MAX_LIMBS * 4));
  SC_HIP_P(alloc((void**)&p->d_comp, sizeof(CompDesc)));
  SC_HIP_P(alloc((void**)&p->d_state, sizeof(ProofState)));
  SC_HIP_P(hipEventCreate(&p->ev0));
  SC_HIP_P(hipEventCreate(&p->ev1));
  hipStream_t s = ctx->stream;
  SC_HIP_P(hipMemcpyAsync(p->d_comp, &desc, sizeof(desc), hipMemcpyHostToDevice, s));
  SC_HIP_P(hipMemsetAsync(p->d_state, 0, sizeof(ProofState), s));

  const FieldConsts* fc = ctx->d_fc[fidx];
  if (input->kind == SC_INPUT_HOST_CANONICAL || input->kind == SC_INPUT_HOST_MONT) {
    for (int v = 0; v < p->T; ++v) {  // one copy per table (a 2D copy's pitch is limited)
      SC_HIP_P(hipMemcpyAsync((uint8_t*)p->d_tables + v * stride_bytes, (const uint8_t*)input->data + v * table_bytes,
                              table_bytes, hipMemcpyHostToDevice, s));
    }
    if (input->kind == SC_INPUT_HOST_CANONICAL) {
      // every table in place; the gaps between tables (if any) are encoded too, unread
      const uint64_t span = (uint64_t)(p->T - 1) * p->in_stride + p->n0;
      SC_HIP_P(p->ln->field_op(s, 3, span, p->d_tables, nullptr, p->d_tables, fc));
    }
  } else if (input->kind == SC_INPUT_DEVICE_MONT_LIMB) {
    SC_HIP_P(p->ln->limb_to_element(s, (const uint32_t*)input->data, p->d_tables, p->n0, p->in_stride, p->T));
  } else if (input->kind == SC_INPUT_HOST_MONT_LIMB) {
    // staged through a temporary device buffer of the input's size, freed after the transpose
    uint32_t* tmp = nullptr;
    SC_HIP_P(hipMalloc(&tmp, words * 4));
    hipError_t e = hipMemcpyAsync(tmp, input->data, words * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = p->ln->limb_to_element(s, tmp, p->d_tables, p->n0, p->in_stride, p->T);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(tmp);
    SC_HIP_P(e);
  } else if (input->kind == SC_INPUT_AFFINE || input->kind == SC_INPUT_UNIFORM) {
    FillArgs f;
    f.tables = p->d_tables;
    f.n = p->n0;
    f.stride = p->in_stride;
    f.num_tables = p->T;
    f.kind = input->kind == SC_INPUT_AFFINE ? FILL_AFFINE : FILL_UNIFORM;
    f.seed = input->seed;
    f.selector_mask = input->selector_mask;
    f.instance = input->instance;
    f.pad0 = 0;
    f.fc = fc;
    SC_HIP_P(p->ln->fill(s, f));
    if (input->kind == SC_INPUT_UNIFORM)  // fill_kernel writes 0 or 1 to these tables
      p->sel01 = (uint32_t)(input->selector_mask & (p->T >= 32 ? ~0ull : (1ull << p->T) - 1));
  }
  SC_HIP_P(hipStreamSynchronize(s));
#undef SC_HIP_P
  *out = p;
  return SC_OK;
}

int sc_proof_set_schedule(sc_proof* p, int32_t max_pass_blocks, int64_t tail_pairs, int32_t fused_finalize) {
  if (p == nullptr) return fail(SC_ERR_ARG, "proof must not be NULL");
  if (fused_finalize < 0 || fused_finalize > 1) return fail(SC_ERR_ARG, "fused_finalize must be 0 or 1");
  if (max_pass_blocks < 0 || max_pass_blocks > MAX_PASS_BLOCKS)
    return fail(SC_ERR_ARG, "max_pass_blocks must be in [0, 4096]");
  if (tail_pairs < 0) return fail(SC_ERR_ARG, "tail_pairs must be >= 0");
  p->fused_finalize = fused_finalize != 0;
  p->max_pass_blocks = max_pass_blocks ? max_pass_blocks : default_pass_blocks(p->L);
  p->tail_pairs = tail_pairs ? tail_pairs : default_proof_tail_pairs(p->L, p->tmax > 8);
  drop_graph(p);
  return SC_OK;
}

int sc_proof_set_point_parallel(sc_proof* p, int64_t max_pairs) {
  if (p == nullptr) return fail(SC_ERR_ARG, "proof must not be NULL");
  p->default_wide = max_pairs < 0;
  p->wide_pairs = max_pairs < 0 ? default_wide_pairs(p) : max_pairs;
  drop_graph(p);
  return SC_OK;
}

int sc_proof_set_fold_chunk(sc_proof* p, int32_t chunk_log2) {
  if (p == nullptr) return fail(SC_ERR_ARG, "proof must not be NULL");
  // a folding round reads 4 consecutive entries of a table, so a chunk holds at least 4
  if (chunk_log2 < 0 || chunk_log2 == 1 || chunk_log2 > 63)
    return fail(SC_ERR_ARG, "chunk_log2 must be 0 (default) or in [2, 63]");
  p->fold_chunk_shift = chunk_log2 == 0 ? default_fold_chunk_shift(p->L) : chunk_log2;
  drop_graph(p);
  return SC_OK;
}

int sc_proof_run(sc_proof* p, int32_t mode, const uint32_t* explicit_r) {
  if (p == nullptr) return fail(SC_ERR_ARG, "proof must not be NULL");
  int rc = check_mode(p, mode, explicit_r);
  if (rc != SC_OK) return rc;
  SC_HIP(hipSetDevice(p->ctx->device));
  rc = upload_explicit(p, mode, explicit_r);
  if (rc != SC_OK) return rc;
  hipStream_t s = p->ctx->stream;
  p->last_mode = mode;
  SC_HIP(hipEventRecord(p->ev0, s));
  if (p->graph_exec != nullptr && p->graph_mode == mode && !p->profiling) {
    SC_HIP(hipGraphLaunch(p->graph_exec, s));
    p->profiled_last = false;
  } else {
    rc = enqueue_rounds(p, mode);
    if (rc != SC_OK) return rc;
    p->profiled_last = p->profiling;
  }
  SC_HIP(hipEventRecord(p->ev1, s));
  return SC_OK;
}

int sc_proof_capture(sc_proof* p, int32_t mode, const uint32_t* explicit_r) {
  if (p == nullptr) return fail(SC_ERR_ARG, "proof must not be NULL");
  int rc = check_mode(p, mode, explicit_r);
  if (rc != SC_OK) return rc;
  SC_HIP(hipSetDevice(p->ctx->device));
  rc = upload_explicit(p, mode, explicit_r);
  if (rc != SC_OK) return rc;
  drop_graph(p);
  hipStream_t s = p->ctx->stream;
  const bool prof = p->profiling;
  p->profiling = false;  // no event records inside the graph
  SC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  rc = enqueue_rounds(p, mode);
  p->profiling = prof;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (rc != SC_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(SC_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  p->graph = g;
  SC_HIP(hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0));
  p->graph_mode = mode;
  return SC_OK;
}

int sc_proof_wait(sc_proof* p, float* ms) {
  if (p == nullptr) return fail(SC_ERR_ARG, "proof must not be NULL");
  SC_HIP(hipSetDevice(p->ctx->device));
  SC_HIP(hipStreamSynchronize(p->ctx->stream));
  if (ms != nullptr) {
    *ms = 0.0f;
    SC_HIP(hipEventElapsedTime(ms, p->ev0, p->ev1));
  }
  return SC_OK;
}

namespace {

// Bytes of ProofState a result slot holds: r (all rounds) and polys up to the proof's last
// round, one contiguous range.
uint64_t slot_bytes(const sc_proof* p) {
  return offsetof(ProofState, polys) - offsetof(ProofState, r) +
         (uint64_t)p->num_vars * sizeof(((ProofState*)nullptr)->polys[0]);
}

// The challenges and round polynomials of a ProofState image (its r and polys arrays) in the
// layout of sc_proof_result, with the checksum.
void decode_result(const sc_proof* p, const uint32_t (*r)[MAX_LIMBS], const uint32_t (*polys)[MAX_POINTS][MAX_LIMBS],
                   uint32_t* round_polys, uint32_t* challenges, uint64_t* checksum) {
  uint64_t sum = 0;
  for (int j = 0; j < p->num_vars; ++j) {
    for (int t = 0; t < p->npts; ++t) {
      for (int i = 0; i < p->L; ++i) {
        const uint32_t w = polys[j][t][i];
        sum += (uint64_t)(i + 1) * w;
        if (round_polys) round_polys[((size_t)j * p->npts + t) * p->L + i] = w;
      }
    }
    if (challenges && j + 1 < p->num_vars) {
      for (int i = 0; i < p->L; ++i) challenges[(size_t)j * p->L + i] = r[j][i];
    }
  }
  if (checksum) *checksum = sum;
}

}  // namespace

int sc_proof_result(sc_proof* p, uint32_t* round_polys, uint32_t* challenges, uint64_t* checksum) {
  if (p == nullptr) return fail(SC_ERR_ARG, "proof must not be NULL");
  SC_HIP(hipSetDevice(p->ctx->device));
  SC_HIP(hipStreamSynchronize(p->ctx->stream));
  std::vector<ProofState> host(1);
  SC_HIP(hipMemcpy(host.data(), p->d_state, sizeof(ProofState), hipMemcpyDeviceToHost));
  const ProofState& st = host[0];
  decode_result(p, st.r, st.polys, round_polys, challenges, checksum);
  return SC_OK;
}

int sc_proof_run_async(sc_proof* p, int32_t mode, const uint32_t* explicit_r, int32_t* slot) {
  if (p == nullptr || slot == nullptr) return fail(SC_ERR_ARG, "proof and slot must not be NULL");
  SC_HIP(hipSetDevice(p->ctx->device));
  if (p->h_slot[0] == nullptr) {
    for (int k = 0; k < 2; ++k) {
      SC_HIP(hipHostMalloc((void**)&p->h_slot[k], slot_bytes(p), hipHostMallocDefault));
      SC_HIP(hipEventCreateWithFlags(&p->ev_slot[k], hipEventDisableTiming));
    }
  }
  int rc = sc_proof_run(p, mode, explicit_r);
  if (rc != SC_OK) return rc;
  const int k = (int)(p->async_runs & 1u);
  hipStream_t s = p->ctx->stream;
  SC_HIP(hipMemcpyAsync(p->h_slot[k], &p->d_state->r[0][0], slot_bytes(p), hipMemcpyDeviceToHost, s));
  SC_HIP(hipEventRecord(p->ev_slot[k], s));
  ++p->async_runs;
  *slot = k;
  return SC_OK;
}

int sc_proof_slot_result(sc_proof* p, int32_t slot, uint32_t* round_polys, uint32_t* challenges,
                         uint64_t* checksum) {
  if (p == nullptr) return fail(SC_ERR_ARG, "proof must not be NULL");
  if (slot < 0 || slot > 1 || p->h_slot[slot] == nullptr)
    return fail(SC_ERR_ARG, "slot must be one returned by sc_proof_run_async");
  SC_HIP(hipSetDevice(p->ctx->device));
  SC_HIP(hipEventSynchronize(p->ev_slot[slot]));
  const uint8_t* base = reinterpret_cast<const uint8_t*>(p->h_slot[slot]);
  const auto* r = reinterpret_cast<const uint32_t(*)[MAX_LIMBS]>(base);
  const auto* polys = reinterpret_cast<const uint32_t(*)[MAX_POINTS][MAX_LIMBS]>(
      base + offsetof(ProofState, polys) - offsetof(ProofState, r));
  decode_result(p, r, polys, round_polys, challenges, checksum);
  return SC_OK;
}

int sc_proof_final(sc_proof* p, uint32_t* r_final, uint32_t* final_values) {
  if (p == nullptr) return fail(SC_ERR_ARG, "proof must not be NULL");
  SC_HIP(hipSetDevice(p->ctx->device));
  SC_HIP(hipStreamSynchronize(p->ctx->stream));
  std::vector<ProofState> host(1);
  SC_HIP(hipMemcpy(host.data(), p->d_state, sizeof(ProofState), hipMemcpyDeviceToHost));
  const ProofState& st = host[0];
  if (st.final_ready == 0u || p->last_mode == SC_CHALLENGE_EXPLICIT)
    return fail(SC_ERR_UNSUPPORTED, "no final point: run the proof with fixed or sha3 challenges first");
  for (int i = 0; i < p->L; ++i) {
    if (r_final) r_final[i] = st.r[p->num_vars - 1][i];
  }
  for (int v = 0; v < p->T; ++v) {
    for (int i = 0; i < p->L; ++i) {
      if (final_values) final_values[(size_t)v * p->L + i] = st.final_vals[v][i];
    }
  }
  return SC_OK;
}

int sc_proof_tables(sc_proof* p, uint32_t* out) {
  if (p == nullptr || out == nullptr) return fail(SC_ERR_ARG, "proof and out must not be NULL");
  SC_HIP(hipSetDevice(p->ctx->device));
  SC_HIP(hipStreamSynchronize(p->ctx->stream));
  const uint64_t table_bytes = p->n0 * p->L * 4;
  for (int v = 0; v < p->T; ++v) {
    SC_HIP(hipMemcpy((uint8_t*)out + v * table_bytes, (const uint8_t*)p->d_tables + v * p->in_stride * p->L * 4,
                     table_bytes, hipMemcpyDeviceToHost));
  }
  return SC_OK;
}

int sc_proof_info(sc_proof* p, int32_t* degree, int32_t* pass_rounds, uint64_t* device_bytes,
                  int64_t* point_parallel_pairs) {
  if (p == nullptr) return fail(SC_ERR_ARG, "proof must not be NULL");
  if (degree) *degree = p->npts - 1;
  if (pass_rounds) *pass_rounds = tail_first_round(p);
  if (device_bytes) *device_bytes = p->device_bytes;
  if (point_parallel_pairs) *point_parallel_pairs = p->wide_pairs;
  return SC_OK;
}

int sc_proof_use_compiled(sc_proof* p, int32_t enable) {
  if (p == nullptr) return fail(SC_ERR_ARG, "proof must not be NULL");
  const int sid = enable ? p->matched_sid : 0;
  if (sid != p->sid) drop_graph(p);
  p->sid = sid;
  if (p->default_wide) p->wide_pairs = default_wide_pairs(p);
  return p->matched_sid;
}

int sc_proof_set_profiling(sc_proof* p, int32_t enable) {
  if (p == nullptr) return fail(SC_ERR_ARG, "proof must not be NULL");
  SC_HIP(hipSetDevice(p->ctx->device));
  if (enable && p->prof_ev[0] == nullptr) {
    // timestamps only: without the system-scope fence of a default event, the records
    // between rounds cost ~1.5% of a 2^26-row proof instead of ~4%
    for (hipEvent_t& e : p->prof_ev) SC_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    for (hipEvent_t& e : p->prof_fin) SC_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
  }
  p->profiling = enable != 0;
  return SC_OK;
}

int sc_proof_kernel_times(sc_proof* p, float* pass_ms, int32_t capacity, float* tail_ms) {
  if (p == nullptr) return fail(SC_ERR_ARG, "proof must not be NULL");
  if (!p->profiled_last) return fail(SC_ERR_ARG, "the last run was not profiled (sc_proof_set_profiling)");
  SC_HIP(hipSetDevice(p->ctx->device));
  SC_HIP(hipStreamSynchronize(p->ctx->stream));
  const int jt = tail_first_round(p);
  for (int j = 0; j < jt && j < capacity; ++j) {
    float ms = 0.0f;
    SC_HIP(hipEventElapsedTime(&ms, p->prof_ev[2 * j], p->prof_ev[2 * j + 1]));
    if (pass_ms) pass_ms[j] = ms;
  }
  if (tail_ms) {
    *tail_ms = 0.0f;
    if (jt < p->num_vars) SC_HIP(hipEventElapsedTime(tail_ms, p->prof_ev[2 * MAX_ROUNDS], p->prof_ev[2 * MAX_ROUNDS + 1]));
  }
  return SC_OK;
}

int sc_proof_round_times(sc_proof* p, float* pass_ms, float* finalize_ms, int32_t capacity, float* tail_ms) {
  int rc = sc_proof_kernel_times(p, pass_ms, capacity, tail_ms);
  if (rc != SC_OK) return rc;
  const int jt = tail_first_round(p);
  for (int j = 0; j < jt && j < capacity; ++j) {
    float ms = 0.0f;
    SC_HIP(hipEventElapsedTime(&ms, p->prof_ev[2 * j + 1], p->prof_fin[j]));
    if (finalize_ms) finalize_ms[j] = ms;
  }
  return SC_OK;
}

int sc_proof_destroy(sc_proof* p) {
  if (p == nullptr) return SC_OK;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->ctx->stream);
  free_proof(p);
  return SC_OK;
}

int sc_prove(sc_context* ctx, int32_t bit_width, const sc_composition* comp, int32_t num_vars,
             const sc_input* input, int32_t mode, const uint32_t* explicit_r, uint32_t* round_polys,
             uint32_t* challenges, uint64_t* checksum) {
  sc_proof* p = nullptr;
  int rc = sc_proof_create(ctx, bit_width, comp, num_vars, input, &p);
  if (rc != SC_OK) return rc;
  rc = sc_proof_run(p, mode, explicit_r);
  if (rc == SC_OK) rc = sc_proof_wait(p, nullptr);
  if (rc == SC_OK) rc = sc_proof_result(p, round_polys, challenges, checksum);
  std::string err = g_error;
  sc_proof_destroy(p);
  g_error = err;
  return rc;
}

int sc_prove_noalloc(sc_context* ctx, int32_t bit_width, const sc_composition* comp, int32_t num_vars,
                     const uint32_t* tables, int32_t mode, uint32_t* round_polys, uint32_t* challenges,
                     uint64_t* checksum) {
  if (ctx == nullptr || tables == nullptr || round_polys == nullptr)
    return fail(SC_ERR_ARG, "ctx, tables and round_polys must not be NULL");
  const int fidx = field_index(bit_width);
  if (fidx < 0) return fail(SC_ERR_ARG, "bit_width must be 32, 64, 128 or 256");
  if (num_vars < 1 || num_vars > MAX_ROUNDS) return fail(SC_ERR_UNSUPPORTED, "num_vars must be in [1, 40]");
  if (mode != SC_CHALLENGE_FIXED && mode != SC_CHALLENGE_SHA3)
    return fail(SC_ERR_UNSUPPORTED, "sc_prove_noalloc takes challenge_mode 0 (fixed) or 1 (sha3)");
  CompDesc desc;
  int degree = 0;
  int rc = build_comp_any(fidx, comp, ctx->host_fc[fidx], desc, degree);
  if (rc != SC_OK) return rc;
  const uint64_t L = (uint64_t)launchers_for(fidx).limbs;
  const uint64_t n0 = (uint64_t)1 << num_vars;
  const uint64_t fold_words = (uint64_t)desc.num_tables * (n0 / 2 + n0 / 4) * L;
  if (fold_words > NOALLOC_FOLD_WORDS)
    return fail(SC_ERR_UNSUPPORTED, "sc_prove_noalloc: the intermediate tables (0.75x the input, " +
                                        std::to_string(fold_words * 4) + " bytes) exceed the library's fixed " +
                                        std::to_string(NOALLOC_FOLD_WORDS * 4) + "-byte area");
  if (ctx->device < 0 || ctx->device >= MAX_DEVICES) return fail(SC_ERR_UNSUPPORTED, "device index above 63");
  uint8_t* base = g_scratch[ctx->device];
  if (base == nullptr) return fail(SC_ERR_HIP, g_scratch_error[ctx->device]);
  SC_HIP(hipSetDevice(ctx->device));
  // A host-side description only: every device pointer is the caller's memory or scratch.
  sc_proof p;
  init_shape(&p, ctx, fidx, desc, num_vars);
  p.d_tables = const_cast<uint32_t*>(tables);
  p.in_stride = n0;  // the caller's tables, packed
  p.d_a = reinterpret_cast<uint32_t*>(base);
  p.d_b = p.d_a + (uint64_t)p.T * (n0 / 2) * L;
  p.d_partials = reinterpret_cast<uint32_t*>(base + NOALLOC_OFF_PARTIALS);
  p.d_comp = reinterpret_cast<CompDesc*>(base + NOALLOC_OFF_COMP);
  p.d_state = reinterpret_cast<ProofState*>(base + NOALLOC_OFF_STATE);
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(noalloc_setup_kernel, dim3(1), dim3(BLOCK), 0, s, desc, p.d_comp, p.d_state);
  SC_HIP(hipGetLastError());
  rc = enqueue_rounds(&p, mode);
  if (rc != SC_OK) return rc;
  hipLaunchKernelGGL(noalloc_export_kernel, dim3(1), dim3(BLOCK), 0, s, (const ProofState*)p.d_state, num_vars,
                     p.npts, (int)L, round_polys, challenges, checksum);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int sc_field_op(sc_context* ctx, int32_t bit_width, int32_t op, uint64_t n, const uint32_t* a,
                const uint32_t* b, uint32_t* out) {
  if (ctx == nullptr || a == nullptr || out == nullptr) return fail(SC_ERR_ARG, "NULL argument");
  const int fidx = field_index(bit_width);
  if (fidx < 0) return fail(SC_ERR_ARG, "bit_width must be 32, 64, 128 or 256");
  if (op < 0 || op > 6) return fail(SC_ERR_ARG, "op must be in [0, 6]");
  if ((op <= 2 || op >= 5) && b == nullptr) return fail(SC_ERR_ARG, "b must not be NULL for this op");
  if (n == 0) return SC_OK;
  SC_HIP(hipSetDevice(ctx->device));
  const Launchers& ln = launchers_for(fidx);
  const uint64_t bytes = n * ln.limbs * 4;
  uint32_t* d = nullptr;
  SC_HIP(hipMalloc(&d, 3 * bytes));
  hipStream_t s = ctx->stream;
  hipError_t e = hipMemcpyAsync(d, a, bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && b != nullptr) e = hipMemcpyAsync(d + n * ln.limbs, b, bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = ln.field_op(s, op, n, d, d + n * ln.limbs, d + 2 * n * ln.limbs, ctx->d_fc[fidx]);
  if (e == hipSuccess) e = hipMemcpyAsync(out, d + 2 * n * ln.limbs, bytes, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(SC_ERR_HIP, std::string("sc_field_op: ") + hipGetErrorString(e));
  return SC_OK;
}

int sc_transcript_challenge(sc_context* ctx, int32_t bit_width, int32_t round, const uint32_t* g_mont,
                            int32_t npts, uint32_t* r_mont) {
  if (ctx == nullptr || g_mont == nullptr || r_mont == nullptr) return fail(SC_ERR_ARG, "NULL argument");
  const int fidx = field_index(bit_width);
  if (fidx < 0) return fail(SC_ERR_ARG, "bit_width must be 32, 64, 128 or 256");
  if (npts < 1 || npts > MAX_POINTS) return fail(SC_ERR_ARG, "npts must be in [1, 8]");
  if (round < 0 || round + 1 >= MAX_ROUNDS) return fail(SC_ERR_ARG, "round out of range");
  SC_HIP(hipSetDevice(ctx->device));
  const Launchers& ln = launchers_for(fidx);
  uint32_t* d_g = nullptr;
  ProofState* d_st = nullptr;
  SC_HIP(hipMalloc(&d_g, (size_t)npts * ln.limbs * 4));
  hipError_t e = hipMalloc(&d_st, sizeof(ProofState));
  hipStream_t s = ctx->stream;
  if (e == hipSuccess) e = hipMemcpyAsync(d_g, g_mont, (size_t)npts * ln.limbs * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = ln.transcript(s, d_g, npts, round, ctx->d_fc[fidx], d_st);
  if (e == hipSuccess) e = hipMemcpyAsync(r_mont, d_st->r[round], (size_t)ln.limbs * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d_g);
  if (d_st) (void)hipFree(d_st);
  if (e != hipSuccess) return fail(SC_ERR_HIP, std::string("sc_transcript_challenge: ") + hipGetErrorString(e));
  return SC_OK;
}

int sc_calibrate(sc_context* ctx, double* rates, int32_t n) {
  if (ctx == nullptr || rates == nullptr) return fail(SC_ERR_ARG, "NULL argument");
  if (n < 0 || n > CALIB_COUNT) return fail(SC_ERR_ARG, "n must be in [0, 7]");
  SC_HIP(hipSetDevice(ctx->device));
  for (int k = 0; k < n; ++k) rates[k] = 0.0;
  SC_HIP(calibrate(ctx->stream, rates, n));
  return SC_OK;
}

// ---------------------------------------------------------------------------------------------
// Batched proving
// ---------------------------------------------------------------------------------------------

}  // extern "C"

#define SC_HIP_B(expr)                                                                        \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) {                                                                   \
      free_batch(b);                                                                          \
      return fail(_e == hipErrorOutOfMemory ? SC_ERR_MEMORY : SC_ERR_HIP,                     \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                         \
    }                                                                                         \
  } while (0)

struct sc_batch {
  sc_context* ctx = nullptr;
  int fidx = 0, L = 0, n = 0, tmax = 0, pmax = 0, max_rounds = 0;
  const Launchers* ln = nullptr;
  std::vector<int> num_vars, npts, num_tables;
  std::vector<uint64_t> n0;
  uint8_t* pool = nullptr;
  uint64_t pool_bytes = 0;
  InstDesc* d_insts = nullptr;
  CompDesc* d_comps = nullptr;
  ProofState* d_states = nullptr;
  StageEntry* d_entries = nullptr;
  uint32_t* d_partials = nullptr;
  // Stage s: its lists start at stage_offset[s] in d_entries, laid out as StagePlan::layout.
  std::vector<uint32_t> stage_offset, stage_pass, stage_eval, stage_fold;
  std::vector<uint64_t> stage_pass_tiles, stage_eval_tiles, stage_fold_tiles;
  // Per-instance schedule: tile (pairs per thread), the pair count at or below which a round
  // runs point-parallel (-1: default_instance_wide_pairs) and the tail threshold (0: default).
  std::vector<uint32_t> items_i;
  std::vector<int64_t> wide_i, tail_i;
  // Each instance's last rounds (from tail_round[i] on, at most tail_pairs pairs) run as one
  // tail entry of the last stage: one workgroup loops over them, as a single proof's tail
  // kernel does. All tails share that one launch (a tail launch per stage measured slower:
  // each adds the latency of its slowest instance's several rounds to the critical path).
  std::vector<uint32_t> tail_offset, tail_entries;
  std::vector<int> tail_round;
  int num_stages = 0;
  int mode = -1;  // challenge mode baked into the instance descriptors
  std::vector<InstDesc> insts;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
};

namespace {

void free_batch(sc_batch* b) {
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->pool) (void)hipFree(b->pool);
  if (b->d_insts) (void)hipFree(b->d_insts);
  if (b->d_comps) (void)hipFree(b->d_comps);
  if (b->d_states) (void)hipFree(b->d_states);
  if (b->d_entries) (void)hipFree(b->d_entries);
  if (b->d_partials) (void)hipFree(b->d_partials);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  delete b;
}

void drop_batch_graph(sc_batch* b) {
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
}

// Work lists of every stage for the current tile size, uploaded to the device.
int64_t batch_wide_pairs(const sc_batch* b, int i) {
  return b->wide_i[i] < 0 ? default_instance_wide_pairs(b->L, b->tmax, b->npts[i]) : b->wide_i[i];
}

// First round of instance i that runs in the batch's tail launch: the first whose pairs fit
// its tail threshold (as for a single proof), at most its last round.
void batch_tail_rounds(sc_batch* b) {
  for (int i = 0; i < b->n; ++i) {
    const int64_t tp = b->tail_i[i] > 0 ? b->tail_i[i] : default_tail_pairs(b->L, b->tmax > 8);
    int j = 0;
    while (j < b->num_vars[i] - 1 && (int64_t)(b->n0[i] >> (j + 1)) > tp) ++j;
    b->tail_round[i] = j;
  }
}

int build_stages(sc_batch* b) {
  std::vector<StageEntry> all;
  batch_tail_rounds(b);
  b->num_stages = 0;
  for (int i = 0; i < b->n; ++i) b->num_stages = std::max(b->num_stages, b->tail_round[i] + 1);
  const int last = b->num_stages - 1;
  const int ns = b->num_stages;
  b->stage_offset.assign(ns, 0);
  b->stage_pass.assign(ns, 0);
  b->stage_eval.assign(ns, 0);
  b->stage_fold.assign(ns, 0);
  b->stage_pass_tiles.assign(ns, 0);
  b->stage_eval_tiles.assign(ns, 0);
  b->stage_fold_tiles.assign(ns, 0);
  b->tail_offset.assign(ns, 0);
  b->tail_entries.assign(ns, 0);
  uint64_t max_tiles = 1;
  for (int s = 0; s < ns; ++s) {
    b->stage_offset[s] = (uint32_t)all.size();
    StagePlan plan;
    for (int i = 0; i < b->n; ++i) {
      if (s >= b->tail_round[i]) continue;
      const uint64_t pairs = b->n0[i] >> (s + 1);
      plan.add((uint32_t)i, (uint32_t)s, pairs, b->num_tables[i], (int64_t)pairs <= batch_wide_pairs(b, i),
               b->items_i[i], b->npts[i]);
      if (plan.partial_tiles() > 0x7FFFFFFFull || plan.fold_tiles > 0x7FFFFFFFull)
        return fail(SC_ERR_UNSUPPORTED, "batch stage has too many tiles");
    }
    plan.layout(all);
    b->stage_pass[s] = (uint32_t)plan.pass.size();
    b->stage_eval[s] = (uint32_t)plan.eval.size();
    b->stage_fold[s] = (uint32_t)plan.fold.size();
    b->stage_pass_tiles[s] = plan.pass_tiles;
    b->stage_eval_tiles[s] = plan.eval_tiles;
    b->stage_fold_tiles[s] = plan.fold_tiles;
    max_tiles = std::max(max_tiles, plan.partial_tiles());
    b->tail_offset[s] = (uint32_t)all.size();
    if (s == last) {
      for (int i = 0; i < b->n; ++i) all.push_back(StageEntry{(uint32_t)i, (uint32_t)b->tail_round[i], 0u, 0u, 1u});
    }
    b->tail_entries[s] = (uint32_t)all.size() - b->tail_offset[s];
  }
  if (b->d_entries) (void)hipFree(b->d_entries);
  if (b->d_partials) (void)hipFree(b->d_partials);
  b->d_entries = nullptr;
  b->d_partials = nullptr;
  SC_HIP(hipMalloc(&b->d_entries, sizeof(StageEntry) * (all.empty() ? 1 : all.size())));
  SC_HIP(hipMalloc(&b->d_partials, max_tiles * MAX_POINTS * MAX_LIMBS * 4));
  if (!all.empty())
    SC_HIP(hipMemcpy(b->d_entries, all.data(), sizeof(StageEntry) * all.size(), hipMemcpyHostToDevice));
  drop_batch_graph(b);
  return SC_OK;
}

int set_batch_mode(sc_batch* b, int mode) {
  if (mode != SC_CHALLENGE_FIXED && mode != SC_CHALLENGE_SHA3)
    return fail(SC_ERR_ARG, "batched proofs take challenge_mode 0 (fixed) or 1 (sha3)");
  if (mode == b->mode) return SC_OK;
  for (auto& d : b->insts) d.challenge_mode = mode;
  SC_HIP(hipMemcpy(b->d_insts, b->insts.data(), sizeof(InstDesc) * b->n, hipMemcpyHostToDevice));
  b->mode = mode;
  drop_batch_graph(b);
  return SC_OK;
}

int enqueue_batch(sc_batch* b) {
  hipStream_t s = b->ctx->stream;
  for (int st = 0; st < b->num_stages; ++st) {
    BatchArgs a;
    a.insts = b->d_insts;
    a.chunk_shift = default_fold_chunk_shift(b->L);
    a.partials = b->d_partials;
    a.fc = b->ctx->d_fc[b->fidx];
    a.tile_base = 0;
    a.pad0 = 0;
    int rc = enqueue_stage(*b->ln, s, a, b->tmax, b->pmax, b->d_entries + b->stage_offset[st], b->stage_pass[st],
                           b->stage_eval[st], b->stage_fold[st], b->stage_pass_tiles[st], b->stage_eval_tiles[st],
                           b->stage_fold_tiles[st]);
    if (rc != SC_OK) return rc;
    if (b->tail_entries[st] != 0) {
      a.entries = b->d_entries + b->tail_offset[st];
      a.n_entries = b->tail_entries[st];
      SC_HIP(b->ln->batch_tail(s, b->tmax, b->pmax, a, b->tail_entries[st]));
    }
  }
  return SC_OK;
}

}  // namespace

extern "C" {

int sc_batch_create(sc_context* ctx, int32_t bit_width, int32_t n, const sc_composition* comps,
                    const int32_t* num_vars, const sc_input* inputs, sc_batch** out) {
  if (ctx == nullptr || comps == nullptr || num_vars == nullptr || inputs == nullptr || out == nullptr)
    return fail(SC_ERR_ARG, "NULL argument");
  *out = nullptr;
  const int fidx = field_index(bit_width);
  if (fidx < 0) return fail(SC_ERR_ARG, "bit_width must be 32, 64, 128 or 256");
  if (n < 1 || n > 65536) return fail(SC_ERR_ARG, "instance count must be in [1, 65536]");
  std::vector<CompDesc> descs(n);
  sc_batch* b = new (std::nothrow) sc_batch();
  if (b == nullptr) return fail(SC_ERR_MEMORY, "host allocation failed");
  b->ctx = ctx;
  b->fidx = fidx;
  b->ln = &launchers_for(fidx);
  b->L = b->ln->limbs;
  b->n = n;
  b->num_vars.resize(n);
  b->npts.resize(n);
  b->num_tables.resize(n);
  b->n0.resize(n);
  b->tail_round.resize(n);
  int maxT = 0, maxP = 0;
  std::vector<uint64_t> offs(n);
  uint64_t total = 0;
  auto align = [](uint64_t x) { return (x + 255) & ~(uint64_t)255; };
  for (int i = 0; i < n; ++i) {
    int degree = 0;
    int rc = build_comp_any(fidx, &comps[i], ctx->host_fc[fidx], descs[i], degree);
    if (rc != SC_OK) {
      free_batch(b);
      return rc;
    }
    if (num_vars[i] < 1 || num_vars[i] > MAX_ROUNDS) {
      free_batch(b);
      return fail(SC_ERR_UNSUPPORTED, "num_vars must be in [1, 40]");
    }
    if (inputs[i].kind != SC_INPUT_AFFINE && inputs[i].kind != SC_INPUT_UNIFORM) {
      free_batch(b);
      return fail(SC_ERR_ARG, "batched instances take generated inputs (affine or uniform)");
    }
    b->num_vars[i] = num_vars[i];
    b->npts[i] = descs[i].npts;
    b->num_tables[i] = descs[i].num_tables;
    b->n0[i] = (uint64_t)1 << num_vars[i];
    maxT = std::max(maxT, descs[i].num_tables);
    maxP = std::max(maxP, descs[i].npts);
    b->max_rounds = std::max(b->max_rounds, num_vars[i]);
    offs[i] = total;
    const uint64_t words = (uint64_t)descs[i].num_tables * b->n0[i] * b->L;
    total += align(words * 4) + align(words / 2 * 4) + align(words / 4 * 4 + 4);
  }
  b->tmax = pick_tmax(maxT);
  b->pmax = maxP <= 4 ? 4 : 8;
  b->items_i.assign(n, 4u);
  b->wide_i.assign(n, -1);
  b->tail_i.assign(n, 0);
  SC_HIP_B(hipSetDevice(ctx->device));
  SC_HIP_B(hipMalloc(&b->pool, total));
  b->pool_bytes = total;
  SC_HIP_B(hipMalloc(&b->d_insts, sizeof(InstDesc) * n));
  SC_HIP_B(hipMalloc(&b->d_comps, sizeof(CompDesc) * n));
  SC_HIP_B(hipMalloc(&b->d_states, sizeof(ProofState) * n));
  SC_HIP_B(hipMemcpy(b->d_comps, descs.data(), sizeof(CompDesc) * n, hipMemcpyHostToDevice));
  SC_HIP_B(hipMemset(b->d_states, 0, sizeof(ProofState) * n));
  SC_HIP_B(hipEventCreate(&b->ev0));
  SC_HIP_B(hipEventCreate(&b->ev1));
  b->insts.resize(n);
  for (int i = 0; i < n; ++i) {
    const uint64_t words = (uint64_t)descs[i].num_tables * b->n0[i] * b->L;
    uint8_t* base = b->pool + offs[i];
    InstDesc& d = b->insts[i];
    d.input = (const uint32_t*)base;
    d.buf_a = (uint32_t*)(base + align(words * 4));
    d.buf_b = (uint32_t*)(base + align(words * 4) + align(words / 2 * 4));
    d.n0 = b->n0[i];
    d.num_vars = num_vars[i];
    d.challenge_mode = SC_CHALLENGE_SHA3;
    d.comp = b->d_comps + i;
    d.st = b->d_states + i;
    FillArgs f;
    f.tables = (uint32_t*)base;
    f.n = b->n0[i];
    f.stride = b->n0[i];
    f.num_tables = descs[i].num_tables;
    f.kind = inputs[i].kind == SC_INPUT_AFFINE ? FILL_AFFINE : FILL_UNIFORM;
    f.seed = inputs[i].seed;
    f.selector_mask = inputs[i].selector_mask;
    f.instance = inputs[i].instance;
    f.pad0 = 0;
    f.fc = ctx->d_fc[fidx];
    SC_HIP_B(b->ln->fill(ctx->stream, f));
  }
  SC_HIP_B(hipMemcpy(b->d_insts, b->insts.data(), sizeof(InstDesc) * n, hipMemcpyHostToDevice));
  b->mode = SC_CHALLENGE_SHA3;
  SC_HIP_B(hipStreamSynchronize(ctx->stream));
  int rc = build_stages(b);
  if (rc != SC_OK) {
    free_batch(b);
    return rc;
  }
  *out = b;
  return SC_OK;
}

int sc_batch_set_point_parallel(sc_batch* b, int64_t max_pairs) {
  if (b == nullptr) return fail(SC_ERR_ARG, "batch must not be NULL");
  b->wide_i.assign(b->n, max_pairs < 0 ? -1 : max_pairs);
  SC_HIP(hipStreamSynchronize(b->ctx->stream));
  return build_stages(b);
}

int sc_batch_set_tile(sc_batch* b, int32_t items_per_thread) {
  if (b == nullptr) return fail(SC_ERR_ARG, "batch must not be NULL");
  if (items_per_thread < 1 || items_per_thread > 64) return fail(SC_ERR_ARG, "items_per_thread must be in [1, 64]");
  b->items_i.assign(b->n, (uint32_t)items_per_thread);
  SC_HIP(hipStreamSynchronize(b->ctx->stream));
  return build_stages(b);
}

int sc_batch_set_schedules(sc_batch* b, const int32_t* items_per_thread, const int64_t* max_pairs,
                           const int64_t* tail_pairs) {
  if (b == nullptr) return fail(SC_ERR_ARG, "batch must not be NULL");
  for (int i = 0; i < b->n; ++i) {
    if (items_per_thread && (items_per_thread[i] < 0 || items_per_thread[i] > 64))
      return fail(SC_ERR_ARG, "items_per_thread must be in [0, 64]");
    if (tail_pairs && tail_pairs[i] < 0) return fail(SC_ERR_ARG, "tail_pairs must be >= 0");
  }
  for (int i = 0; i < b->n; ++i) {
    if (items_per_thread && items_per_thread[i] > 0) b->items_i[i] = (uint32_t)items_per_thread[i];
    if (max_pairs) b->wide_i[i] = max_pairs[i] < 0 ? -1 : max_pairs[i];
    if (tail_pairs) b->tail_i[i] = tail_pairs[i];
  }
  SC_HIP(hipStreamSynchronize(b->ctx->stream));
  return build_stages(b);
}

int sc_batch_instance_schedule(sc_batch* b, int32_t i, int32_t* items_per_thread, int64_t* max_pairs,
                               int32_t* tail_round) {
  if (b == nullptr) return fail(SC_ERR_ARG, "batch must not be NULL");
  if (i < 0 || i >= b->n) return fail(SC_ERR_ARG, "instance index out of range");
  if (items_per_thread) *items_per_thread = (int32_t)b->items_i[i];
  if (max_pairs) *max_pairs = batch_wide_pairs(b, i);
  if (tail_round) *tail_round = b->tail_round[i];
  return SC_OK;
}

int sc_batch_run(sc_batch* b, int32_t mode) {
  if (b == nullptr) return fail(SC_ERR_ARG, "batch must not be NULL");
  SC_HIP(hipSetDevice(b->ctx->device));
  SC_HIP(hipStreamSynchronize(b->ctx->stream));
  int rc = set_batch_mode(b, mode);
  if (rc != SC_OK) return rc;
  hipStream_t s = b->ctx->stream;
  SC_HIP(hipEventRecord(b->ev0, s));
  if (b->graph_exec) {
    SC_HIP(hipGraphLaunch(b->graph_exec, s));
  } else {
    rc = enqueue_batch(b);
    if (rc != SC_OK) return rc;
  }
  SC_HIP(hipEventRecord(b->ev1, s));
  return SC_OK;
}

int sc_batch_capture(sc_batch* b, int32_t mode) {
  if (b == nullptr) return fail(SC_ERR_ARG, "batch must not be NULL");
  SC_HIP(hipSetDevice(b->ctx->device));
  SC_HIP(hipStreamSynchronize(b->ctx->stream));
  int rc = set_batch_mode(b, mode);
  if (rc != SC_OK) return rc;
  drop_batch_graph(b);
  hipStream_t s = b->ctx->stream;
  SC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  rc = enqueue_batch(b);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (rc != SC_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(SC_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  b->graph = g;
  SC_HIP(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
  return SC_OK;
}

int sc_batch_wait(sc_batch* b, float* ms) {
  if (b == nullptr) return fail(SC_ERR_ARG, "batch must not be NULL");
  SC_HIP(hipSetDevice(b->ctx->device));
  SC_HIP(hipStreamSynchronize(b->ctx->stream));
  if (ms != nullptr) {
    *ms = 0.0f;
    SC_HIP(hipEventElapsedTime(ms, b->ev0, b->ev1));
  }
  return SC_OK;
}

int sc_batch_result(sc_batch* b, int32_t i, uint32_t* round_polys, uint32_t* challenges, uint64_t* checksum) {
  if (b == nullptr) return fail(SC_ERR_ARG, "batch must not be NULL");
  if (i < 0 || i >= b->n) return fail(SC_ERR_ARG, "instance index out of range");
  SC_HIP(hipSetDevice(b->ctx->device));
  SC_HIP(hipStreamSynchronize(b->ctx->stream));
  std::vector<ProofState> host(1);
  SC_HIP(hipMemcpy(host.data(), b->d_states + i, sizeof(ProofState), hipMemcpyDeviceToHost));
  const ProofState& st = host[0];
  const int npts = b->npts[i], L = b->L, nv = b->num_vars[i];
  uint64_t sum = 0;
  for (int j = 0; j < nv; ++j) {
    for (int t = 0; t < npts; ++t) {
      for (int k = 0; k < L; ++k) {
        const uint32_t w = st.polys[j][t][k];
        sum += (uint64_t)(k + 1) * w;
        if (round_polys) round_polys[((size_t)j * npts + t) * L + k] = w;
      }
    }
    if (challenges && j + 1 < nv)
      for (int k = 0; k < L; ++k) challenges[(size_t)j * L + k] = st.r[j][k];
  }
  if (checksum) *checksum = sum;
  return SC_OK;
}

int sc_batch_info(sc_batch* b, uint64_t* device_bytes, int32_t* stages, uint64_t* tiles) {
  if (b == nullptr) return fail(SC_ERR_ARG, "batch must not be NULL");
  if (device_bytes) *device_bytes = b->pool_bytes;
  if (stages) *stages = b->num_stages;
  if (tiles) {
    uint64_t t = 0;
    for (int s = 0; s < b->num_stages; ++s) t += b->stage_pass_tiles[s] + b->stage_eval_tiles[s];
    *tiles = t;
  }
  return SC_OK;
}

int sc_batch_destroy(sc_batch* b) {
  if (b == nullptr) return SC_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  free_batch(b);
  return SC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Serving: continuous batching of independent proof requests
// ---------------------------------------------------------------------------------------------

namespace {

// First-fit arena over one device allocation; offsets are 256-byte aligned.
struct Arena {
  uint8_t* base = nullptr;
  uint64_t size = 0;
  std::map<uint64_t, uint64_t> free_blocks;  // offset -> length
  uint64_t used = 0;

  void init(uint8_t* b, uint64_t n) {
    base = b;
    size = n;
    free_blocks.clear();
    free_blocks[0] = n;
    used = 0;
  }
  bool alloc(uint64_t n, uint64_t& off) {
    n = (n + 255) & ~(uint64_t)255;
    for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
      if (it->second >= n) {
        off = it->first;
        const uint64_t rest = it->second - n;
        free_blocks.erase(it);
        if (rest) free_blocks[off + n] = rest;
        used += n;
        return true;
      }
    }
    return false;
  }
  void release(uint64_t off, uint64_t n) {
    n = (n + 255) & ~(uint64_t)255;
    used -= n;
    auto it = free_blocks.emplace(off, n).first;
    auto next = std::next(it);
    if (next != free_blocks.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_blocks.erase(next);
    }
    if (it != free_blocks.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_blocks.erase(it);
      }
    }
  }
};

enum ReqState { REQ_QUEUED = 0, REQ_RUNNING = 1, REQ_DONE = 2, REQ_FINISHING = 3 };

// A server's device memory: chunks allocated on demand up to the budget, sizes doubling from
// 4 GiB (so the whole budget takes a handful of allocations, made only as the load requires;
// one allocation of the whole 200 GB budget took 5.3 s at creation). First fit within a
// chunk; a request never spans chunks, and one larger than the next chunk size gets a chunk of
// its own, released when it finishes. Empty chunks are released when a new chunk would
// exceed the budget.
struct ChunkArena {
  struct Chunk {
    uint8_t* base = nullptr;
    uint64_t size = 0;
    bool dedicated = false;
    Arena fit;
  };
  std::map<int, Chunk> chunks;
  int next_id = 0;
  uint64_t budget = 0, held = 0, next_size = uint64_t(4) << 30;

  uint64_t used() const {
    uint64_t u = 0;
    for (const auto& kv : chunks) u += kv.second.fit.used;
    return u;
  }
  // SC_OK with *chunk, *off and *ptr set; 1 when it does not fit now; SC_ERR_* on failure.
  int alloc(uint64_t n, int* chunk, uint64_t* off, uint8_t** ptr) {
    n = (n + 255) & ~(uint64_t)255;
    for (auto& kv : chunks) {
      if (kv.second.dedicated) continue;
      if (kv.second.fit.alloc(n, *off)) {
        *chunk = kv.first;
        *ptr = kv.second.base + *off;
        return SC_OK;
      }
    }
    const bool dedicated = n > next_size;
    uint64_t size = dedicated ? ((n + (uint64_t(256) << 20) - 1) & ~((uint64_t(256) << 20) - 1)) : next_size;
    if (held + size > budget) {
      release_empty();
      if (held + size > budget) {
        if (!dedicated && held + n <= budget) {
          size = (budget - held) & ~(uint64_t)255;  // the budget's remainder as a last chunk
        } else {
          return 1;
        }
      }
    }
    Chunk c;
    const hipError_t e = hipMalloc(&c.base, size);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      return 1;
    }
    if (e != hipSuccess) return fail(SC_ERR_HIP, std::string("server arena chunk: ") + hipGetErrorString(e));
    c.size = size;
    c.dedicated = dedicated;
    c.fit.init(c.base, size);
    held += size;
    if (!dedicated && next_size < (uint64_t(64) << 30)) next_size *= 2;
    const int id = next_id++;
    auto& slot = chunks[id] = c;
    if (!slot.fit.alloc(n, *off)) return fail(SC_ERR_HIP, "server arena: new chunk too small");
    *chunk = id;
    *ptr = slot.base + *off;
    return SC_OK;
  }
  void release(int chunk, uint64_t off, uint64_t n) {
    auto it = chunks.find(chunk);
    if (it == chunks.end()) return;
    it->second.fit.release(off, n);
    if (it->second.dedicated && it->second.fit.used == 0) drop(it);
  }
  void release_empty() {
    for (auto it = chunks.begin(); it != chunks.end();) {
      auto cur = it++;
      if (cur->second.fit.used == 0) drop(cur);
    }
  }
  void drop(std::map<int, Chunk>::iterator it) {
    (void)hipFree(it->second.base);
    held -= it->second.size;
    chunks.erase(it);
  }
  void free_all() {
    for (auto& kv : chunks) (void)hipFree(kv.second.base);
    chunks.clear();
    held = 0;
  }
};

struct Request {
  int64_t id = 0;
  int fidx = 0, L = 0, npts = 0, num_vars = 0, mode = SC_CHALLENGE_SHA3;
  CompDesc desc;
  sc_input input;
  int state = REQ_QUEUED;
  int slot = -1;
  int round = 0;  // next round to run
  uint32_t items = 4;  // tile of its stage entries (pairs per thread)
  int64_t wide = -1;   // rounds with at most this many pairs point-parallel; -1: default
  uint64_t arena_off = 0, arena_bytes = 0;
  int arena_chunk = -1;
  ProofState* result = nullptr;  // pinned host copy of the final state
};

struct StepRecord {
  hipEvent_t ev = nullptr;
  std::vector<int64_t> finishing;
  int staging = -1;
};

}  // namespace

struct sc_server {
  sc_context* ctx = nullptr;
  hipStream_t stream = nullptr;  // the server's own: two servers on one device run side by side
  int max_active = 0;
  uint32_t items = 4;
  ChunkArena arena;
  InstDesc* d_insts = nullptr;       // [max_active]
  CompDesc* d_comps = nullptr;       // [max_active]
  ProofState* d_states = nullptr;    // [max_active]
  StageEntry* d_entries = nullptr;   // [2 * max_active]: a step's lists and tails
  uint32_t* d_partials = nullptr;    // [max_tiles][MAX_POINTS][MAX_LIMBS]
  uint64_t max_tiles = 0;
  std::vector<int> free_slots;
  std::deque<int64_t> queue;
  std::map<int64_t, Request> reqs;
  std::vector<int64_t> slot_req;
  int64_t next_id = 1;
  // pinned staging ring for descriptors and entry lists
  struct Staging {
    InstDesc* insts = nullptr;
    CompDesc* comps = nullptr;
    StageEntry* entries = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
  };
  std::vector<Staging> staging;
  std::deque<StepRecord> inflight;
  uint64_t steps = 0;
  size_t max_inflight = 3;  // steps enqueued ahead of the device before the host waits
  int policy = SC_SERVER_ALL_RUNNING;
};

namespace {

void free_server(sc_server* sv) {
  if (sv->stream) (void)hipStreamSynchronize(sv->stream);
  for (auto& kv : sv->reqs)
    if (kv.second.result) (void)hipHostFree(kv.second.result);
  for (auto& st : sv->staging) {
    if (st.insts) (void)hipHostFree(st.insts);
    if (st.comps) (void)hipHostFree(st.comps);
    if (st.entries) (void)hipHostFree(st.entries);
    if (st.ev) (void)hipEventDestroy(st.ev);
  }
  for (auto& r : sv->inflight)
    if (r.ev) (void)hipEventDestroy(r.ev);
  sv->arena.free_all();
  if (sv->d_insts) (void)hipFree(sv->d_insts);
  if (sv->d_comps) (void)hipFree(sv->d_comps);
  if (sv->d_states) (void)hipFree(sv->d_states);
  if (sv->d_entries) (void)hipFree(sv->d_entries);
  if (sv->d_partials) (void)hipFree(sv->d_partials);
  if (sv->stream) (void)hipStreamDestroy(sv->stream);
  delete sv;
}

uint64_t request_bytes(const Request& r) {
  const uint64_t words = (uint64_t)r.desc.num_tables * ((uint64_t)1 << r.num_vars) * r.L;
  auto align = [](uint64_t x) { return (x + 255) & ~(uint64_t)255; };
  return align(words * 4) + align(words / 2 * 4) + align(words / 4 * 4 + 4);
}

// Relative device time of a request's current round: pairs x tables x points x a per-width
// weight (limbs, and 16 at 255 bits, whose products are compute-bound: MI355X, a*b*c + d*e,
// 3.5 ms for 2^24 pairs over all rounds at 255 bits against 2.2 ms for 2^26 at 64 bits).
// About 0.25-0.3 ms of device time in these units on MI355X (a*b*c + d*e round 0: 2.7e9 at
// 255 bits, num_vars 24, 0.65 ms; 6.7e8 at 64 bits, num_vars 26, ~0.2 ms).
constexpr double kStepFill = 1e9;

double round_work(const Request& r) {
  const double pairs = std::ldexp(1.0, r.num_vars - r.round - 1);
  return pairs * r.desc.num_tables * r.npts * (r.L == 8 ? 16.0 : (double)r.L);
}

// A free staging buffer, waiting for the oldest if all are in flight.
int take_staging(sc_server* sv) {
  for (size_t i = 0; i < sv->staging.size(); ++i) {
    auto& st = sv->staging[i];
    if (!st.busy) return (int)i;
    if (hipEventQuery(st.ev) == hipSuccess) {
      st.busy = false;
      return (int)i;
    }
  }
  // all busy: wait on each in turn
  for (size_t i = 0; i < sv->staging.size(); ++i) {
    if (hipEventSynchronize(sv->staging[i].ev) == hipSuccess) {
      sv->staging[i].busy = false;
      return (int)i;
    }
  }
  return -1;
}

// Retire steps whose work finished: finishing requests become done, their memory and slot
// return to the pools. wait = block for every in-flight step.
int retire(sc_server* sv, bool wait) {
  while (!sv->inflight.empty()) {
    StepRecord& rec = sv->inflight.front();
    if (wait) {
      SC_HIP(hipEventSynchronize(rec.ev));
    } else if (hipEventQuery(rec.ev) != hipSuccess) {
      break;
    }
    for (int64_t id : rec.finishing) {
      Request& r = sv->reqs[id];
      r.state = REQ_DONE;
      sv->arena.release(r.arena_chunk, r.arena_off, r.arena_bytes);
      sv->free_slots.push_back(r.slot);
      sv->slot_req[r.slot] = 0;
      r.slot = -1;
    }
    (void)hipEventDestroy(rec.ev);
    sv->inflight.pop_front();
  }
  return SC_OK;
}

// A request's largest round (round 0, point-parallel: the most tiles) must fit the step's
// partials buffer on its own, or it could never run: raise its tile until it does, or refuse.
int fit_request_tiles(const sc_server* sv, Request& r) {
  const uint64_t pairs0 = ((uint64_t)1 << r.num_vars) >> 1;
  while (StagePlan::tiles(pairs0, true, r.items, r.npts, 0) > sv->max_tiles) {
    if (r.items >= 64) return fail(SC_ERR_UNSUPPORTED, "request too large for one server step (num_vars)");
    r.items *= 2;
  }
  return SC_OK;
}

}  // namespace

extern "C" {

int sc_server_set_request_schedule(sc_server* sv, int64_t handle, int32_t items_per_thread, int64_t max_pairs) {
  if (sv == nullptr) return fail(SC_ERR_ARG, "server must not be NULL");
  auto it = sv->reqs.find(handle);
  if (it == sv->reqs.end() || it->second.state == REQ_DONE) return fail(SC_ERR_ARG, "no such queued or running request");
  if (items_per_thread < 0 || items_per_thread > 64) return fail(SC_ERR_ARG, "items_per_thread must be in [0, 64]");
  Request r = it->second;
  if (items_per_thread > 0) r.items = (uint32_t)items_per_thread;
  r.wide = max_pairs < 0 ? -1 : max_pairs;
  const int rc = fit_request_tiles(sv, r);
  if (rc != SC_OK) return rc;
  it->second.items = r.items;
  it->second.wide = r.wide;
  return SC_OK;
}

int sc_server_set_policy(sc_server* sv, int32_t policy) {
  if (sv == nullptr) return fail(SC_ERR_ARG, "server must not be NULL");
  if (policy != SC_SERVER_ALL_RUNNING && policy != SC_SERVER_SHORTEST_FIRST)
    return fail(SC_ERR_ARG, "policy must be SC_SERVER_ALL_RUNNING or SC_SERVER_SHORTEST_FIRST");
  sv->policy = policy;
  return SC_OK;
}

int sc_server_create(sc_context* ctx, uint64_t hbm_budget_bytes, int32_t max_active, sc_server** out) {
  if (ctx == nullptr || out == nullptr) return fail(SC_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (max_active < 1 || max_active > 4096) return fail(SC_ERR_ARG, "max_active must be in [1, 4096]");
  if (hbm_budget_bytes < (1u << 20)) return fail(SC_ERR_ARG, "hbm_budget_bytes too small");
  SC_HIP(hipSetDevice(ctx->device));
  sc_server* sv = new (std::nothrow) sc_server();
  if (sv == nullptr) return fail(SC_ERR_MEMORY, "host allocation failed");
  sv->ctx = ctx;
  sv->max_active = max_active;
  sv->max_tiles = (uint64_t)1 << 20;
#define SC_HIP_S(expr)                                                                        \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) {                                                                   \
      free_server(sv);                                                                        \
      return fail(_e == hipErrorOutOfMemory ? SC_ERR_MEMORY : SC_ERR_HIP,                     \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                         \
    }                                                                                         \
  } while (0)
  sv->arena.budget = hbm_budget_bytes;
  SC_HIP_S(hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking));
  SC_HIP_S(hipMalloc(&sv->d_insts, sizeof(InstDesc) * max_active));
  SC_HIP_S(hipMalloc(&sv->d_comps, sizeof(CompDesc) * max_active));
  SC_HIP_S(hipMalloc(&sv->d_states, sizeof(ProofState) * max_active));
  SC_HIP_S(hipMalloc(&sv->d_entries, sizeof(StageEntry) * 2 * max_active));
  SC_HIP_S(hipMalloc(&sv->d_partials, sv->max_tiles * MAX_POINTS * MAX_LIMBS * 4));
  // a step takes one buffer for its admissions and one per width; with 3 steps in flight,
  // 16 buffers keep the host from waiting on a buffer that is still queued behind earlier
  // steps' work (a wait that would also hold up any other server the host loop drives)
  sv->staging.resize(16);
  for (auto& st : sv->staging) {
    SC_HIP_S(hipHostMalloc(&st.insts, sizeof(InstDesc) * max_active, hipHostMallocDefault));
    SC_HIP_S(hipHostMalloc(&st.comps, sizeof(CompDesc) * max_active, hipHostMallocDefault));
    SC_HIP_S(hipHostMalloc(&st.entries, sizeof(StageEntry) * 2 * max_active, hipHostMallocDefault));
    SC_HIP_S(hipEventCreateWithFlags(&st.ev, hipEventDisableTiming));
  }
#undef SC_HIP_S
  sv->slot_req.assign(max_active, 0);
  for (int i = max_active - 1; i >= 0; --i) sv->free_slots.push_back(i);
  *out = sv;
  return SC_OK;
}

int sc_server_submit(sc_server* sv, int32_t bit_width, const sc_composition* comp, int32_t num_vars,
                     const sc_input* input, int32_t challenge_mode, int64_t* handle) {
  if (sv == nullptr || input == nullptr || handle == nullptr) return fail(SC_ERR_ARG, "NULL argument");
  const int fidx = field_index(bit_width);
  if (fidx < 0) return fail(SC_ERR_ARG, "bit_width must be 32, 64, 128 or 256");
  if (num_vars < 1 || num_vars > MAX_ROUNDS) return fail(SC_ERR_UNSUPPORTED, "num_vars must be in [1, 40]");
  if (input->kind != SC_INPUT_AFFINE && input->kind != SC_INPUT_UNIFORM)
    return fail(SC_ERR_ARG, "served requests take generated inputs (affine or uniform)");
  if (challenge_mode != SC_CHALLENGE_FIXED && challenge_mode != SC_CHALLENGE_SHA3)
    return fail(SC_ERR_ARG, "challenge_mode must be 0 (fixed) or 1 (sha3)");
  Request r;
  int degree = 0;
  int rc = build_comp_any(fidx, comp, sv->ctx->host_fc[fidx], r.desc, degree);
  if (rc != SC_OK) return rc;
  r.id = sv->next_id++;
  r.fidx = fidx;
  r.L = launchers_for(fidx).limbs;
  r.npts = r.desc.npts;
  r.num_vars = num_vars;
  r.mode = challenge_mode;
  r.input = *input;
  if (request_bytes(r) > sv->arena.budget) return fail(SC_ERR_MEMORY, "request exceeds the HBM budget");
  r.items = sv->items;
  rc = fit_request_tiles(sv, r);
  if (rc != SC_OK) return rc;
  SC_HIP(hipHostMalloc(&r.result, sizeof(ProofState), hipHostMallocDefault));
  sv->reqs[r.id] = r;
  sv->queue.push_back(r.id);
  *handle = r.id;
  return SC_OK;
}

}  // extern "C"

namespace {

// One scheduling step: retire finished steps, admit queued requests that fit (FIFO), and
// enqueue one round of every running request (batched stage launches per field width). With
// max_inflight steps already enqueued it waits for the oldest, or (wait = false) returns with
// *stepped = 0 and enqueues nothing.
int server_step(sc_server* sv, bool wait, int32_t* n_running, int32_t* n_queued, int32_t* stepped) {
  if (sv == nullptr) return fail(SC_ERR_ARG, "server must not be NULL");
  sc_context* ctx = sv->ctx;
  SC_HIP(hipSetDevice(ctx->device));
  if (stepped) *stepped = 0;
  int rc = retire(sv, false);
  if (rc != SC_OK) return rc;
  while (sv->inflight.size() >= sv->max_inflight) {
    if (!wait) {
      int running = 0;
      for (int slot = 0; slot < sv->max_active; ++slot)
        if (sv->slot_req[slot] != 0) ++running;
      if (n_running) *n_running = running;
      if (n_queued) *n_queued = (int32_t)sv->queue.size();
      return SC_OK;
    }
    SC_HIP(hipEventSynchronize(sv->inflight.front().ev));
    rc = retire(sv, false);
    if (rc != SC_OK) return rc;
  }
  if (stepped) *stepped = 1;
  hipStream_t s = sv->stream;
  // admit: the step's admissions share one staging buffer, entry k for the k-th request
  int adm_si = -1;
  int adm_k = 0;
  while (!sv->queue.empty() && !sv->free_slots.empty()) {
    Request& r = sv->reqs[sv->queue.front()];
    const uint64_t bytes = request_bytes(r);
    uint64_t off = 0;
    int chunk = -1;
    uint8_t* base = nullptr;
    const int arc = sv->arena.alloc(bytes, &chunk, &off, &base);
    if (arc != SC_OK && arc != 1) return arc;
    if (arc == 1) {
      // does not fit now; with nothing running it never will (the device holds other memory)
      bool running = false;
      for (int64_t id : sv->slot_req) running = running || id != 0;
      if (!running && sv->inflight.empty())
        return fail(SC_ERR_MEMORY, "server arena: the device cannot hold the next request");
      break;
    }
    sv->queue.pop_front();
    r.slot = sv->free_slots.back();
    sv->free_slots.pop_back();
    sv->slot_req[r.slot] = r.id;
    r.arena_off = off;
    r.arena_bytes = bytes;
    r.arena_chunk = chunk;
    r.state = REQ_RUNNING;
    r.round = 0;
    const Launchers& ln = launchers_for(r.fidx);
    const uint64_t n0 = (uint64_t)1 << r.num_vars;
    const uint64_t words = (uint64_t)r.desc.num_tables * n0 * r.L;
    auto align = [](uint64_t x) { return (x + 255) & ~(uint64_t)255; };
    if (adm_si < 0) {
      adm_si = take_staging(sv);
      if (adm_si < 0) return fail(SC_ERR_HIP, "no staging buffer");
    }
    auto& st = sv->staging[adm_si];
    const int k = adm_k++;  // at most max_active admissions per step: one per free slot
    InstDesc& d = st.insts[k];
    d.input = (const uint32_t*)base;
    d.buf_a = (uint32_t*)(base + align(words * 4));
    d.buf_b = (uint32_t*)(base + align(words * 4) + align(words / 2 * 4));
    d.n0 = n0;
    d.num_vars = r.num_vars;
    d.challenge_mode = r.mode;
    d.comp = sv->d_comps + r.slot;
    d.st = sv->d_states + r.slot;
    st.comps[k] = r.desc;
    SC_HIP(hipMemcpyAsync(sv->d_insts + r.slot, &st.insts[k], sizeof(InstDesc), hipMemcpyHostToDevice, s));
    SC_HIP(hipMemcpyAsync(sv->d_comps + r.slot, &st.comps[k], sizeof(CompDesc), hipMemcpyHostToDevice, s));
    SC_HIP(hipMemsetAsync(sv->d_states + r.slot, 0, sizeof(ProofState), s));
    SC_HIP(hipEventRecord(st.ev, s));
    st.busy = true;
    FillArgs f;
    f.tables = (uint32_t*)base;
    f.n = n0;
    f.stride = n0;
    f.num_tables = r.desc.num_tables;
    f.kind = r.input.kind == SC_INPUT_AFFINE ? FILL_AFFINE : FILL_UNIFORM;
    f.seed = r.input.seed;
    f.selector_mask = r.input.selector_mask;
    f.instance = r.input.instance;
    f.pad0 = 0;
    f.fc = ctx->d_fc[r.fidx];
    SC_HIP(ln.fill(s, f));
  }
  // Shortest remaining work first: the request with the least work left; every request whose
  // current round costs at most a quarter of that one's (they delay it little); and further
  // requests in order of work left while the step holds at most max(kStepFill, 1.25x the
  // first one's round), so a step led by a nearly finished request still fills the device.
  // The rest wait. Otherwise one round of every running request.
  std::vector<char> take(sv->max_active, 1);
  if (sv->policy == SC_SERVER_SHORTEST_FIRST) {
    std::vector<std::pair<double, int>> order;  // (current round's work, slot); work left ~2x
    for (int slot = 0; slot < sv->max_active; ++slot) {
      const int64_t id = sv->slot_req[slot];
      take[slot] = 0;
      if (id == 0 || sv->reqs[id].state != REQ_RUNNING) continue;
      order.emplace_back(round_work(sv->reqs[id]), slot);
    }
    std::sort(order.begin(), order.end());
    if (!order.empty()) {
      const double lead_work = order[0].first;
      const double budget = std::max(kStepFill, 1.25 * lead_work);
      double acc = 0.0;
      for (size_t k = 0; k < order.size(); ++k) {
        const double w = order[k].first;
        if (k == 0 || w <= 0.25 * lead_work || acc + w <= budget) {
          take[order[k].second] = 1;
          acc += w;
        }
      }
    }
  }
  // one round of every request taken, grouped by width
  StepRecord rec;
  int staged = 0;
  bool any = false;
  for (int fidx = 0; fidx < 4; ++fidx) {
    std::vector<int> slots;
    int maxT = 0, maxP = 0;
    for (int slot = 0; slot < sv->max_active; ++slot) {
      const int64_t id = sv->slot_req[slot];
      if (id == 0) continue;
      const Request& r = sv->reqs[id];
      if (r.state != REQ_RUNNING || r.fidx != fidx || !take[slot]) continue;
      slots.push_back(slot);
      maxT = std::max(maxT, r.desc.num_tables);
      maxP = std::max(maxP, r.npts);
    }
    if (slots.empty()) continue;
    const int tmax = pick_tmax(maxT), pmax = maxP <= 4 ? 4 : 8;
    // oldest request first: one that did not fit a step's partials gets them in the next
    std::sort(slots.begin(), slots.end(),
              [&](int a2, int b2) { return sv->slot_req[a2] < sv->slot_req[b2]; });
    // one round of each request (per-pair or point-parallel), then tail entries: requests
    // whose current round has at most the tail's pairs finish all their remaining rounds
    StagePlan plan;
    std::vector<StageEntry> tails;
    for (int slot : slots) {
      const Request& r = sv->reqs[sv->slot_req[slot]];
      const uint64_t pairs = ((uint64_t)1 << r.num_vars) >> (r.round + 1);
      const int64_t tail_pairs = default_tail_pairs(r.L, pick_tmax(r.desc.num_tables) > 8);
      if ((int64_t)pairs <= tail_pairs) {
        tails.push_back(StageEntry{(uint32_t)slot, (uint32_t)r.round, 0u, 0u, 1u});
        continue;
      }
      const int64_t wide_pairs = r.wide < 0 ? default_instance_wide_pairs(r.L, tmax, r.npts) : r.wide;
      const bool wide = (int64_t)pairs <= wide_pairs;
      if (plan.partial_tiles() + StagePlan::tiles(pairs, wide, r.items, r.npts, (uint32_t)r.round) > sv->max_tiles)
        continue;  // runs in a later step (alone it fits: fit_request_tiles)
      plan.add((uint32_t)slot, (uint32_t)r.round, pairs, r.desc.num_tables, wide, r.items, r.npts);
    }
    if (plan.size() == 0 && tails.empty()) continue;
    std::vector<StageEntry> ents;
    plan.layout(ents);
    const uint32_t n_plan = (uint32_t)ents.size();
    ents.insert(ents.end(), tails.begin(), tails.end());
    any = true;
    const int si = take_staging(sv);
    if (si < 0) return fail(SC_ERR_HIP, "no staging buffer");
    auto& st = sv->staging[si];
    std::copy(ents.begin(), ents.end(), st.entries);
    SC_HIP(hipMemcpyAsync(sv->d_entries, st.entries, sizeof(StageEntry) * ents.size(), hipMemcpyHostToDevice, s));
    SC_HIP(hipEventRecord(st.ev, s));
    st.busy = true;
    ++staged;
    BatchArgs a;
    a.tile_base = 0;
    a.pad0 = 0;
    a.insts = sv->d_insts;
    a.chunk_shift = default_fold_chunk_shift(launchers_for(fidx).limbs);
    a.partials = sv->d_partials;
    a.fc = ctx->d_fc[fidx];
    const Launchers& ln = launchers_for(fidx);
    const uint32_t n_pass = (uint32_t)plan.pass.size(), n_eval = (uint32_t)plan.eval.size();
    rc = enqueue_stage(ln, s, a, tmax, pmax, sv->d_entries, n_pass, n_eval, (uint32_t)plan.fold.size(),
                       plan.pass_tiles, plan.eval_tiles, plan.fold_tiles);
    if (rc != SC_OK) return rc;
    if (!tails.empty()) {
      a.entries = sv->d_entries + n_plan;
      a.n_entries = (uint32_t)tails.size();
      SC_HIP(ln.batch_tail(s, tmax, pmax, a, a.n_entries));
    }
    // each request is in exactly one of [pass][eval] or the tails (fold entries repeat eval's)
    for (size_t k = 0; k < ents.size(); ++k) {
      if (k >= n_pass + n_eval && k < n_plan) continue;
      const StageEntry& e = ents[k];
      Request& r = sv->reqs[sv->slot_req[e.inst]];
      r.round = k < n_plan ? r.round + 1 : r.num_vars;
      if (r.round >= r.num_vars) {
        SC_HIP(hipMemcpyAsync(r.result, sv->d_states + r.slot, sizeof(ProofState), hipMemcpyDeviceToHost, s));
        r.state = REQ_FINISHING;
        rec.finishing.push_back(r.id);
      }
    }
  }
  if (any) {
    SC_HIP(hipEventCreateWithFlags(&rec.ev, hipEventDisableTiming));
    SC_HIP(hipEventRecord(rec.ev, s));
    sv->inflight.push_back(rec);
    ++sv->steps;
  }
  int running = 0;
  for (int slot = 0; slot < sv->max_active; ++slot)
    if (sv->slot_req[slot] != 0) ++running;
  if (n_running) *n_running = running;
  if (n_queued) *n_queued = (int32_t)sv->queue.size();
  return SC_OK;
}

}  // namespace

extern "C" {

int sc_server_step(sc_server* sv, int32_t* n_running, int32_t* n_queued) {
  return server_step(sv, true, n_running, n_queued, nullptr);
}

int sc_server_try_step(sc_server* sv, int32_t* n_running, int32_t* n_queued, int32_t* stepped) {
  return server_step(sv, false, n_running, n_queued, stepped);
}

int sc_server_poll(sc_server* sv, int64_t handle, int32_t* state) {
  if (sv == nullptr || state == nullptr) return fail(SC_ERR_ARG, "NULL argument");
  int rc = retire(sv, false);
  if (rc != SC_OK) return rc;
  auto it = sv->reqs.find(handle);
  if (it == sv->reqs.end()) return fail(SC_ERR_ARG, "unknown request handle");
  *state = it->second.state == REQ_FINISHING ? REQ_RUNNING : it->second.state;
  return SC_OK;
}

int sc_server_drain(sc_server* sv) {
  if (sv == nullptr) return fail(SC_ERR_ARG, "server must not be NULL");
  return retire(sv, true);
}

int sc_server_result(sc_server* sv, int64_t handle, uint32_t* round_polys, uint32_t* challenges,
                     uint64_t* checksum) {
  if (sv == nullptr) return fail(SC_ERR_ARG, "server must not be NULL");
  int rc = retire(sv, false);
  if (rc != SC_OK) return rc;
  auto it = sv->reqs.find(handle);
  if (it == sv->reqs.end()) return fail(SC_ERR_ARG, "unknown request handle");
  Request& r = it->second;
  if (r.state != REQ_DONE) return fail(SC_ERR_ARG, "request not finished");
  const ProofState& st = *r.result;
  uint64_t sum = 0;
  for (int j = 0; j < r.num_vars; ++j) {
    for (int t = 0; t < r.npts; ++t) {
      for (int k = 0; k < r.L; ++k) {
        const uint32_t w = st.polys[j][t][k];
        sum += (uint64_t)(k + 1) * w;
        if (round_polys) round_polys[((size_t)j * r.npts + t) * r.L + k] = w;
      }
    }
    if (challenges && j + 1 < r.num_vars)
      for (int k = 0; k < r.L; ++k) challenges[(size_t)j * r.L + k] = st.r[j][k];
  }
  if (checksum) *checksum = sum;
  (void)hipHostFree(r.result);
  sv->reqs.erase(it);
  return SC_OK;
}

int sc_server_info(sc_server* sv, uint64_t* arena_used, uint64_t* arena_size, uint64_t* steps) {
  if (sv == nullptr) return fail(SC_ERR_ARG, "server must not be NULL");
  if (arena_used) *arena_used = sv->arena.used();
  if (arena_size) *arena_size = sv->arena.held;
  if (steps) *steps = sv->steps;
  return SC_OK;
}

int sc_server_destroy(sc_server* sv) {
  if (sv == nullptr) return SC_OK;
  (void)hipSetDevice(sv->ctx->device);
  free_server(sv);
  return SC_OK;
}

}  // extern "C"
