/* C ABI of libsumcheck_hip.so, the MI355X (gfx950) SumCheck prover.
 *
 * It replaces the reference's CUDA executable build/field_sumcheck_cuda
 * (cuda/src/field_sumcheck.cu), which the reference's Python drivers run as a subprocess
 * (benchmarks/field_sweep/compare_triton_cuda.py:108-136,
 * experiments/h100/tools/validate_cuda_vs_registry.py:66-91), and the in-process Triton round
 * loop (benchmarks/field_sweep/triton_sweep.py:521-816). Python binds it with ctypes; plain
 * pointers and sizes only.
 *
 * Conventions
 *   - Every function returns SC_OK (0) or a nonzero SC_ERR_* code; sc_last_error() then
 *     describes the failure (thread-local). The reference calls std::exit instead
 *     (cuda/src/field_sumcheck.cu:18-26, 1112-1173).
 *   - Field elements are little-endian arrays of L 32-bit limbs, L = bit_width / 32, for
 *     bit_width 32 (p = 2^32-5), 64 (Goldilocks), 128 (2^127-1), 256 (BLS12-381 scalar)
 *     (src/zkduel/fields.py:258-294). "Montgomery" means x*2^(32L) mod p.
 *   - Tables are element-major: table v, entry i at words [(v*2^num_vars + i)*L, +L).
 *   - The caller owns host buffers; a context owns device buffers it allocates.
 *   - Calls on one context are serialized by the caller; contexts are independent.
 */
#ifndef SUMCHECK_HIP_H
#define SUMCHECK_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SC_OK 0
#define SC_ERR_ARG 1         /* invalid argument (ValueError on the Python side) */
#define SC_ERR_UNSUPPORTED 2 /* outside the compiled limits (ValueError) */
#define SC_ERR_HIP 3         /* HIP runtime failure (RuntimeError) */
#define SC_ERR_MEMORY 4      /* device allocation failed (MemoryError) */

/* Composition sum_k coeff_k * prod_m table[term_vars[term_offsets[k] + m]],
 * m < term_lens[k]: the reference's WorkloadSpec encoding
 * (benchmarks/field_sweep/workload_specs.py:6-17, src/zkduel/_expr.py:87-110). Coefficients are
 * canonical field elements in [0, p) as L limbs each (negative integers are taken mod p, as
 * reference.py:17-20 does). Degree d = max term length <= 7 (cuda/src/field_sumcheck.cu:29);
 * num_tables <= 32; num_terms <= 32. */
typedef struct sc_composition {
  int32_t num_tables;
  int32_t num_terms;
  const int32_t* term_lens;    /* [num_terms] */
  const int32_t* term_offsets; /* [num_terms] */
  const int32_t* term_vars;    /* [sum of term_lens], values in [0, num_tables) */
  const uint32_t* coeffs;      /* [num_terms][L] canonical */
} sc_composition;

#define SC_INPUT_HOST_CANONICAL 0 /* data: host tables, canonical values in [0, p) */
#define SC_INPUT_HOST_MONT 1      /* data: host tables, Montgomery form */
#define SC_INPUT_DEVICE_MONT 2    /* data: device tables, Montgomery form, borrowed */
#define SC_INPUT_AFFINE 3         /* generated: seed + 7919(v+1) + 104729(o+1) (field_sumcheck.cu:123-128) */
#define SC_INPUT_UNIFORM 4        /* generated: counter-based uniform in [0,p); selector tables in {0,1} */
/* Limb-major tables, the reference's --layout limb order [table][limb][entry]
 * (cuda/src/field_sumcheck.cu:165-190), Montgomery form: transposed once on the device into
 * the element-major layout the kernels stream; results are the same bits. */
#define SC_INPUT_HOST_MONT_LIMB 5   /* data: host tables */
#define SC_INPUT_DEVICE_MONT_LIMB 6 /* data: device tables, read only */

typedef struct sc_input {
  int32_t kind;
  uint32_t instance;      /* UNIFORM: instance number in the generator key */
  const void* data;       /* HOST_* / DEVICE_MONT */
  uint64_t seed;          /* AFFINE: the value seed (the reference passes seed + rounds); UNIFORM: key seed */
  uint64_t selector_mask; /* UNIFORM: bit v set -> table v drawn from {0, 1} */
} sc_input;

#define SC_CHALLENGE_FIXED 0    /* r_j = 1 (cuda/src/field_sumcheck.cu:1058) */
#define SC_CHALLENGE_SHA3 1     /* SHA3-256 transcript (field_sumcheck.cu:857-907, triton_sweep.py:474-484) */
#define SC_CHALLENGE_EXPLICIT 2 /* caller's list, canonical (reference.py:44-55 `challenges`) */

typedef struct sc_context sc_context;
typedef struct sc_proof sc_proof;

const char* sc_last_error(void);
const char* sc_version(void);

int sc_context_create(int device, sc_context** out);
int sc_context_destroy(sc_context* ctx);
int sc_context_sync(sc_context* ctx);

/* Allocates a proof's device buffers and prepares its input (generation, upload, Montgomery
 * encode): the reference's per-case setup (field_sumcheck.cu:1178-1221). num_vars in [1, 40]. */
int sc_proof_create(sc_context* ctx, int32_t bit_width, const sc_composition* comp,
                    int32_t num_vars, const sc_input* input, sc_proof** out);

/* Tunables of the round schedule: pass grid cap and the pair count at or below which the
 * remaining rounds run in the single-workgroup tail kernel (0 keeps the default), and
 * fused_finalize = 1 to let each pass kernel's last block finish the round instead of a
 * separate finalize launch. */
int sc_proof_set_schedule(sc_proof* proof, int32_t max_pass_blocks, int64_t tail_pairs,
                          int32_t fused_finalize);

/* Kernel choice per round: rounds with at most max_pairs pairs run as a fold launch plus the
 * point-parallel evaluation (one lane per pair and evaluated point) instead of the per-pair
 * pass kernel with the fold fused in; 0 never, a negative value restores the default (every
 * round for interpreted compositions of more than 8 tables and for interpreted 255-bit
 * compositions of degree 3 or more, rounds below 2^20 pairs for other interpreted and below
 * 2^18 for compiled 255-bit compositions, none otherwise). Results do not depend on it. No
 * counterpart in the reference, which runs every round as an eval, a reduce and a separate
 * fold kernel (field_sumcheck.cu:614-694). */
int sc_proof_set_point_parallel(sc_proof* proof, int64_t max_pairs);

/* Layout of the intermediate (folded) tables: chunks of 2^chunk_log2 entries of every table
 * side by side (chunk q of table v after chunk q of tables 0..v-1), so a wave's pairs of all
 * tables form one contiguous run; a chunk at least a table's size makes them table-major.
 * chunk_log2 in [2, 63] (a folding round reads 4 consecutive entries of each table); 0 restores
 * the default (2^7 up to 128 bits, table-major at 256 bits).
 * Results do not depend on it. No
 * counterpart in the reference, whose tables are table-major throughout
 * (field_sumcheck.cu:130-190). */
int sc_proof_set_fold_chunk(sc_proof* proof, int32_t chunk_log2);

/* Enqueues one whole proof (all rounds, transcript on the device) on the context's stream
 * and returns without waiting: the reference's run_once (field_sumcheck.cu:1012-1088).
 * explicit_r: [num_vars-1][L] canonical, only for SC_CHALLENGE_EXPLICIT (else NULL). */
int sc_proof_run(sc_proof* proof, int32_t challenge_mode, const uint32_t* explicit_r);

/* Captures the same sequence into a hipGraph; later sc_proof_run calls with the same mode
 * replay it. */
int sc_proof_capture(sc_proof* proof, int32_t challenge_mode, const uint32_t* explicit_r);

/* Waits for the enqueued runs; *ms = device time of the last run (hipEvents), may be NULL. */
int sc_proof_wait(sc_proof* proof, float* ms);

/* Results of the last run: round_polys [num_vars][degree+1][L] Montgomery g_j(0..d);
 * challenges [num_vars-1][L] Montgomery r_j; checksum = sum over rounds, points and limbs of
 * (i+1)*limb_i (field_sumcheck.cu:1043-1049). Any output pointer may be NULL. */
int sc_proof_result(sc_proof* proof, uint32_t* round_polys, uint32_t* challenges,
                    uint64_t* checksum);

/* Repeated proving with the host's work overlapped with the device's: sc_proof_run_async
 * enqueues the proof as sc_proof_run does and, after it on the same stream, a copy of its
 * round polynomials and challenges into one of two pinned host slots the proof owns; *slot
 * names it. sc_proof_slot_result waits for that copy only (not for later runs) and returns the
 * run's results as sc_proof_result does. A slot is reused by the run after next: read a run's
 * results before enqueuing two more. No counterpart in the reference, whose driver waits for
 * each run (field_sumcheck.cu:1012-1088). */
int sc_proof_run_async(sc_proof* proof, int32_t challenge_mode, const uint32_t* explicit_r, int32_t* slot);
int sc_proof_slot_result(sc_proof* proof, int32_t slot, uint32_t* round_polys, uint32_t* challenges,
                         uint64_t* checksum);

/* Copies the prepared input tables (Montgomery, element-major) to host memory. */
int sc_proof_tables(sc_proof* proof, uint32_t* out);

/* Describes the prepared proof: degree, number of rounds run before the tail kernel, device
 * bytes held, and the pair count at or below which rounds run point-parallel (the value in
 * effect, see sc_proof_set_point_parallel). Any pointer may be NULL. */
int sc_proof_info(sc_proof* proof, int32_t* degree, int32_t* pass_rounds, uint64_t* device_bytes,
                  int64_t* point_parallel_pairs);

/* Compositions with compiled kernels (a*b, a*b*c, a*b*c + d*e) run them by default; enable = 0
 * forces the interpreted kernels (same results). Returns the compiled composition id that
 * matches (0: none), or a negative error code. */
int sc_proof_use_compiled(sc_proof* proof, int32_t enable);

/* Per-kernel timing: when enabled, eager runs record a hipEvent pair around every pass
 * kernel and around the tail kernel, on the stream they run on. sc_proof_kernel_times then
 * returns, for the last run, pass_ms[j] for rounds j < pass_rounds and *tail_ms. */
int sc_proof_set_profiling(sc_proof* proof, int32_t enable);
int sc_proof_kernel_times(sc_proof* proof, float* pass_ms, int32_t capacity, float* tail_ms);
/* As sc_proof_kernel_times, plus finalize_ms[j]: round j's finalize launch (block partials
 * summed, g_j stored, r_j derived on the device), the reference's per-round challenge phase
 * (field_sumcheck.cu:1053-1066) in round-sum timing. Any pointer may be NULL. */
int sc_proof_round_times(sc_proof* proof, float* pass_ms, float* finalize_ms, int32_t capacity,
                         float* tail_ms);

/* Final-point step of the last run (no counterpart in the reference, which stops after the
 * last round polynomial, field_sumcheck.cu:1057): r_final [L] = the challenge derived from the last
 * round's polynomial by the same transcript rule, final_values [T][L] = every table folded at
 * (r_0, ..., r_final), i.e. f_v at the random point. Montgomery limbs. A verifier checks
 * g_last(r_final) = C(final_values). SC_ERR_UNSUPPORTED after a run with explicit challenges. */
int sc_proof_final(sc_proof* proof, uint32_t* r_final, uint32_t* final_values);

int sc_proof_destroy(sc_proof* proof);

/* One-shot: create, run, wait, read results, destroy. */
int sc_prove(sc_context* ctx, int32_t bit_width, const sc_composition* comp, int32_t num_vars,
             const sc_input* input, int32_t challenge_mode, const uint32_t* explicit_r,
             uint32_t* round_polys, uint32_t* challenges, uint64_t* checksum);

/* sc_prove without allocation: no counterpart in the reference. For callers that cannot
 * allocate device memory: code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc. It allocates nothing (no hipMalloc/hipFree, no host copy) and enqueues the
 * whole proof on the context's stream without waiting (sc_context_sync waits). All buffers are
 * device memory of the caller:
 *   tables       [T][2^num_vars][L] Montgomery, element-major, read only (SC_INPUT_DEVICE_MONT)
 *   round_polys  [num_vars][degree+1][L] Montgomery, out
 *   challenges   [num_vars-1][L] Montgomery, out, may be NULL
 *   checksum     one uint64, out, may be NULL (as sc_prove)
 * challenge_mode SC_CHALLENGE_FIXED or SC_CHALLENGE_SHA3 (else SC_ERR_UNSUPPORTED). The
 * intermediate tables, block partials, descriptor and round state live in one fixed device
 * array in the library, one per GPU, held until the process ends: the intermediate tables
 * (0.75x the input) must fit in SC_NOALLOC_FOLD_BYTES, else SC_ERR_UNSUPPORTED. Every call on
 * a GPU uses that GPU's array: calls that overlap on one GPU (other contexts or streams) race;
 * calls on one context are ordered by its stream. */
#define SC_NOALLOC_FOLD_BYTES (128u << 20)
int sc_prove_noalloc(sc_context* ctx, int32_t bit_width, const sc_composition* comp, int32_t num_vars,
                     const uint32_t* tables, int32_t challenge_mode, uint32_t* round_polys,
                     uint32_t* challenges, uint64_t* checksum);

/* Elementwise device field operations on n elements (host buffers, [n][L]):
 * op 0: a*b*R^-1 (Montgomery product), 1: a+b, 2: a-b, 3: a -> a*R (encode),
 * 4: a -> a*R^-1 (decode). The device twin of src/zkduel/fields.py:63-115, and what the
 * reference's tests/test_uint256.py exercises. b may be NULL for ops 3 and 4.
 * ops 5 and 6 run the prover kernels' internal arithmetic path on Montgomery inputs and return
 * Montgomery outputs: 5: a*b*R^-1 (as op 0), 6: a + a*(b - a)*R^-1 (a fold with r = a). */
int sc_field_op(sc_context* ctx, int32_t bit_width, int32_t op, uint64_t n, const uint32_t* a,
                const uint32_t* b, uint32_t* out);

/* Device transcript challenge for one round: g_mont [npts][L] Montgomery limbs in;
 * r_mont [L] out (field_sumcheck.cu:857-907 computed on the GPU). */
int sc_transcript_challenge(sc_context* ctx, int32_t bit_width, int32_t round, const uint32_t* g_mont,
                            int32_t npts, uint32_t* r_mont);

/* Batched proving (no counterpart in the reference, which proves one instance per run):
 * n independent instances of one width, each with its own composition, num_vars and
 * generated input (SC_INPUT_AFFINE or SC_INPUT_UNIFORM), advance round by round; each stage
 * is one launch over all active instances' tiles plus one launch finishing their rounds.
 * Results per instance are the same bits as sc_prove on that instance. */
typedef struct sc_batch sc_batch;
int sc_batch_create(sc_context* ctx, int32_t bit_width, int32_t n, const sc_composition* comps,
                    const int32_t* num_vars, const sc_input* inputs, sc_batch** out);
/* Pairs per thread in a tile of every instance (tile = 256 * items pairs; 256 / (points
 * evaluated per pair: degree + 1 in round 0, degree later) * items pairs in point-parallel
 * rounds); default 4. */
int sc_batch_set_tile(sc_batch* batch, int32_t items_per_thread);
/* Instance rounds with at most max_pairs pairs run point-parallel (a fold launch plus an
 * evaluation launch with one lane per pair and point, as sc_proof_set_point_parallel for one
 * proof), for every instance; 0 never, negative each instance's default (see
 * sc_batch_set_schedules). Results do not depend on it.
 * No counterpart in the reference (one instance per run). */
int sc_batch_set_point_parallel(sc_batch* batch, int64_t max_pairs);
/* Per-instance schedules (no counterpart in the reference), arrays of the batch's size, any
 * of them NULL to keep the current values: items_per_thread[i] the instance's tile ([1, 64];
 * 0 keeps it), max_pairs[i] the pair count at or below which its rounds run point-parallel
 * (negative: the default, every round when the batch holds an instance of more than 8 tables
 * or, at 256 bits, when the instance's degree is 3 or more; otherwise never), tail_pairs[i]
 * the pair count at or below which its remaining rounds run in the batch's tail launch (0:
 * the default, 512 pairs, 64 at 256 bits, 1 with more than 8 tables). Results do not depend
 * on them. sc_batch_instance_schedule reads back one instance's tile, the point-parallel
 * threshold in effect and the first round of its tail. */
int sc_batch_set_schedules(sc_batch* batch, const int32_t* items_per_thread, const int64_t* max_pairs,
                           const int64_t* tail_pairs);
int sc_batch_instance_schedule(sc_batch* batch, int32_t instance, int32_t* items_per_thread,
                               int64_t* max_pairs, int32_t* tail_round);
/* challenge_mode SC_CHALLENGE_FIXED or SC_CHALLENGE_SHA3; enqueues without waiting. */
int sc_batch_run(sc_batch* batch, int32_t challenge_mode);
int sc_batch_capture(sc_batch* batch, int32_t challenge_mode);
int sc_batch_wait(sc_batch* batch, float* ms);
int sc_batch_result(sc_batch* batch, int32_t instance, uint32_t* round_polys, uint32_t* challenges,
                    uint64_t* checksum);
int sc_batch_info(sc_batch* batch, uint64_t* device_bytes, int32_t* stages, uint64_t* tiles);
int sc_batch_destroy(sc_batch* batch);

/* Serving (no counterpart in the reference): continuous batching of proof requests. Requests
 * are queued by sc_server_submit and admitted at step boundaries, first in first out, while a
 * slot (max_active) and HBM (hbm_budget_bytes: their tables take 1.75x the input) are free.
 * The server's device memory is allocated on demand, in chunks of doubling size from 4 GiB up
 * to the budget (a request larger than the next chunk gets one of its own, released when it
 * finishes), so creation allocates no table memory. Each sc_server_step enqueues one round of
 * every running request, each at its own round, oldest first, using the batched stage kernels
 * grouped by width; tables and transcript state stay on the device between steps. A request's
 * result is the same bits as sc_prove on it. Results are read once (sc_server_result frees
 * the request). Poll states: 0 queued, 1 running, 2 done. sc_server_info: arena_used = bytes
 * of the admitted requests, arena_size = device bytes the server holds. Each server runs on a
 * stream of its own, so two servers on one context (say one for short requests, one for long
 * ones) run side by side on the GPU. Calls on one server must be serialized. */
typedef struct sc_server sc_server;
int sc_server_create(sc_context* ctx, uint64_t hbm_budget_bytes, int32_t max_active, sc_server** out);
int sc_server_submit(sc_server* server, int32_t bit_width, const sc_composition* comp, int32_t num_vars,
                     const sc_input* input, int32_t challenge_mode, int64_t* handle);
int sc_server_step(sc_server* server, int32_t* n_running, int32_t* n_queued);
/* sc_server_step that does not wait for in-flight steps (no counterpart in the reference; for
 * a host loop driving several servers): with the server's in-flight steps (3) already enqueued
 * it enqueues nothing and sets *stepped = 0; otherwise it steps as sc_server_step and sets
 * *stepped = 1. A step takes one of the server's 16 pinned staging buffers for its admissions
 * and one per field width; only when all 16 are still queued (three steps in flight, each
 * over all four widths with admissions) does it wait for the oldest to be consumed. */
int sc_server_try_step(sc_server* server, int32_t* n_running, int32_t* n_queued, int32_t* stepped);
/* Which running requests a step advances (no counterpart in the reference):
 * SC_SERVER_ALL_RUNNING (the default) one round of every running request;
 * SC_SERVER_SHORTEST_FIRST the request with the least work left, every request whose current
 * round costs at most a quarter of that one's, and further requests in order of work left
 * while the step holds about 0.25 ms of device work (work = pairs x tables x points x a
 * per-width weight). Shortest-first lets the smaller requests of a burst finish first instead
 * of all finishing together; under sustained overload a large request can wait long. Results
 * do not depend on the policy. */
#define SC_SERVER_ALL_RUNNING 0
#define SC_SERVER_SHORTEST_FIRST 1
int sc_server_set_policy(sc_server* server, int32_t policy);
int sc_server_poll(sc_server* server, int64_t handle, int32_t* state);
/* Schedule of a queued or running request, as sc_batch_set_schedules for one instance (tile and
 * point-parallel threshold; the tail threshold is the default). A request's largest round
 * must fit one step's partials (2^20 tiles): submit and this call raise its tile to fit, and
 * refuse a request that cannot (SC_ERR_UNSUPPORTED). No counterpart in the reference. */
int sc_server_set_request_schedule(sc_server* server, int64_t handle, int32_t items_per_thread,
                                   int64_t max_pairs);
int sc_server_drain(sc_server* server);
int sc_server_result(sc_server* server, int64_t handle, uint32_t* round_polys, uint32_t* challenges,
                     uint64_t* checksum);
int sc_server_info(sc_server* server, uint64_t* arena_used, uint64_t* arena_size, uint64_t* steps);
int sc_server_destroy(sc_server* server);

/* Measured sustained lane rates (operations per second) of the VALU operations the field
 * arithmetic is built from: rates[0] v_mad_u64_u32 (32x32+64), rates[1] v_mul_hi_u32,
 * rates[2] 64-bit add, rates[3] v_fma_f64; and streaming HBM bandwidth in bytes/s:
 * rates[4] read-only, rates[5] copy (read + write counted), rates[6] 2 bytes read per byte
 * written (a fold round's mix; read + write counted). The INT32 roofline of the wide fields is
 * priced against rates[0]. n <= 7. */
int sc_calibrate(sc_context* ctx, double* rates, int32_t n);

#ifdef __cplusplus
}
#endif

#endif /* SUMCHECK_HIP_H */
