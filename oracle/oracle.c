/* C twin of the SumCheck oracle — TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
 *
 * Same algorithm as oracle/sumcheck_ref.py (which restates src/zkduel/reference.py:44-110,
 * the transcript of benchmarks/field_sweep/triton_sweep.py:474-484 and the checksum of
 * cuda/src/field_sumcheck.cu:1043-1049), multi-threaded with pthreads so it reaches
 * num_vars 16..26. It shares no code with the product: field products use separated
 * operand scanning (SOS) Montgomery reduction, and Keccak derives its rotation offsets and
 * round constants from FIPS 202's formulas (Algorithms 2-5) at start-up instead of tables.
 * Pinned by tests/test_oracle_c.py against the Python oracle and the reference's goldens.
 *
 * Values are held in Montgomery form, L little-endian 32-bit limbs per element.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 8
#define MAXT 32
#define MAXP 8

typedef struct {
  int L;
  uint32_t p[MAXL];
  uint32_t npr;      /* -p^-1 mod 2^32 */
  uint32_t r2[MAXL]; /* R^2 mod p */
  uint32_t one[MAXL];
  int bits;
} field_t;

static const char* P_HEX[4] = {
    "fffffffb", "ffffffff00000001", "7fffffffffffffffffffffffffffffff",
    "73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001"};

static int hexval(char c) { return c <= '9' ? c - '0' : (c | 32) - 'a' + 10; }

/* ---------------------------------------------------------------- multiprecision helpers */

static int geq(const uint32_t* a, const uint32_t* b, int L) {
  for (int i = L - 1; i >= 0; --i)
    if (a[i] != b[i]) return a[i] > b[i];
  return 1;
}

static void sub_in(uint32_t* a, const uint32_t* b, int L) {
  int64_t br = 0;
  for (int i = 0; i < L; ++i) {
    int64_t d = (int64_t)a[i] - b[i] - br;
    br = d < 0;
    a[i] = (uint32_t)d;
  }
}

static void fadd(const field_t* f, uint32_t* o, const uint32_t* a, const uint32_t* b) {
  uint64_t c = 0;
  uint32_t t[MAXL + 1];
  for (int i = 0; i < f->L; ++i) {
    c += (uint64_t)a[i] + b[i];
    t[i] = (uint32_t)c;
    c >>= 32;
  }
  if (c || geq(t, f->p, f->L)) sub_in(t, f->p, f->L);
  memcpy(o, t, 4 * f->L);
}

static void fsub(const field_t* f, uint32_t* o, const uint32_t* a, const uint32_t* b) {
  uint32_t t[MAXL];
  int64_t br = 0;
  for (int i = 0; i < f->L; ++i) {
    int64_t d = (int64_t)a[i] - b[i] - br;
    br = d < 0;
    t[i] = (uint32_t)d;
  }
  if (br) {
    uint64_t c = 0;
    for (int i = 0; i < f->L; ++i) {
      c += (uint64_t)t[i] + f->p[i];
      t[i] = (uint32_t)c;
      c >>= 32;
    }
  }
  memcpy(o, t, 4 * f->L);
}

/* SOS: full product, then L word-by-word reductions, then one conditional subtraction. */
static void fmul(const field_t* f, uint32_t* o, const uint32_t* a, const uint32_t* b) {
  const int L = f->L;
  uint32_t t[2 * MAXL + 2];
  memset(t, 0, sizeof(t));
  for (int i = 0; i < L; ++i) {
    uint64_t c = 0;
    for (int j = 0; j < L; ++j) {
      c += (uint64_t)a[i] * b[j] + t[i + j];
      t[i + j] = (uint32_t)c;
      c >>= 32;
    }
    for (int k = i + L; c; ++k) {
      c += t[k];
      t[k] = (uint32_t)c;
      c >>= 32;
    }
  }
  for (int i = 0; i < L; ++i) {
    uint32_t m = t[i] * f->npr;
    uint64_t c = 0;
    for (int j = 0; j < L; ++j) {
      c += (uint64_t)m * f->p[j] + t[i + j];
      t[i + j] = (uint32_t)c;
      c >>= 32;
    }
    for (int k = i + L; c; ++k) {
      c += t[k];
      t[k] = (uint32_t)c;
      c >>= 32;
    }
  }
  uint32_t* r = t + L; /* value < 2p, with t[2L] the carry word */
  if (t[2 * L] || geq(r, f->p, L)) sub_in(r, f->p, L);
  memcpy(o, r, 4 * L);
}

static void field_init(field_t* f, int bw) {
  memset(f, 0, sizeof(*f));
  int idx = bw == 32 ? 0 : bw == 64 ? 1 : bw == 128 ? 2 : 3;
  f->L = bw / 32;
  const char* h = P_HEX[idx];
  int n = (int)strlen(h);
  for (int k = 0; k < n; ++k) {
    int bit = 4 * (n - 1 - k);
    f->p[bit / 32] |= (uint32_t)hexval(h[k]) << (bit % 32);
  }
  /* nprime by Newton iteration: x <- x (2 - p x) mod 2^32 */
  uint32_t x = 1;
  for (int i = 0; i < 5; ++i) x *= 2u - f->p[0] * x;
  f->npr = (uint32_t)(0u - x);
  /* R mod p and R^2 mod p by doubling 1 */
  uint32_t v[MAXL] = {1};
  for (int i = 0; i < 32 * f->L; ++i) fadd(f, v, v, v);
  memcpy(f->one, v, 4 * f->L);
  for (int i = 0; i < 32 * f->L; ++i) fadd(f, v, v, v);
  memcpy(f->r2, v, 4 * f->L);
  f->bits = 32 * f->L;
  while (!((f->p[(f->bits - 1) / 32] >> ((f->bits - 1) % 32)) & 1)) --f->bits;
}

static void to_mont(const field_t* f, uint32_t* o, const uint32_t* canon) { fmul(f, o, canon, f->r2); }

/* ---------------------------------------------------------------- Keccak from FIPS 202 */

static int RHO[25];
static uint64_t RC[24];

static int rc_bit(int t) { /* Algorithm 5 */
  if (t % 255 == 0) return 1;
  unsigned R = 0x80; /* bits R[0..7] stored MSB-first: R = 10000000 */
  for (int i = 1; i <= t % 255; ++i) {
    /* R = 0 || R; R[0] ^= R[8]; R[4] ^= R[8]; R[5] ^= R[8]; R[6] ^= R[8]; truncate to 8 */
    unsigned r8 = R & 1;
    R >>= 1;
    if (r8) R ^= 0x80 | 0x08 | 0x04 | 0x02;
  }
  return (R >> 7) & 1;
}

static void keccak_tables(void) {
  static int done = 0;
  if (done) return;
  memset(RHO, 0, sizeof(RHO));
  int x = 1, y = 0; /* Algorithm 2 */
  for (int t = 0; t < 24; ++t) {
    RHO[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
    int nx = y, ny = (2 * x + 3 * y) % 5;
    x = nx;
    y = ny;
  }
  for (int ir = 0; ir < 24; ++ir) { /* Algorithm 6, iota */
    uint64_t rc = 0;
    for (int j = 0; j <= 6; ++j)
      if (rc_bit(j + 7 * ir)) rc |= (uint64_t)1 << ((1 << j) - 1);
    RC[ir] = rc;
  }
  done = 1;
}

static uint64_t rol(uint64_t v, int s) { return s ? (v << s) | (v >> (64 - s)) : v; }

static void keccak_p(uint64_t* A) {
  for (int ir = 0; ir < 24; ++ir) {
    uint64_t C[5], D[5], B[25];
    for (int x = 0; x < 5; ++x) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
    for (int x = 0; x < 5; ++x) D[x] = C[(x + 4) % 5] ^ rol(C[(x + 1) % 5], 1);
    for (int i = 0; i < 25; ++i) A[i] ^= D[i % 5];
    for (int x = 0; x < 5; ++x)
      for (int y = 0; y < 5; ++y) B[y + 5 * ((2 * x + 3 * y) % 5)] = rol(A[x + 5 * y], RHO[x + 5 * y]);
    for (int x = 0; x < 5; ++x)
      for (int y = 0; y < 5; ++y) A[x + 5 * y] = B[x + 5 * y] ^ (~B[(x + 1) % 5 + 5 * y] & B[(x + 2) % 5 + 5 * y]);
    A[0] ^= RC[ir];
  }
}

void oracle_sha3_256(const uint8_t* msg, uint64_t len, uint8_t* out) {
  keccak_tables();
  uint64_t A[25];
  memset(A, 0, sizeof(A));
  const int rate = 136;
  uint8_t block[136];
  uint64_t off = 0;
  for (;;) {
    uint64_t take = len - off < (uint64_t)rate ? len - off : (uint64_t)rate;
    memset(block, 0, sizeof(block));
    memcpy(block, msg + off, take);
    off += take;
    int last = take < (uint64_t)rate;
    if (last) {
      block[take] ^= 0x06;
      block[rate - 1] ^= 0x80;
    }
    for (int i = 0; i < rate / 8; ++i) {
      uint64_t w = 0;
      for (int k = 0; k < 8; ++k) w |= (uint64_t)block[8 * i + k] << (8 * k);
      A[i] ^= w;
    }
    keccak_p(A);
    if (last) break;
  }
  for (int i = 0; i < 32; ++i) out[i] = (uint8_t)(A[i / 8] >> (8 * (i % 8)));
}

/* ---------------------------------------------------------------- inputs */

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static uint64_t key_of(uint64_t seed, uint32_t inst, uint32_t table) {
  uint64_t k = mix64(seed ^ 0x5851F42D4C957F2Dull);
  return mix64(k ^ (((uint64_t)inst << 32) | table));
}

static uint64_t draw(uint64_t key, uint64_t ctr) { return mix64(key + ctr * 0x9E3779B97F4A7C15ull); }

static void uniform_canon(const field_t* f, uint64_t key, uint64_t idx, uint32_t* out) {
  uint32_t w[MAXL];
  for (int attempt = 0; attempt < 32; ++attempt) {
    for (int k = 0; k < (f->L + 1) / 2; ++k) {
      uint64_t d = draw(key, idx * 128 + attempt * 4 + k);
      w[2 * k] = (uint32_t)d;
      if (2 * k + 1 < f->L) w[2 * k + 1] = (uint32_t)(d >> 32);
    }
    int top = f->bits - 32 * (f->L - 1);
    if (top < 32) w[f->L - 1] &= (1u << top) - 1;
    if (!geq(w, f->p, f->L)) {
      memcpy(out, w, 4 * f->L);
      return;
    }
  }
  sub_in(w, f->p, f->L);
  memcpy(out, w, 4 * f->L);
}

/* ---------------------------------------------------------------- the prover */

typedef struct {
  const field_t* f;
  int T, K, npts, fold;
  const int* exps; /* [K][T] */
  const uint32_t* coeff; /* [K][L] Montgomery */
  const uint32_t* src; uint32_t* dst;
  uint64_t src_n, lo, hi;
  const uint32_t* r;
  uint32_t acc[MAXP][MAXL];
} job_t;

static void compose(const job_t* j, const uint32_t* x /* [T][L] */, uint32_t* out) {
  const field_t* f = j->f;
  const int L = f->L;
  uint32_t total[MAXL] = {0}, prod[MAXL];
  for (int k = 0; k < j->K; ++k) {
    memcpy(prod, j->coeff + (size_t)k * L, 4 * L);
    for (int v = 0; v < j->T; ++v)
      for (int e = 0; e < j->exps[k * j->T + v]; ++e) fmul(f, prod, prod, x + (size_t)v * L);
    fadd(f, total, total, prod);
  }
  memcpy(out, total, 4 * L);
}

static void* run_job(void* arg) {
  job_t* j = (job_t*)arg;
  const field_t* f = j->f;
  const int L = f->L, T = j->T;
  uint32_t e[MAXT * MAXL], o[MAXT * MAXL], d[MAXT * MAXL], c[MAXL], tmp[MAXL];
  memset(j->acc, 0, sizeof(j->acc));
  for (uint64_t i = j->lo; i < j->hi; ++i) {
    for (int v = 0; v < T; ++v) {
      const uint32_t* s = j->src + (size_t)v * j->src_n * L;
      uint32_t* ev = e + v * L;
      uint32_t* ov = o + v * L;
      if (j->fold) {
        /* f'[m] = f[2m] + r (f[2m+1] - f[2m]) for m = 2i, 2i+1 */
        for (int h = 0; h < 2; ++h) {
          uint32_t* out = h ? ov : ev;
          const uint32_t* a = s + (4 * i + 2 * h) * L;
          const uint32_t* b = a + L;
          fsub(f, tmp, b, a);
          fmul(f, tmp, tmp, j->r);
          fadd(f, out, a, tmp);
          memcpy(j->dst + ((size_t)v * (j->src_n / 2) + 2 * i + h) * L, out, 4 * L);
        }
      } else {
        memcpy(ev, s + (2 * i) * L, 4 * L);
        memcpy(ov, s + (2 * i + 1) * L, 4 * L);
      }
      fsub(f, d + v * L, ov, ev);
    }
    compose(j, e, c);
    fadd(f, j->acc[0], j->acc[0], c);
    if (j->npts > 1) {
      compose(j, o, c);
      fadd(f, j->acc[1], j->acc[1], c);
    }
    for (int t = 2; t < j->npts; ++t) {
      for (int v = 0; v < T; ++v) fadd(f, o + v * L, o + v * L, d + v * L);
      compose(j, o, c);
      fadd(f, j->acc[t], j->acc[t], c);
    }
  }
  return NULL;
}

static void challenge_from_digest(const field_t* f, const uint8_t* dg, uint32_t* r_mont) {
  /* digest as little-endian integer mod p, then Montgomery: Horner over bytes, high first */
  uint32_t acc[MAXL] = {0}, b[MAXL];
  for (int i = 31; i >= 0; --i) {
    for (int s = 0; s < 8; ++s) fadd(f, acc, acc, acc);
    memset(b, 0, sizeof(b));
    b[0] = dg[i];
    if (!geq(b, f->p, f->L)) fadd(f, acc, acc, b);
    else { sub_in(b, f->p, f->L); fadd(f, acc, acc, b); }
  }
  to_mont(f, r_mont, acc);
}

/* Returns 0 on success. input_kind: 0 affine (value seed), 1 uniform, 2 host canonical.
 * challenge_mode: 0 fixed (r = 1), 1 SHA3 transcript, 2 explicit canonical list. */
int oracle_prove(int bit_width, int T, int K, const int32_t* term_lens, const int32_t* term_offsets,
                 const int32_t* term_vars, const uint32_t* coeffs_canon, int num_vars, int input_kind,
                 uint64_t seed, uint32_t instance, uint64_t selector_mask, const uint32_t* host_canon,
                 int challenge_mode, const uint32_t* explicit_canon, int threads,
                 uint32_t* out_polys, uint32_t* out_r, uint32_t* out_final) {
  if (bit_width != 32 && bit_width != 64 && bit_width != 128 && bit_width != 256) return 1;
  if (T < 1 || T > MAXT || K < 1 || num_vars < 1 || num_vars > 40) return 1;
  field_t f;
  field_init(&f, bit_width);
  const int L = f.L;
  int* exps = calloc((size_t)K * T, sizeof(int));
  uint32_t* coeff = calloc((size_t)K * L, 4);
  int degree = 0;
  for (int k = 0; k < K; ++k) {
    if (term_lens[k] > degree) degree = term_lens[k];
    for (int m = 0; m < term_lens[k]; ++m) exps[k * T + term_vars[term_offsets[k] + m]]++;
    to_mont(&f, coeff + (size_t)k * L, coeffs_canon + (size_t)k * L);
  }
  const int npts = degree + 1;
  if (npts > MAXP) { free(exps); free(coeff); return 1; }
  const uint64_t n0 = (uint64_t)1 << num_vars;
  uint32_t* cur = malloc((size_t)T * n0 * L * 4);
  uint32_t* nxt = malloc((size_t)T * (n0 / 2 + 1) * L * 4);
  if (!cur || !nxt) { free(exps); free(coeff); free(cur); free(nxt); return 2; }
  for (int v = 0; v < T; ++v) {
    uint64_t key = key_of(seed, instance, (uint32_t)v);
    for (uint64_t i = 0; i < n0; ++i) {
      uint32_t x[MAXL] = {0};
      if (input_kind == 0) {
        uint64_t val = seed + 7919ull * (v + 1) + 104729ull * (i + 1);
        if (L == 1) val %= f.p[0];
        x[0] = (uint32_t)val;
        if (L > 1) x[1] = (uint32_t)(val >> 32);
      } else if (input_kind == 1) {
        if ((selector_mask >> v) & 1) x[0] = (uint32_t)(draw(key, i * 128) & 1);
        else uniform_canon(&f, key, i, x);
      } else {
        memcpy(x, host_canon + ((size_t)v * n0 + i) * L, 4 * L);
      }
      to_mont(&f, cur + ((size_t)v * n0 + i) * L, x);
    }
  }
  if (threads < 1) threads = 1;
  job_t* jobs = calloc((size_t)threads, sizeof(job_t));
  pthread_t* tid = calloc((size_t)threads, sizeof(pthread_t));
  uint32_t r[MAXL];
  memcpy(r, f.one, 4 * L);
  uint64_t src_n = n0;
  for (int j = 0; j < num_vars; ++j) {
    const int fold = j > 0;
    const uint64_t pairs = n0 >> (j + 1);
    int nt = (uint64_t)threads > pairs ? (int)pairs : threads;
    if (nt < 1) nt = 1;
    for (int t = 0; t < nt; ++t) {
      job_t* jb = &jobs[t];
      jb->f = &f; jb->T = T; jb->K = K; jb->npts = npts; jb->fold = fold;
      jb->exps = exps; jb->coeff = coeff; jb->src = cur; jb->dst = nxt; jb->src_n = src_n;
      jb->lo = pairs * t / nt; jb->hi = pairs * (t + 1) / nt; jb->r = r;
      pthread_create(&tid[t], NULL, run_job, jb);
    }
    uint32_t g[MAXP][MAXL];
    memset(g, 0, sizeof(g));
    for (int t = 0; t < nt; ++t) {
      pthread_join(tid[t], NULL);
      for (int p = 0; p < npts; ++p) fadd(&f, g[p], g[p], jobs[t].acc[p]);
    }
    if (fold) {
      uint32_t* sw = cur; cur = nxt; nxt = sw;
      src_n /= 2;
    }
    for (int p = 0; p < npts; ++p) memcpy(out_polys + ((size_t)j * npts + p) * L, g[p], 4 * L);
    if (j + 1 < num_vars) {
      if (challenge_mode == 1) {
        uint8_t msg[32 + 4 + 4 * MAXP * MAXL], dg[32];
        size_t len = 0;
        memcpy(msg, "zkduel:field-sweep:v1", 21);
        len = 21;
        for (int k = 0; k < 4; ++k) msg[len++] = (uint8_t)((uint32_t)j >> (8 * k));
        for (int p = 0; p < npts; ++p)
          for (int l = 0; l < L; ++l)
            for (int k = 0; k < 4; ++k) msg[len++] = (uint8_t)(g[p][l] >> (8 * k));
        oracle_sha3_256(msg, len, dg);
        challenge_from_digest(&f, dg, r);
      } else if (challenge_mode == 2) {
        to_mont(&f, r, explicit_canon + (size_t)j * L);
      } else {
        memcpy(r, f.one, 4 * L);
      }
      memcpy(out_r + (size_t)j * L, r, 4 * L);
    }
  }
  if (out_final) {
    /* the tables after the last round hold one pair each: report f[0], f[1] per table */
    for (int v = 0; v < T; ++v) memcpy(out_final + (size_t)v * 2 * L, cur + (size_t)v * src_n * L, 8 * L);
  }
  free(jobs); free(tid); free(exps); free(coeff); free(cur); free(nxt);
  return 0;
}

/* Elementwise Montgomery product, for pinning fmul against big integers. */
void oracle_mont_mul(int bit_width, uint64_t n, const uint32_t* a, const uint32_t* b, uint32_t* out) {
  field_t f;
  field_init(&f, bit_width);
  for (uint64_t i = 0; i < n; ++i) fmul(&f, out + i * f.L, a + i * f.L, b + i * f.L);
}

/* ---------------------------------------------------------------- claim and final point */

typedef struct {
  const field_t* f;
  int T, K;
  const int* exps;
  const uint32_t* coeff;
  uint64_t seed, selector_mask;
  uint32_t instance;
  uint64_t lo, hi; /* pairs */
  const uint32_t* r0;
  uint32_t* folded; /* [T][n0/2][L] */
  uint64_t half;
  uint32_t claim[MAXL];
} claim_job_t;

static void gen_mont(const field_t* f, uint64_t seed, uint32_t inst, uint64_t sel_mask, int v, uint64_t i,
                     uint32_t* out) {
  uint32_t x[MAXL] = {0};
  const uint64_t key = key_of(seed, inst, (uint32_t)v);
  if ((sel_mask >> v) & 1) x[0] = (uint32_t)(draw(key, i * 128) & 1);
  else uniform_canon(f, key, i, x);
  to_mont(f, out, x);
}

static void* run_claim_job(void* arg) {
  claim_job_t* j = (claim_job_t*)arg;
  const field_t* f = j->f;
  const int L = f->L;
  uint32_t a[MAXT * MAXL], b[MAXT * MAXL], c[MAXL], tmp[MAXL];
  job_t shape; /* compose() reads only f, T, K, exps and coeff */
  memset(&shape, 0, sizeof(shape));
  shape.f = f; shape.T = j->T; shape.K = j->K; shape.exps = j->exps; shape.coeff = j->coeff;
  memset(j->claim, 0, sizeof(j->claim));
  for (uint64_t m = j->lo; m < j->hi; ++m) {
    for (int v = 0; v < j->T; ++v) {
      gen_mont(f, j->seed, j->instance, j->selector_mask, v, 2 * m, a + v * L);
      gen_mont(f, j->seed, j->instance, j->selector_mask, v, 2 * m + 1, b + v * L);
      fsub(f, tmp, b + v * L, a + v * L);
      fmul(f, tmp, tmp, j->r0);
      fadd(f, j->folded + ((size_t)v * j->half + m) * L, a + v * L, tmp);
    }
    compose(&shape, a, c);
    fadd(f, j->claim, j->claim, c);
    compose(&shape, b, c);
    fadd(f, j->claim, j->claim, c);
  }
  return NULL;
}

typedef struct {
  const field_t* f;
  uint32_t* t; /* one table, n entries */
  uint64_t n;
  const uint32_t* point; /* [num_vars][L], point[0] already applied */
  int num_vars;
} mle_job_t;

static void* run_mle_job(void* arg) {
  mle_job_t* j = (mle_job_t*)arg;
  const field_t* f = j->f;
  const int L = f->L;
  uint32_t tmp[MAXL];
  uint64_t n = j->n;
  for (int k = 1; k < j->num_vars; ++k, n /= 2) {
    for (uint64_t m = 0; m < n / 2; ++m) {
      const uint32_t* a = j->t + 2 * m * L;
      const uint32_t* b = a + L;
      fsub(f, tmp, b, a);
      fmul(f, tmp, tmp, j->point + (size_t)k * L);
      fadd(f, j->t + m * L, a, tmp);
    }
  }
  return NULL;
}

/* For the uniform (counter-based) tables keyed by (seed, instance), selector tables in {0, 1}:
 * out_claim [L] = sum over the hypercube of the composition, and out_final [T][L] = every
 * table's multilinear extension at point_mont [num_vars][L] (variable 0 = the lowest index
 * bit, the first round's challenge), all Montgomery. Tables are generated on the fly and
 * folded once by the first coordinate, so memory is half the input. Returns 0 on success. */
int oracle_claim_final(int bit_width, int T, int K, const int32_t* term_lens, const int32_t* term_offsets,
                       const int32_t* term_vars, const uint32_t* coeffs_canon, int num_vars, uint64_t seed,
                       uint32_t instance, uint64_t selector_mask, const uint32_t* point_mont, int threads,
                       uint32_t* out_claim, uint32_t* out_final) {
  if (bit_width != 32 && bit_width != 64 && bit_width != 128 && bit_width != 256) return 1;
  if (T < 1 || T > MAXT || K < 1 || num_vars < 1 || num_vars > 40) return 1;
  field_t f;
  field_init(&f, bit_width);
  const int L = f.L;
  int* exps = calloc((size_t)K * T, sizeof(int));
  uint32_t* coeff = calloc((size_t)K * L, 4);
  for (int k = 0; k < K; ++k) {
    for (int m = 0; m < term_lens[k]; ++m) exps[k * T + term_vars[term_offsets[k] + m]]++;
    to_mont(&f, coeff + (size_t)k * L, coeffs_canon + (size_t)k * L);
  }
  const uint64_t half = (uint64_t)1 << (num_vars - 1);
  uint32_t* folded = malloc((size_t)T * half * L * 4);
  if (!folded) { free(exps); free(coeff); return 2; }
  if (threads < 1) threads = 1;
  int nt = (uint64_t)threads > half ? (int)half : threads;
  claim_job_t* jobs = calloc((size_t)nt, sizeof(claim_job_t));
  pthread_t* tid = calloc((size_t)(nt > T ? nt : T), sizeof(pthread_t));
  for (int t = 0; t < nt; ++t) {
    claim_job_t* jb = &jobs[t];
    jb->f = &f; jb->T = T; jb->K = K; jb->exps = exps; jb->coeff = coeff;
    jb->seed = seed; jb->selector_mask = selector_mask; jb->instance = instance;
    jb->lo = half * t / nt; jb->hi = half * (t + 1) / nt; jb->r0 = point_mont;
    jb->folded = folded; jb->half = half;
    pthread_create(&tid[t], NULL, run_claim_job, jb);
  }
  uint32_t claim[MAXL] = {0};
  for (int t = 0; t < nt; ++t) {
    pthread_join(tid[t], NULL);
    fadd(&f, claim, claim, jobs[t].claim);
  }
  memcpy(out_claim, claim, 4 * L);
  mle_job_t* mj = calloc((size_t)T, sizeof(mle_job_t));
  for (int v = 0; v < T; ++v) {
    mj[v].f = &f; mj[v].t = folded + (size_t)v * half * L; mj[v].n = half;
    mj[v].point = point_mont; mj[v].num_vars = num_vars;
    pthread_create(&tid[v], NULL, run_mle_job, &mj[v]);
  }
  for (int v = 0; v < T; ++v) {
    pthread_join(tid[v], NULL);
    memcpy(out_final + (size_t)v * L, folded + (size_t)v * half * L, 4 * L);
  }
  free(mj); free(jobs); free(tid); free(folded); free(exps); free(coeff);
  return 0;
}
