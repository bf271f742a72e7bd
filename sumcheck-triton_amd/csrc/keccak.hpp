// SHA3-256 (FIPS 202) for host and device, written from the specification.
//
// The transcript hashes at most a few hundred bytes per round. keccak_f1600 is the
// straightforward 24-round permutation over 25 64-bit words (host code and checks);
// keccak_f1600_wave is the device form, one state word per lane of a wave.
#pragma once

#include <stdint.h>

#include "field.hpp"

namespace sc {

SC_HD uint64_t rotl64(uint64_t x, int s) { return s == 0 ? x : ((x << s) | (x >> (64 - s))); }

SC_HD uint64_t keccak_rc(int i) {
  // Round constants of FIPS 202, Section 3.2.5 (the output of rc(t) folded into lanes).
  constexpr uint64_t RC[24] = {
      0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull,
      0x8000000080008000ull, 0x000000000000808Bull, 0x0000000080000001ull,
      0x8000000080008081ull, 0x8000000000008009ull, 0x000000000000008Aull,
      0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
      0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull,
      0x8000000000008003ull, 0x8000000000008002ull, 0x8000000000000080ull,
      0x000000000000800Aull, 0x800000008000000Aull, 0x8000000080008081ull,
      0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};
  return RC[i];
}

// Keccak-f[1600] on A[x + 5y].
SC_HD void keccak_f1600(uint64_t A[25]) {
  // rho offsets r[x][y] (FIPS 202 Table 2) stored at index x + 5y
  constexpr int R[25] = {0,  1,  62, 28, 27, 36, 44, 6,  55, 20, 3,  10, 43,
                         25, 39, 41, 45, 15, 21, 8,  18, 2,  61, 56, 14};
#pragma unroll 1
  for (int round = 0; round < 24; ++round) {
    uint64_t C[5], D[5], B[25];
#pragma unroll
    for (int x = 0; x < 5; ++x) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
#pragma unroll
    for (int x = 0; x < 5; ++x) D[x] = C[(x + 4) % 5] ^ rotl64(C[(x + 1) % 5], 1);
    // theta, rho and pi: B[y, 2x+3y] = rot(A[x,y] ^ D[x], r[x,y])
#pragma unroll
    for (int y = 0; y < 5; ++y) {
#pragma unroll
      for (int x = 0; x < 5; ++x) {
        const int nx = y, ny = (2 * x + 3 * y) % 5;
        B[nx + 5 * ny] = rotl64(A[x + 5 * y] ^ D[x], R[x + 5 * y]);
      }
    }
    // chi
#pragma unroll
    for (int y = 0; y < 5; ++y) {
#pragma unroll
      for (int x = 0; x < 5; ++x) {
        A[x + 5 * y] = B[x + 5 * y] ^ (~B[(x + 1) % 5 + 5 * y] & B[(x + 2) % 5 + 5 * y]);
      }
    }
    // iota
    A[0] ^= keccak_rc(round);
  }
}

#if defined(__HIPCC__)
// Keccak-f[1600] spread over one wave: lane x + 5y (< 25) holds A[x + 5y]. A round is two
// exchanges through 25 words of LDS (xs), each one 64-bit store per lane and a few 64-bit
// loads: theta reads columns x-1 and x+1 of every row (ten words) and forms
// D[x] = C[x-1] ^ rot(C[x+1], 1) directly; rho, pi and chi then read the theta output at the
// three lanes whose words pi moves to this lane and to its two row neighbours, each rotated
// by that source's rho offset. The wave runs in lockstep and LDS serves one wave's accesses
// in order, so wave-scope fences (which also keep the compiler from reordering the accesses)
// are the only synchronization. Lanes 25..63 follow along on copies of lane 0..24 data and
// never store.
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, 64);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, 64);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t rotl64_var(uint64_t x, int s) {
  return s == 0 ? x : ((x << s) | (x >> (64 - s)));
}

// The lane whose word pi moves to lane (X, Y): B[y, 2x+3y] = rot(A[x, y]), so lane (X, Y)
// receives from (x, y) = ((Y - 3X) * 3 mod 5, X).
__device__ __forceinline__ int keccak_pi_src(int lane) {
  const int x = lane % 5, y = lane / 5;
  return ((3 * ((y - 3 * x + 15) % 5)) % 5) + 5 * x;
}

__device__ __forceinline__ void keccak_f1600_wave(uint64_t& a, uint64_t* xs) {
  const bool owner = (threadIdx.x & 63) < 25;
  const int lane = (int)(threadIdx.x & 63) % 25;
  const int x = lane % 5, y = lane / 5;
  constexpr int R[25] = {0,  1,  62, 28, 27, 36, 44, 6,  55, 20, 3,  10, 43,
                         25, 39, 41, 45, 15, 21, 8,  18, 2,  61, 56, 14};
  const int cm = (x + 4) % 5, cp = (x + 1) % 5;
  const int s0 = keccak_pi_src(lane);
  const int s1 = keccak_pi_src((x + 1) % 5 + 5 * y);
  const int s2 = keccak_pi_src((x + 2) % 5 + 5 * y);
  const int r0 = R[s0], r1 = R[s1], r2 = R[s2];
  // round constant i sits in lane i; lane 0 reads round r's from lane r (a register read,
  // not a memory load on the round's critical path)
  const uint64_t rc = keccak_rc((int)(threadIdx.x & 63) < 24 ? (int)(threadIdx.x & 63) : 0);
  for (int round = 0; round < 24; ++round) {
    // theta: a ^= C[x-1] ^ rot(C[x+1], 1), both column parities read in one exchange
    if (owner) xs[lane] = a;
    wave_sync_lds();
    uint64_t c_m = 0, c_p = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      c_m ^= xs[cm + 5 * k];
      c_p ^= xs[cp + 5 * k];
    }
    const uint64_t t = a ^ c_m ^ rotl64(c_p, 1);
    wave_sync_lds();
    // rho, pi, chi: this lane's B and its row neighbours' B, read from their pi sources
    if (owner) xs[lane] = t;
    wave_sync_lds();
    const uint64_t b0 = rotl64_var(xs[s0], r0);
    const uint64_t b1 = rotl64_var(xs[s1], r1);
    const uint64_t b2 = rotl64_var(xs[s2], r2);
    wave_sync_lds();
    a = b0 ^ (~b1 & b2);
    // (readlane returns int: zero-extend each half, never sign-extend)
    const uint32_t rc_lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)rc, round);
    const uint32_t rc_hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(rc >> 32), round);
    const uint64_t rc_round = (uint64_t)rc_lo | ((uint64_t)rc_hi << 32);
    if ((threadIdx.x & 63) == 0) a ^= rc_round;  // iota (lane 0 holds A[0, 0])
  }
}
#endif

// SHA3-256 over a message already padded (pad10*1, SHA3 domain bits) into whole
// 136-byte blocks of 34 little-endian 32-bit words.
SC_HD void sha3_256_padded(const uint32_t* words, int nblocks, uint32_t digest[8]) {
  uint64_t A[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) A[i] = 0;
  for (int b = 0; b < nblocks; ++b) {
#pragma unroll
    for (int i = 0; i < 17; ++i) {
      A[i] ^= (uint64_t)words[34 * b + 2 * i] | ((uint64_t)words[34 * b + 2 * i + 1] << 32);
    }
    keccak_f1600(A);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    digest[2 * i] = (uint32_t)A[i];
    digest[2 * i + 1] = (uint32_t)(A[i] >> 32);
  }
}

// The per-round transcript message (cuda/src/field_sumcheck.cu:857-907,
// benchmarks/field_sweep/triton_sweep.py:474-484) is
//   "zkduel:field-sweep:v1" (21 bytes) || u32le(round) || u32le(each Montgomery limb of g).
// transcript_blocks gives its padded length in blocks; transcript_word(k) gives word k of
// the padded message, so the words can be produced in parallel (one per lane).
SC_HD int transcript_blocks(int nvals) { return (25 + 4 * nvals) / 136 + 1; }

SC_HD uint32_t transcript_word(int k, uint32_t round, const uint32_t* vals, int nvals, int nblocks) {
  uint32_t w;
  if (k < 5) {
    // "zkdu" "el:f" "ield" "-swe" "ep:v"
    w = k == 0 ? 0x75646B7Au : k == 1 ? 0x663A6C65u : k == 2 ? 0x646C6569u
        : k == 3 ? 0x6577732Du : 0x763A7065u;
  } else if (k <= 5 + nvals) {
    // bytes: '1' || round || vals: word k = high byte of x_{k-6} | x_{k-5} << 8 with
    // x_0 = round, x_i = vals[i-1]; the high byte before x_0 is '1'.
    const int i = k - 5;
    const uint32_t xi = i == 0 ? round : vals[i - 1];
    const uint32_t prev = i == 0 ? 0x31u : ((i == 1 ? round : vals[i - 2]) >> 24);
    w = prev | (xi << 8);
  } else if (k == 6 + nvals) {
    const uint32_t last = nvals == 0 ? round : vals[nvals - 1];
    w = (last >> 24) | (0x06u << 8);
  } else {
    w = 0u;
  }
  if (k == 34 * nblocks - 1) w |= 0x80000000u;
  return w;
}

}  // namespace sc
