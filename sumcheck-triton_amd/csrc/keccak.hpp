// SHA3-256 (FIPS 202) for host and device, written from the specification.
//
// The transcript hashes at most a few hundred bytes per round, so the permutation is the
// straightforward 24-round Keccak-f[1600] over 25 64-bit lanes, fully unrolled. On the device
// it is called with wave-uniform inputs so the compiler can keep the state in scalar
// registers and run it on the scalar ALU.
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
// Keccak-f[1600] spread over one wave: lane x + 5y (< 25) holds A[x + 5y] as two 32-bit
// halves, and the column sums of theta, the pi permutation and the row terms of chi are lane
// exchanges. Lanes 25..63 follow along on copies of lane 0..24 data and are ignored.
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, 64);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, 64);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

__device__ __forceinline__ uint64_t rotl64_var(uint64_t x, int s) {
  return s == 0 ? x : ((x << s) | (x >> (64 - s)));
}

__device__ __forceinline__ void keccak_f1600_wave(uint64_t& a) {
  const int lane = (int)(threadIdx.x & 63) % 25;
  const int x = lane % 5, y = lane / 5;
  // rho offset of this lane's word, and the lane whose word pi moves here: B[y, 2x+3y] =
  // rot(A[x, y]) means lane (X, Y) receives from (x, y) = ((Y - 3X) * 3 mod 5, X) after rho.
  constexpr int R[25] = {0,  1,  62, 28, 27, 36, 44, 6,  55, 20, 3,  10, 43,
                         25, 39, 41, 45, 15, 21, 8,  18, 2,  61, 56, 14};
  const int rot = R[lane];
  const int pi_src = ((3 * ((y - 3 * x + 15) % 5)) % 5) + 5 * x;
  const int c1 = (x + 4) % 5, c2 = (x + 1) % 5;
  const int n1 = (x + 1) % 5 + 5 * y, n2 = (x + 2) % 5 + 5 * y;
  for (int round = 0; round < 24; ++round) {
    // theta: column parity C[x], then D[x] = C[x-1] ^ rot(C[x+1], 1)
    uint64_t c = a;
#pragma unroll
    for (int k = 1; k < 5; ++k) c ^= shfl64(a, x + 5 * ((y + k) % 5));
    a ^= shfl64(c, c1) ^ rotl64(shfl64(c, c2), 1);
    // rho and pi
    const uint64_t b = shfl64(rotl64_var(a, rot), pi_src);
    // chi and iota
    a = b ^ (~shfl64(b, n1) & shfl64(b, n2));
    if (lane == 0 && (threadIdx.x & 63) == 0) a ^= keccak_rc(round);
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
