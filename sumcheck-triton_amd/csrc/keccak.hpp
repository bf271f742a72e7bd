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

// Bit interleaving (the Keccak team's "Keccak implementation overview", section 2.1): a 64-bit
// lane word w is held as two 32-bit halves, E = its even bits and O = its odd bits
// (E bit i = w bit 2i, O bit i = w bit 2i + 1). XOR, AND and NOT act on each half alone, and
// a 64-bit rotation becomes two 32-bit rotations:
//   rotl64 by 2m:     E' = rotl32(E, m),      O' = rotl32(O, m)
//   rotl64 by 2m + 1: E' = rotl32(O, m + 1),  O' = rotl32(E, m)
// unshuffle32 moves the even bits of x to bits 0..15 and the odd bits to 16..31; shuffle32
// undoes it (each step swaps two bit groups, so the inverse runs the steps in reverse).
SC_HD uint32_t unshuffle32(uint32_t x) {
  uint32_t t;
  t = (x ^ (x >> 1)) & 0x22222222u; x ^= t ^ (t << 1);
  t = (x ^ (x >> 2)) & 0x0C0C0C0Cu; x ^= t ^ (t << 2);
  t = (x ^ (x >> 4)) & 0x00F000F0u; x ^= t ^ (t << 4);
  t = (x ^ (x >> 8)) & 0x0000FF00u; x ^= t ^ (t << 8);
  return x;
}

SC_HD uint32_t shuffle32(uint32_t x) {
  uint32_t t;
  t = (x ^ (x >> 8)) & 0x0000FF00u; x ^= t ^ (t << 8);
  t = (x ^ (x >> 4)) & 0x00F000F0u; x ^= t ^ (t << 4);
  t = (x ^ (x >> 2)) & 0x0C0C0C0Cu; x ^= t ^ (t << 2);
  t = (x ^ (x >> 1)) & 0x22222222u; x ^= t ^ (t << 1);
  return x;
}

// Half h (0: even bits, 1: odd bits) of the 64-bit word w.
SC_HD uint32_t interleaved_half(uint64_t w, int h) {
  const uint32_t lo = unshuffle32((uint32_t)w), hi = unshuffle32((uint32_t)(w >> 32));
  return h == 0 ? ((lo & 0xFFFFu) | (hi << 16)) : ((lo >> 16) | (hi & 0xFFFF0000u));
}

// The 64-bit word whose even bits are e and odd bits are o.
SC_HD uint64_t interleaved_join(uint32_t e, uint32_t o) {
  const uint32_t lo = shuffle32((e & 0xFFFFu) | (o << 16));
  const uint32_t hi = shuffle32((e >> 16) | (o & 0xFFFF0000u));
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

#if defined(__HIPCC__)
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The lane whose word pi moves to lane (X, Y): B[y, 2x+3y] = rot(A[x, y]), so lane (X, Y)
// receives from (x, y) = ((Y - 3X) * 3 mod 5, X).
__device__ __forceinline__ int keccak_pi_src(int lane) {
  const int x = lane % 5, y = lane / 5;
  return ((3 * ((y - 3 * x + 15) % 5)) % 5) + 5 * x;
}

// Lane position helpers of the interleaved layout.
__device__ __forceinline__ int keccak_half() { return (int)(threadIdx.x & 63) >> 5; }
__device__ __forceinline__ int keccak_word() {
  const int p = (int)(threadIdx.x & 31);
  return p < 25 ? p : p - 25;
}

// rotl32(v, k) for sh = (32 - k) & 31 (funnel shift of v:v right by sh).
__device__ __forceinline__ uint32_t rotl32_sh(uint32_t v, uint32_t sh) { return __builtin_amdgcn_alignbit(v, v, sh); }

// Keccak-f[1600] spread over one wave in interleaved form: lane h * 32 + p holds half h of
// state word p = x + 5y (p < 25); lanes h * 32 + 25..31 repeat words 0..6 (same values, so
// every lane can store without a branch). A round is two exchanges through 50 words of LDS
// (xs[h * 25 + p]): theta reads the column parities of columns x - 1 (own half) and x + 1
// (the other half, which rotl64 by 1 needs) and forms this lane's half of
// A ^ C[x-1] ^ rotl64(C[x+1], 1); rho, pi and chi then read, for this lane and its two row
// neighbours, the half of the pi source that the source's rho offset maps to this half,
// rotated by that offset's half (one v_alignbit). The wave runs in lockstep and LDS serves
// one wave's accesses in order, so wave-scope fences (which also keep the compiler from
// reordering the accesses) are the only synchronization.
// Against one 64-bit word per lane (two ops per XOR/AND, 64-bit variable rotations) this
// halves the per-round VALU chain: 3.0 vs 4.6 us per permutation on MI355X.
__device__ __forceinline__ void keccak_f1600_wave(uint32_t& a, uint32_t* xs) {
  const int h = keccak_half(), p = keccak_word();
  const int x = p % 5, y = p / 5;
  constexpr int R[25] = {0,  1,  62, 28, 27, 36, 44, 6,  55, 20, 3,  10, 43,
                         25, 39, 41, 45, 15, 21, 8,  18, 2,  61, 56, 14};
  const uint32_t* col_m = xs + h * 25 + (x + 4) % 5;
  const uint32_t* col_p = xs + (1 - h) * 25 + (x + 1) % 5;
  const uint32_t d_sh = h == 0 ? 31u : 0u;  // rotl64 by 1: the even half takes rotl32(O, 1)
  const int src[3] = {keccak_pi_src(p), keccak_pi_src((x + 1) % 5 + 5 * y), keccak_pi_src((x + 2) % 5 + 5 * y)};
  const uint32_t* bsrc[3];
  uint32_t bsh[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int r = R[src[j]];
    const int hs = (r & 1) ? 1 - h : h;
    const int k = (r & 1) ? (h == 0 ? (r + 1) / 2 : (r - 1) / 2) : r / 2;
    bsrc[j] = xs + hs * 25 + src[j];
    bsh[j] = (uint32_t)(32 - k) & 31u;
  }
  uint32_t* mine = xs + h * 25 + p;
  // round constant i, both halves, sits in lane i; a round reads its pair with readlane
  const int l = (int)(threadIdx.x & 63);
  const uint64_t rc = keccak_rc(l < 24 ? l : 0);
  const uint32_t rc_e = interleaved_half(rc, 0), rc_o = interleaved_half(rc, 1);
  for (int round = 0; round < 24; ++round) {
    *mine = a;
    wave_sync_lds();
    uint32_t c_m = 0, c_p = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      c_m ^= col_m[5 * k];
      c_p ^= col_p[5 * k];
    }
    const uint32_t t = a ^ c_m ^ rotl32_sh(c_p, d_sh);
    wave_sync_lds();
    *mine = t;
    wave_sync_lds();
    const uint32_t b0 = rotl32_sh(*bsrc[0], bsh[0]);
    const uint32_t b1 = rotl32_sh(*bsrc[1], bsh[1]);
    const uint32_t b2 = rotl32_sh(*bsrc[2], bsh[2]);
    wave_sync_lds();
    a = b0 ^ (~b1 & b2);
    const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)rc_e, round);
    const uint32_t o = (uint32_t)__builtin_amdgcn_readlane((int)rc_o, round);
    if (p == 0) a ^= h == 0 ? e : o;  // iota
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
