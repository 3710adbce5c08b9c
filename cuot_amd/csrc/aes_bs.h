// aes_bs.h -- AES-128 encryption bitsliced over 32 blocks (no tables).
//
// The variant of AES for gfx950 without lookups: one lane keeps 32 blocks as
// 128 bit-slice words, S[8 j + i] = bit i of byte j of the 32 blocks (bit k of
// the word belongs to block k).  SubBytes is the generated tower-field circuit
// (aes_bs_sbox.h, 106 logic operations per byte position for all 32 blocks);
// ShiftRows is a renaming of words; MixColumns is XORs of words (xtime is a
// renaming plus three XORs); AddRoundKey complements the words whose key bit
// is set.  The S-box constant 0x63 is carried by the round keys: MixColumns
// maps a column of four equal bytes to itself, so SubBytes without the
// constant followed by round key rk_r ^ 0x63.. (r >= 1) is the cipher.
//
// Loading: word w of block k lands in S[32 w + k]; transposing each group of
// 32 words (bit q of word w of block k -> bit k of word 32 w + q) gives the
// slice layout, since byte j = 4 w + q / 8 and bit i = q % 8 make
// 8 j + i = 32 w + q.  The same transposition maps back.
//
// Compiled for the device by aes_bs.hip and for the host by
// tests/aes_bs_host_check.cpp.
#pragma once

#include <stdint.h>
#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#endif

#if defined(__HIP_DEVICE_COMPILE__)
#define BS_HD __device__ __forceinline__
// a ^ b ^ c in one three-input bit operation (table 0x96 is symmetric)
BS_HD uint32_t bs_x3(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
#elif defined(__HIPCC__)
#define BS_HD __host__ __device__ inline
BS_HD uint32_t bs_x3(uint32_t a, uint32_t b, uint32_t c) { return a ^ b ^ c; }
#else
#define BS_HD inline
inline uint32_t bs_x3(uint32_t a, uint32_t b, uint32_t c) { return a ^ b ^ c; }
#endif
#define BS_X3(a, b, c) bs_x3((a), (b), (c))

#include "aes_bs_sbox.h"

// In-place transpose of a 32 x 32 bit matrix: bit q of a[k] <-> bit k of a[q].
BS_HD void bs_transpose32(uint32_t* a) {
  const uint32_t masks[5] = {0x0000FFFFu, 0x00FF00FFu, 0x0F0F0F0Fu, 0x33333333u, 0x55555555u};
#pragma unroll
  for (int st = 0; st < 5; st++) {
    const int j = 16 >> st;
    const uint32_t m = masks[st];
#pragma unroll
    for (int k = 0; k < 32; k++) {
      if (k & j) continue;
      const uint32_t t = ((a[k] >> j) ^ a[k + j]) & m;
      a[k] ^= t << j;
      a[k + j] ^= t;
    }
  }
}

// S ^= round key words rk[0..3]: words of set key bits are complemented.
BS_HD void bs_add_key(uint32_t* S, const uint32_t* rk) {
#pragma unroll
  for (int w = 0; w < 4; w++) {
    const uint32_t kw = rk[w];
#pragma unroll
    for (int q = 0; q < 32; q++) S[32 * w + q] ^= 0u - ((kw >> q) & 1u);
  }
}

// SubBytes (without 0x63) and ShiftRows: byte j = (row j % 4, column j / 4)
// moves to column (j / 4 - j % 4) mod 4 of its row.
BS_HD void bs_sub_shift(const uint32_t* S, uint32_t* T) {
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int r = j & 3, c = j >> 2;
    const int d = 4 * ((c - r) & 3) + r;
    const uint32_t* x = S + 8 * j;
    uint32_t* y = T + 8 * d;
    AES_BS_SBOX(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], y[0], y[1], y[2], y[3], y[4], y[5], y[6], y[7]);
  }
}

// MixColumns from T into S: with t = a0 ^ a1 ^ a2 ^ a3, row r of a column is
// a_r ^ t ^ xtime(a_r ^ a_{r+1}); xtime(u) bit i = u_{i-1}, plus u_7 at bits
// 0, 1, 3, 4 (the reduction by 0x11B).
BS_HD void bs_mix_columns(const uint32_t* T, uint32_t* S) {
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const uint32_t* a = T + 32 * c;  // a[8 r + i]: row r, bit i
    uint32_t t[8], u[4][8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      t[i] = bs_x3(a[i], a[8 + i], a[16 + i]) ^ a[24 + i];
#pragma unroll
      for (int r = 0; r < 4; r++) u[r][i] = a[8 * r + i] ^ a[8 * ((r + 1) & 3) + i];
    }
    uint32_t* o = S + 32 * c;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const uint32_t* ur = u[r];
      o[8 * r + 0] = bs_x3(a[8 * r + 0], t[0], ur[7]);
      o[8 * r + 1] = bs_x3(a[8 * r + 1], t[1], ur[0]) ^ ur[7];
      o[8 * r + 2] = bs_x3(a[8 * r + 2], t[2], ur[1]);
      o[8 * r + 3] = bs_x3(a[8 * r + 3], t[3], ur[2]) ^ ur[7];
      o[8 * r + 4] = bs_x3(a[8 * r + 4], t[4], ur[3]) ^ ur[7];
      o[8 * r + 5] = bs_x3(a[8 * r + 5], t[5], ur[4]);
      o[8 * r + 6] = bs_x3(a[8 * r + 6], t[6], ur[5]);
      o[8 * r + 7] = bs_x3(a[8 * r + 7], t[7], ur[6]);
    }
  }
}

// 0x63 in every byte of round keys 1..10, as bs_encrypt expects them
inline void bs_fold_sbox_constant(const uint32_t rk[44], uint32_t out[44]) {
  for (int w = 0; w < 44; w++) out[w] = rk[w] ^ (w >= 4 ? 0x63636363u : 0u);
}

// AES-128 of the 32 sliced blocks in S; rk: 44 words from bs_fold_sbox_constant.
BS_HD void bs_encrypt(uint32_t* S, const uint32_t* rk) {
  uint32_t T[128];
  bs_add_key(S, rk);
#pragma nounroll
  for (int r = 1; r < 10; r++) {
    bs_sub_shift(S, T);
    bs_mix_columns(T, S);
    bs_add_key(S, rk + 4 * r);
  }
  bs_sub_shift(S, T);
#pragma unroll
  for (int q = 0; q < 128; q++) S[q] = T[q];
  bs_add_key(S, rk + 40);
}
