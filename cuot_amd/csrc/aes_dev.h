// aes_dev.h -- AES-128 encryption on gfx950 with an LDS-resident T-table.
//
// CDNA4 has no AES instructions.  Each round is done with one 256-entry
// 32-bit table T0 (SubBytes+MixColumns of a row-0 byte); the row-1..3 tables
// are byte rotations of T0, applied with a rotate after the lookup.  The table
// is replicated CUOT_AES_REP times in LDS so that lanes with different
// (lane % REP) hit different banks: entry x of copy r is at word x*REP + r.
//
// The tables are computed on the host from the FIPS-197 definition of the
// S-box (multiplicative inverse + affine map) and uploaded at init.
//
// State words are little-endian columns: word c = bytes 4c..4c+3 of the block,
// i.e. the block's memory layout.  Output = FIPS-197 AES-128 of the 16 bytes.
#pragma once

#include "common.h"

#ifndef CUOT_AES_REP
#define CUOT_AES_REP 32
#endif

#define CUOT_AES_LDS_WORDS (256 * CUOT_AES_REP)


struct AesLds {
  const uint32_t* base;  // LDS table, replicated
  uint32_t lane_off;     // (lane % REP) * 4 bytes
};

// Fill the replicated table.  Call from all threads of the block, then sync.
// T0g: the 256-word table in global memory (uploaded by cuot_init).
CUOT_DEV void aes_stage_table(uint32_t* lds, const uint32_t* __restrict__ T0g, int tid, int nthreads) {
  for (int i = tid; i < CUOT_AES_LDS_WORDS; i += nthreads) lds[i] = T0g[i / CUOT_AES_REP];
}

CUOT_DEV AesLds aes_lds(const uint32_t* lds) {
  AesLds a;
  a.base = lds;
  a.lane_off = (threadIdx.x % CUOT_AES_REP) * 4u;
  return a;
}

// Table lookup of byte value b (0..255).
CUOT_DEV uint32_t tlook(const AesLds& a, uint32_t b) {
  uint32_t off = (b * (CUOT_AES_REP * 4u)) + a.lane_off;
  return *(const uint32_t*)((const char*)a.base + off);
}

CUOT_DEV uint32_t rotl32(uint32_t x, int s) { return (x << s) | (x >> (32 - s)); }

// a ^ b ^ c as one three-input bit operation (table 0x96) on VGPR operands;
// the round key (uniform in some kernels, per lane in others) is added by a
// plain XOR, which takes it from either register file
CUOT_DEV uint32_t aes_x3(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

CUOT_DEV uint32_t aes_col(const AesLds& a, uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint32_t k) {
  uint32_t t0 = tlook(a, w0 & 0xffu);
  uint32_t t1 = tlook(a, (w1 >> 8) & 0xffu);
  uint32_t t2 = tlook(a, (w2 >> 16) & 0xffu);
  uint32_t t3 = tlook(a, w3 >> 24);
  return aes_x3(t0, rotl32(t1, 8), rotl32(t2, 16)) ^ rotl32(t3, 24) ^ k;
}

// Last round: S-box bytes recovered from T0 (T0[x] bytes are 2s, s, s, 3s).
CUOT_DEV uint32_t aes_col_last(const AesLds& a, uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3,
                               uint32_t k) {
  uint32_t t0 = tlook(a, w0 & 0xffu);
  uint32_t t1 = tlook(a, (w1 >> 8) & 0xffu);
  uint32_t t2 = tlook(a, (w2 >> 16) & 0xffu);
  uint32_t t3 = tlook(a, w3 >> 24);
  return (((t0 >> 8) & 0x000000ffu) | (t1 & 0x0000ff00u) | (t2 & 0x00ff0000u) | ((t3 << 8) & 0xff000000u)) ^ k;
}

// rk: 44 little-endian round-key words (FIPS w[0..43] in memory order).
CUOT_DEV blk aes_enc(const AesLds& a, const uint32_t* __restrict__ rk, blk in) {
  uint32_t s0 = in.x ^ rk[0], s1 = in.y ^ rk[1], s2 = in.z ^ rk[2], s3 = in.w ^ rk[3];
#pragma unroll
  for (int r = 1; r < 10; r++) {
    uint32_t n0 = aes_col(a, s0, s1, s2, s3, rk[4 * r + 0]);
    uint32_t n1 = aes_col(a, s1, s2, s3, s0, rk[4 * r + 1]);
    uint32_t n2 = aes_col(a, s2, s3, s0, s1, rk[4 * r + 2]);
    uint32_t n3 = aes_col(a, s3, s0, s1, s2, rk[4 * r + 3]);
    s0 = n0;
    s1 = n1;
    s2 = n2;
    s3 = n3;
  }
  blk o;
  o.x = aes_col_last(a, s0, s1, s2, s3, rk[40]);
  o.y = aes_col_last(a, s1, s2, s3, s0, rk[41]);
  o.z = aes_col_last(a, s2, s3, s0, s1, rk[42]);
  o.w = aes_col_last(a, s3, s0, s1, s2, rk[43]);
  return o;
}

// GGM two-key PRG: left = AES_k0(p) ^ p, right = AES_k1(p) ^ p.
// rk2: 88 words, key 0 schedule then key 1 schedule.
CUOT_DEV void ggm_expand(const AesLds& a, const uint32_t* __restrict__ rk2, blk p, blk& l, blk& r) {
  l = bxor(aes_enc(a, rk2, p), p);
  r = bxor(aes_enc(a, rk2 + 44, p), p);
}
