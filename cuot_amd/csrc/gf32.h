// gf32.h -- GF(2^32) = GF(2)[t]/(t^32 + t^7 + t^3 + t^2 + 1): host arithmetic,
// the Cantor basis, and bitsliced device multiplication by a constant.
//
// Bitsliced layout: a field element for 32 independent rows is 32 words
// x[0..31]; bit r of x[j] is coefficient t^j of row r's element.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#endif

#define CUOT_GF_POLY_LOW 0x8Du  // t^7 + t^3 + t^2 + 1

// ---- host ---------------------------------------------------------------------
static inline uint32_t gf32_mul_host(uint32_t a, uint32_t b) {
  uint64_t r = 0;
  for (int i = 0; i < 32; i++)
    if ((b >> i) & 1) r ^= (uint64_t)a << i;
  for (int k = 62; k >= 32; k--)
    if ((r >> k) & 1) r ^= ((uint64_t)CUOT_GF_POLY_LOW | (1ull << 32)) << (k - 32);
  return (uint32_t)r;
}

// Cantor basis: beta_0 = 1, beta_i^2 + beta_i = beta_{i-1}.  Solves the
// GF(2)-linear system x^2 + x = c by elimination; the root with bit 0 clear is
// taken (the other root is x + 1).  Returns 0 on success.
static inline int gf32_cantor_basis(uint32_t beta[32]) {
  uint32_t col[32];
  for (int j = 0; j < 32; j++) {
    uint32_t tj = 1u << j;
    col[j] = gf32_mul_host(tj, tj) ^ tj;
  }
  beta[0] = 1;
  for (int i = 1; i < 32; i++) {
    // reduce: basis vectors with pivot = top bit, tracking combinations
    uint32_t bv[32] = {0}, bm[32] = {0};
    bool has[32] = {false};
    for (int j = 0; j < 32; j++) {
      uint32_t v = col[j], m = 1u << j;
      while (v) {
        int hb = 31 - __builtin_clz(v);
        if (!has[hb]) {
          has[hb] = true;
          bv[hb] = v;
          bm[hb] = m;
          break;
        }
        v ^= bv[hb];
        m ^= bm[hb];
      }
    }
    uint32_t v = beta[i - 1], m = 0;
    while (v) {
      int hb = 31 - __builtin_clz(v);
      if (!has[hb]) return -1;
      v ^= bv[hb];
      m ^= bm[hb];
    }
    if (m & 1) m ^= 1;  // canonical root
    beta[i] = m;
  }
  return 0;
}

#ifdef __HIPCC__
// ---- device -------------------------------------------------------------------
// y = c * x for 32 bitsliced rows.  UNIFORM: c is the same in every lane of the
// wave (branches on its bits are scalar); otherwise per-lane bit masks.
template <bool UNIFORM>
__device__ __forceinline__ void gf32_mulc(const uint32_t* __restrict__ x, uint32_t c, uint32_t* __restrict__ y) {
  uint32_t p[63];
#pragma unroll
  for (int i = 0; i < 63; i++) p[i] = 0;
#pragma unroll
  for (int i = 0; i < 32; i++) {
    if (UNIFORM) {
      if ((c >> i) & 1u) {
#pragma unroll
        for (int j = 0; j < 32; j++) p[i + j] ^= x[j];
      }
    } else {
      const uint32_t m = 0u - ((c >> i) & 1u);
#pragma unroll
      for (int j = 0; j < 32; j++) p[i + j] ^= x[j] & m;
    }
  }
  // reduce t^(32+i) = t^(i+7) + t^(i+3) + t^(i+2) + t^i, from the top
#pragma unroll
  for (int k = 62; k >= 32; k--) {
    const uint32_t v = p[k];
    p[k - 32 + 7] ^= v;
    p[k - 32 + 3] ^= v;
    p[k - 32 + 2] ^= v;
    p[k - 32] ^= v;
  }
#pragma unroll
  for (int i = 0; i < 32; i++) y[i] = p[i];
}
#endif
