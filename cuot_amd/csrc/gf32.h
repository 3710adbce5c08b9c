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

// ---- acc = c * v, Horner from the top bit (host and device) ------------------
// The accumulator is multiplied by t (a renaming of its 32 planes plus three
// xors of the wrapped plane) and v is added for every set bit of c, high bit
// first; v itself is never modified, so only v and the accumulator are live.
// One bit per step behind a branch on that bit of c: a scalar branch when c is
// equal in every lane (UNIFORM), a lane-divergent one otherwise.  With
// GF32_TWO_BIT, uniform constants take two bits per step, adding v, v*t or
// v + v*t (v*t differs from a renaming of v in three planes only).
#if defined(__HIP_DEVICE_COMPILE__)
#define GF32_HD __device__ __forceinline__
#elif defined(__HIPCC__)
#define GF32_HD __host__ __device__ inline
#else
#define GF32_HD inline
#endif

// a ^ b ^ c; on the device one three-input bit operation (table 0x96) written
// as asm so that b ^ c is not common-subexpressioned across the unrolled steps
// (which would keep 32 more registers live for the whole product)
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ uint32_t gf32_x3(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
#else
GF32_HD uint32_t gf32_x3(uint32_t a, uint32_t b, uint32_t c) { return a ^ b ^ c; }
#endif

// one multiplication of the accumulator by t; S is the current slot offset
// (logical plane j lives in A[(j + S) & 31]) and becomes S - 1
template <int S>
GF32_HD void gf32_acc_times_t(uint32_t* A) {
  constexpr int S1 = S - 1;
  const uint32_t top = A[(0 + S1) & 31];  // old plane 31
  A[(2 + S1) & 31] ^= top;
  A[(3 + S1) & 31] ^= top;
  A[(7 + S1) & 31] ^= top;
}

template <int D>  // digits D-1 .. 0 remain; offset before this digit: 2*D - 32
GF32_HD void gf32_top_steps_u(uint32_t* A, const uint32_t* v, const uint32_t* vt, uint32_t c) {
  if constexpr (D > 0) {
    constexpr int S0 = 2 * D;       // offset before the two multiplications by t
    gf32_acc_times_t<S0>(A);
    gf32_acc_times_t<S0 - 1>(A);
    constexpr int S = S0 - 2;
    const uint32_t dg = (c >> (2 * (D - 1))) & 3u;
    if (dg == 1u) {
#pragma unroll
      for (int j = 0; j < 32; j++) A[(j + S) & 31] ^= v[j];
    } else if (dg == 2u) {
#pragma unroll
      for (int j = 0; j < 32; j++) A[(j + S) & 31] ^= vt[j];
    } else if (dg == 3u) {
#pragma unroll
      for (int j = 0; j < 32; j++) A[(j + S) & 31] = gf32_x3(A[(j + S) & 31], v[j], vt[j]);
    }
    gf32_top_steps_u<D - 1>(A, v, vt, c);
  }
}

template <int D>  // bits D-1 .. 0 remain; offset before this bit: D - 32
GF32_HD void gf32_top_steps_n(uint32_t* A, const uint32_t* v, uint32_t c) {
  if constexpr (D > 0) {
    gf32_acc_times_t<D>(A);
    constexpr int S = D - 1;
    // a lane-divergent branch: no mask registers, lanes with the bit clear
    // are switched off for the 32 xors
    if ((c >> (D - 1)) & 1u) {
#pragma unroll
      for (int j = 0; j < 32; j++) A[(j + S) & 31] ^= v[j];
    }
    gf32_top_steps_n<D - 1>(A, v, c);
  }
}

// GF32_TWO_BIT = 1 selects two-bit digits for wave-uniform constants (fewer
// instructions, but measured slower in the FFT pass: off by default)
#ifndef GF32_TWO_BIT
#define GF32_TWO_BIT 0
#endif

// acc[j] = (c * v)[j]; acc and v must not alias
template <bool UNIFORM>
GF32_HD void gf32_mulc_top(const uint32_t* v, uint32_t c, uint32_t* acc) {
#pragma unroll
  for (int j = 0; j < 32; j++) acc[j] = 0u;
  if (UNIFORM && GF32_TWO_BIT) {
    uint32_t vt[32];  // v * t: a renaming of v except planes 2, 3 and 7
#pragma unroll
    for (int j = 1; j < 32; j++) vt[j] = v[j - 1];
    vt[0] = v[31];
    vt[2] ^= v[31];
    vt[3] ^= v[31];
    vt[7] ^= v[31];
    gf32_top_steps_u<16>(acc, v, vt, c);
  } else {
    gf32_top_steps_n<32>(acc, v, c);
  }
}
