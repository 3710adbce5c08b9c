// gf32.h -- GF(2^32) as the tower GF(2^16)[theta] / (theta^2 + theta + nu)
// over GF(2^16) = GF(2)[z] / (z^16 + z^5 + z^3 + z + 1), with nu = z^-1
// (absolute trace 1, so the quadratic is irreducible): host and device scalar
// arithmetic, the Cantor basis, the embedding of 16-bit chunk polynomials,
// and bitsliced device multiplication by a constant.
//
// A 32-bit value holds x0 + x1 * theta: x0 in the low 16 bits, x1 in the high
// 16 bits, both in the z-polynomial basis.  The subfield GF(2^16) is the set
// {x1 = 0}, so a product by a subfield constant acts on the two halves
// independently (two 16-bit products instead of one 32-bit one); the twiddles
// of every FFT layer k >= m - 16 lie in that subfield.
//
// Chunk embedding: a chunk polynomial u(x) of degree < 16 maps to u(theta)
// (theta has degree 32 over GF(2)), so a chunk product of degree <= 30 maps to
// its value at theta without wrap-around and the 31 coefficients are read back
// through the inverse of the basis {theta^i}.
//
// Bitsliced layout: a field element for 32 independent rows is 32 words
// x[0..31]; bit r of x[j] is coordinate j of row r's element.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#endif

#define CUOT_GF16_POLY_LOW 0x002Bu  // z^5 + z^3 + z + 1
#define CUOT_GF16_NU 0x8015u        // z^-1 = z^15 + z^4 + z^2 + 1

// ---- scalar arithmetic (constexpr: host, device and compile time) ----------
constexpr uint32_t gf16_mul(uint32_t a, uint32_t b) {
  uint32_t r = 0;
  for (int i = 0; i < 16; i++)
    if ((b >> i) & 1u) r ^= a << i;
  for (int k = 30; k >= 16; k--)
    if ((r >> k) & 1u) r ^= (0x10000u | CUOT_GF16_POLY_LOW) << (k - 16);
  return r;
}

// x * nu = x / z in GF(2^16): x = x_0 + z x' gives x / z = x_0 z^-1 + x'
constexpr uint32_t gf16_times_nu(uint32_t x) { return (x >> 1) ^ ((x & 1u) ? CUOT_GF16_NU : 0u); }

// (a0 + a1 t)(b0 + b1 t) = a0 b0 + nu a1 b1 + (a0 b1 + a1 b0 + a1 b1) t,  t^2 = t + nu
constexpr uint32_t gf32_mul(uint32_t a, uint32_t b) {
  const uint32_t a0 = a & 0xFFFFu, a1 = a >> 16, b0 = b & 0xFFFFu, b1 = b >> 16;
  const uint32_t p0 = gf16_mul(a0, b0), p1 = gf16_mul(a1, b1), p2 = gf16_mul(a0 ^ a1, b0 ^ b1);
  return (p0 ^ gf16_times_nu(p1)) | ((p2 ^ p0) << 16);
}

static inline uint32_t gf32_mul_host(uint32_t a, uint32_t b) { return gf32_mul(a, b); }

// ---- chunk embedding: pack (u -> u(theta)) and unpack (degree <= 30) ---------
struct GfChunkMaps {
  uint32_t pack[16];    // pack[i] = theta^i (value of coordinate set by chunk bit i)
  uint32_t unpack[31];  // bit j of unpack[i]: coordinate j contributes to coefficient i
};

constexpr GfChunkMaps gf_chunk_maps() {
  GfChunkMaps m{};
  uint32_t pw[32] = {};
  pw[0] = 1u;
  for (int i = 1; i < 32; i++) pw[i] = gf32_mul(pw[i - 1], 0x10000u);  // theta = (0, 1)
  for (int i = 0; i < 16; i++) m.pack[i] = pw[i];
  // invert the 32 x 32 matrix whose column i is pw[i]: rows[r] bit c = bit r of pw[c]
  uint32_t a[32] = {}, inv[32] = {};
  for (int r = 0; r < 32; r++) {
    for (int c = 0; c < 32; c++)
      if ((pw[c] >> r) & 1u) a[r] |= 1u << c;
    inv[r] = 1u << r;
  }
  // row reduction of [a | inv] to [I | a^-1] (columns of a = unknown coefficients)
  for (int c = 0; c < 32; c++) {
    int piv = c;
    while (!((a[piv] >> c) & 1u)) piv++;
    const uint32_t ta = a[piv], ti = inv[piv];
    a[piv] = a[c];
    inv[piv] = inv[c];
    a[c] = ta;
    inv[c] = ti;
    for (int r = 0; r < 32; r++)
      if (r != c && ((a[r] >> c) & 1u)) {
        a[r] ^= a[c];
        inv[r] ^= inv[c];
      }
  }
  for (int i = 0; i < 31; i++) m.unpack[i] = inv[i];
  return m;
}

inline constexpr GfChunkMaps kGfChunk = gf_chunk_maps();

// value of the 16-bit chunk polynomial u at theta
constexpr uint32_t gf32_pack_chunk(uint32_t u) {
  uint32_t v = 0;
  for (int i = 0; i < 16; i++)
    if ((u >> i) & 1u) v ^= kGfChunk.pack[i];
  return v;
}

// ---- Cantor basis: beta_0 = 1, beta_i^2 + beta_i = beta_{i-1} ----------------
// Solves the GF(2)-linear system x^2 + x = c by elimination; the root with
// coordinate 0 clear is taken (the other root is x + 1).  Returns 0 on success.
static inline int gf32_cantor_basis(uint32_t beta[32]) {
  uint32_t col[32];
  for (int j = 0; j < 32; j++) {
    uint32_t tj = 1u << j;
    col[j] = gf32_mul(tj, tj) ^ tj;
  }
  beta[0] = 1;
  for (int i = 1; i < 32; i++) {
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

// ---- bitsliced multiplication by a constant (host and device) -----------------
#if defined(__HIP_DEVICE_COMPILE__)
#define GF32_HD __device__ __forceinline__
#elif defined(__HIPCC__)
#define GF32_HD __host__ __device__ inline
#else
#define GF32_HD inline
#endif

// One multiplication of a 16-plane accumulator half by z: logical plane j
// lives in A[(j + S) & 15] and S becomes S - 1 (a renaming), then the old top
// plane (now logical plane 0) is added to planes 1, 3 and 5.
template <int S>
GF32_HD void gf16_acc_times_z(uint32_t* A) {
  constexpr int S1 = S - 1;
  const uint32_t top = A[(0 + S1) & 15];
  A[(1 + S1) & 15] ^= top;
  A[(3 + S1) & 15] ^= top;
  A[(5 + S1) & 15] ^= top;
}

// A0 = c * v[0..15], A1 = c * v[16..31] for c in GF(2^16): Horner from the top
// bit of c, both halves in one pass (one branch per bit: scalar when c is
// equal in every lane, lane-divergent otherwise; lanes with the bit clear are
// switched off for the 32 xors).  D bits remain; offset before this bit: D.
template <int D>
GF32_HD void gf16x2_steps(uint32_t* A0, uint32_t* A1, const uint32_t* v, uint32_t c) {
  if constexpr (D > 0) {
    gf16_acc_times_z<D>(A0);
    gf16_acc_times_z<D>(A1);
    constexpr int S = D - 1;
    if ((c >> (D - 1)) & 1u) {
#pragma unroll
      for (int j = 0; j < 16; j++) {
        A0[(j + S) & 15] ^= v[j];
        A1[(j + S) & 15] ^= v[16 + j];
      }
    }
    gf16x2_steps<D - 1>(A0, A1, v, c);
  }
}

// A = a * x + b * y on 16-plane halves (x, y: 16 planes each)
template <int D>
GF32_HD void gf16_steps2(uint32_t* A, const uint32_t* x, uint32_t a, const uint32_t* y, uint32_t b) {
  if constexpr (D > 0) {
    gf16_acc_times_z<D>(A);
    constexpr int S = D - 1;
    if ((a >> (D - 1)) & 1u) {
#pragma unroll
      for (int j = 0; j < 16; j++) A[(j + S) & 15] ^= x[j];
    }
    if ((b >> (D - 1)) & 1u) {
#pragma unroll
      for (int j = 0; j < 16; j++) A[(j + S) & 15] ^= y[j];
    }
    gf16_steps2<D - 1>(A, x, a, y, b);
  }
}

// A = c * v on one 16-plane half (c in GF(2^16))
template <int D>
GF32_HD void gf16_steps1(uint32_t* A, const uint32_t* v, uint32_t c) {
  if constexpr (D > 0) {
    gf16_acc_times_z<D>(A);
    constexpr int S = D - 1;
    if ((c >> (D - 1)) & 1u) {
#pragma unroll
      for (int j = 0; j < 16; j++) A[(j + S) & 15] ^= v[j];
    }
    gf16_steps1<D - 1>(A, v, c);
  }
}

// Branch-free form for a wave-uniform c: every bit's mask comes from a scalar
// register and each plane takes one three-input (A ^ (v & mask)) operation.
template <int D>
GF32_HD void gf16_steps1_masked(uint32_t* A, const uint32_t* v, uint32_t c) {
  if constexpr (D > 0) {
    gf16_acc_times_z<D>(A);
    constexpr int S = D - 1;
    const uint32_t msk = 0u - ((c >> (D - 1)) & 1u);
#pragma unroll
    for (int j = 0; j < 16; j++) A[(j + S) & 15] ^= v[j] & msk;
    gf16_steps1_masked<D - 1>(A, v, c);
  }
}

GF32_HD void gf16_mulc_masked(const uint32_t* v, uint32_t c, uint32_t* acc) {
#pragma unroll
  for (int j = 0; j < 16; j++) acc[j] = 0u;
  gf16_steps1_masked<16>(acc, v, c & 0xFFFFu);
}

// acc[0..15] = c * v[0..15] in GF(2^16); acc and v must not alias
GF32_HD void gf16_mulc(const uint32_t* v, uint32_t c, uint32_t* acc) {
#pragma unroll
  for (int j = 0; j < 16; j++) acc[j] = 0u;
  gf16_steps1<16>(acc, v, c & 0xFFFFu);
}

// acc = c * v for a constant c of the subfield GF(2^16) (high half zero);
// acc and v must not alias.  UNIFORM only documents the caller's expectation:
// the branches are scalar exactly when c is wave-uniform.
template <bool UNIFORM>
GF32_HD void gf32_mulc_sub(const uint32_t* v, uint32_t c, uint32_t* acc) {
#pragma unroll
  for (int j = 0; j < 32; j++) acc[j] = 0u;
  gf16x2_steps<16>(acc, acc + 16, v, c & 0xFFFFu);
}

// acc = c * v for a general constant c = c0 + c1 theta (Karatsuba on the
// halves): P0 = c0 v0, P1 = c1 v1, P2 = (c0 + c1)(v0 + v1);
//   acc0 = P0 + nu P1,   acc1 = P2 + P0
// (three 16-bit products instead of four; one 16-plane temporary).  The
// product by nu = z^-1 is a renaming plus three xors: plane j takes plane
// j + 1, plane 15 takes plane 0, and plane 0 is added to planes 0, 2 and 4.
template <bool UNIFORM>
GF32_HD void gf32_mulc(const uint32_t* v, uint32_t c, uint32_t* acc) {
  const uint32_t c0 = c & 0xFFFFu, c1 = c >> 16;
  uint32_t t[16];
#pragma unroll
  for (int j = 0; j < 16; j++) {
    t[j] = v[j] ^ v[16 + j];
    acc[16 + j] = 0u;
    acc[j] = 0u;
  }
  gf16_steps1<16>(acc + 16, t, c0 ^ c1);  // P2
  gf16_steps1<16>(acc, v, c0);            // P0
#pragma unroll
  for (int j = 0; j < 16; j++) {
    acc[16 + j] ^= acc[j];
    t[j] = 0u;
  }
  gf16_steps1<16>(t, v + 16, c1);  // P1
  const uint32_t x0 = t[0];
#pragma unroll
  for (int j = 0; j < 15; j++) acc[j] ^= t[j + 1];
  acc[15] ^= x0;
  acc[0] ^= x0;
  acc[2] ^= x0;
  acc[4] ^= x0;
}

// out[0..31] = pack of the 16 chunk-bit planes in[0..15] (out and in distinct)
GF32_HD void gf32_pack_planes(const uint32_t* in, uint32_t* out) {
#pragma unroll
  for (int j = 0; j < 32; j++) {
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
      if ((kGfChunk.pack[i] >> j) & 1u) s ^= in[i];
    out[j] = s;
  }
}

// out[0..15] = coordinates J0 .. J0 + 15 of the pack of in[0..15]
template <int J0>
GF32_HD void gf32_pack_planes_half(const uint32_t* in, uint32_t* out) {
#pragma unroll
  for (int j = 0; j < 16; j++) {
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
      if ((kGfChunk.pack[i] >> (J0 + j)) & 1u) s ^= in[i];
    out[j] = s;
  }
}

// out[0..30] = the 31 coefficient planes of the chunk products held in in[0..31]
GF32_HD void gf32_unpack_planes(const uint32_t* in, uint32_t* out) {
#pragma unroll
  for (int i = 0; i < 31; i++) {
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < 32; j++)
      if ((kGfChunk.unpack[i] >> j) & 1u) s ^= in[j];
    out[i] = s;
  }
}
