// qc.hip -- exact quasi-cyclic encoder over GF(2) for all 128 bit-rows at once:
//     y_i = x_i ^ XOR_{j : a_j = 1} x_{n + ((i - j) mod n)},   i < n
// (the workload's "2n noisy blocks -> n through a random circulant"; the
// reference computes a float32 cuFFT convolution instead,
// silent/quasi_cyclic.cu:11-49,118-128, whose parity is inexact at this size).
//
// Algorithm: binary polynomial product a(x) * e1(x) by an additive FFT over
// GF(2^32) (Lin-Chung-Han novel-basis FFT on a Cantor basis), then the fold
// mod x^n - 1.  Bits are packed 16 per field element (chunk polynomials of
// degree < 16 multiply without reduction, so the chunk products are exact),
// and the 128 rows are bitsliced: block x_{n+16c+b} IS bit-plane b of chunk c
// for all 128 rows, so no bit transpose is needed.
//
// Data layout (device scratch): X[plane][elem] of 16-byte words (128 rows),
// 32 planes x M elements, M = 2^m >= 2*ceil(n/16) - 1.
//
// Steps (forward, per party):  pack -> basis conversion (division by the
// subspace polynomials s_k, XOR only) -> duplicate -> FFT layers m-2..0 ->
// pointwise * A^ -> inverse FFT layers 0..m-1 -> inverse conversion ->
// overlap-add of 31-bit chunk products, fold mod x^n - 1, xor x_0.
#include <string.h>

#include "gf32.h"
#include "kernels.h"

namespace {

// ---- helpers ----------------------------------------------------------------
struct QcGeo {
  int m;        // log2 M
  int64_t M;    // transform length
  int64_t n;    // code length
};

// lower terms of s_k: e = 2^j for j a proper submask of k
CUOT_DEV int lower_terms(int k, uint32_t* e) {
  int c = 0;
  for (int j = 0; j < k; j++)
    if ((j & ~k) == 0) e[c++] = 1u << j;
  return c;
}

// C(g) = sum_b g_b * beta_{b+1}  (the twist s_k(shift of chunk g))
CUOT_DEV uint32_t chunk_const(const uint32_t* beta, uint64_t g) {
  uint32_t c = 0;
  for (int b = 0; g; b++, g >>= 1)
    if (g & 1) c ^= beta[b + 1];
  return c;
}

struct Beta {
  uint32_t v[32];
};

// ---- pack -------------------------------------------------------------------
// X[b][c] = x1[16c + b] (zero beyond n), c < M/2, planes 0..15.
__global__ void qc_pack_x(const uint4* __restrict__ x1, int64_t n, uint4* __restrict__ X, int64_t M) {
  const int64_t half = M >> 1;
  const int64_t total = half * 16;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t >> 4;
    const int b = (int)(t & 15);
    const int64_t src = 16 * c + b;
    X[(int64_t)b * M + c] = src < n ? x1[src] : make_uint4(0, 0, 0, 0);
  }
}

// single-row variant for a: row 0 (bit 0 of word 0) = bit (16c + b) of a
__global__ void qc_pack_a(const uint8_t* __restrict__ a, int64_t n, uint4* __restrict__ X, int64_t M) {
  const int64_t half = M >> 1;
  const int64_t total = half * 16;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t >> 4;
    const int b = (int)(t & 15);
    const int64_t src = 16 * c + b;
    uint32_t bit = src < n ? (a[src >> 3] >> (src & 7)) & 1u : 0u;
    X[(int64_t)b * M + c] = make_uint4(bit, 0, 0, 0);
  }
}

// ---- basis conversion steps (plane-wise XOR, u32 granularity) ---------------
// Polynomial of Mp elements (chunks of 2^(k+1)), planes [0, P), plane stride M.
// hi_fix: hi_i ^= XOR_{e in T_k, e > i} hi_{i + half - e}   for i < e_max
__global__ void qc_hi_fix(uint32_t* __restrict__ X, int64_t M, int P, int64_t Mp, int k) {
  uint32_t e[16];
  const int ne = lower_terms(k, e);
  if (ne == 0) return;
  const uint32_t emax = e[ne - 1];
  const int64_t half = (int64_t)1 << k;
  const int64_t nchunk = Mp >> (k + 1);
  const int64_t total = (int64_t)P * nchunk * emax * 4;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int w = (int)(t & 3);
    int64_t r = t >> 2;
    const int64_t i = r % emax;
    r /= emax;
    const int64_t g = r % nchunk;
    const int p = (int)(r / nchunk);
    uint32_t* base = X + ((int64_t)p * M + g * 2 * half + half) * 4 + w;
    uint32_t v = base[i * 4];
    for (int q = 0; q < ne; q++)
      if ((int64_t)e[q] > i) v ^= base[(i + half - e[q]) * 4];
    base[i * 4] = v;
  }
}

// lo_update: lo_j ^= XOR_{e in T_k, e <= j} hi_{j - e}
__global__ void qc_lo_update(uint32_t* __restrict__ X, int64_t M, int P, int64_t Mp, int k) {
  uint32_t e[16];
  const int ne = lower_terms(k, e);
  if (ne == 0) return;
  const int64_t half = (int64_t)1 << k;
  const int64_t total = (int64_t)P * (Mp >> 1) * 4;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int w = (int)(t & 3);
    int64_t r = t >> 2;
    const int64_t half_elems = Mp >> 1;
    const int64_t li = r % half_elems;  // index among low halves
    const int p = (int)(r / half_elems);
    const int64_t g = li >> k, j = li & (half - 1);
    uint32_t* lo = X + ((int64_t)p * M + g * 2 * half) * 4 + w;
    const uint32_t* hi = lo + half * 4;
    uint32_t v = lo[j * 4];
    for (int q = 0; q < ne; q++)
      if ((int64_t)e[q] <= j) v ^= hi[(j - e[q]) * 4];
    lo[j * 4] = v;
  }
}

// duplicate the converted lower half into the upper half, zero planes 16..31
__global__ void qc_dup(uint4* __restrict__ X, int64_t M) {
  const int64_t half = M >> 1;
  const int64_t total = 32 * half;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)(t / half);
    const int64_t e = t % half;
    if (p < 16) {
      X[(int64_t)p * M + half + e] = X[(int64_t)p * M + e];
    } else {
      X[(int64_t)p * M + e] = make_uint4(0, 0, 0, 0);
      X[(int64_t)p * M + half + e] = make_uint4(0, 0, 0, 0);
    }
  }
}

// ---- FFT layer --------------------------------------------------------------
// lane = (butterfly b, word w): u = g*2^(k+1) + j, v = u + 2^k, c = C(g).
// forward: u ^= c*v; v ^= u.   inverse: v ^= u; u ^= c*v.
template <bool INV>
__global__ __launch_bounds__(256) void qc_fft_layer(uint32_t* __restrict__ X, int64_t M, int k, Beta beta) {
  const int64_t nb = M >> 1;
  const int64_t total = nb * 4;
  const int64_t half = (int64_t)1 << k;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int w = (int)(t & 3);
    const int64_t b = t >> 2;
    const int64_t g = b >> k, j = b & (half - 1);
    const int64_t u = g * 2 * half + j, v = u + half;
    const uint32_t c = chunk_const(beta.v, (uint64_t)g);
    uint32_t xu[32], xv[32], cv[32];
#pragma unroll
    for (int p = 0; p < 32; p++) {
      xu[p] = X[((int64_t)p * M + u) * 4 + w];
      xv[p] = X[((int64_t)p * M + v) * 4 + w];
    }
    if (!INV) {
      gf32_mulc<false>(xv, c, cv);
#pragma unroll
      for (int p = 0; p < 32; p++) {
        xu[p] ^= cv[p];
        xv[p] ^= xu[p];
      }
    } else {
#pragma unroll
      for (int p = 0; p < 32; p++) xv[p] ^= xu[p];
      gf32_mulc<false>(xv, c, cv);
#pragma unroll
      for (int p = 0; p < 32; p++) xu[p] ^= cv[p];
    }
#pragma unroll
    for (int p = 0; p < 32; p++) {
      X[((int64_t)p * M + u) * 4 + w] = xu[p];
      X[((int64_t)p * M + v) * 4 + w] = xv[p];
    }
  }
}

__global__ __launch_bounds__(256) void qc_pointwise(uint32_t* __restrict__ X, int64_t M,
                                                    const uint32_t* __restrict__ ahat) {
  const int64_t total = M * 4;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int w = (int)(t & 3);
    const int64_t e = t >> 2;
    uint32_t x[32], y[32];
#pragma unroll
    for (int p = 0; p < 32; p++) x[p] = X[((int64_t)p * M + e) * 4 + w];
    gf32_mulc<false>(x, ahat[e], y);
#pragma unroll
    for (int p = 0; p < 32; p++) X[((int64_t)p * M + e) * 4 + w] = y[p];
  }
}

// A^[e] = element of row 0: bit j = bit 0 of plane j word 0
__global__ void qc_extract_row0(const uint32_t* __restrict__ X, int64_t M, uint32_t* __restrict__ ahat) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < M; e += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = 0;
    for (int p = 0; p < 32; p++) v |= (X[((int64_t)p * M + e) * 4] & 1u) << p;
    ahat[e] = v;
  }
}

// y_i = x0_i ^ H(i) ^ H(i + n), H(s) = plane (s&15) of chunk s>>4 ^ plane
// (s&15)+16 of chunk (s>>4)-1.
CUOT_DEV uint4 hcontrib(const uint4* __restrict__ X, int64_t M, int64_t s) {
  const int64_t c = s >> 4;
  const int b = (int)(s & 15);
  uint4 v = make_uint4(0, 0, 0, 0);
  if (c < M) v = X[(int64_t)b * M + c];
  if (c >= 1 && c - 1 < M) {
    uint4 w = X[(int64_t)(b + 16) * M + c - 1];
    v.x ^= w.x;
    v.y ^= w.y;
    v.z ^= w.z;
    v.w ^= w.w;
  }
  return v;
}

__global__ void qc_output(const uint4* __restrict__ X, int64_t M, const uint4* __restrict__ x0, int64_t n,
                          uint4* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint4 a = hcontrib(X, M, i), b = hcontrib(X, M, i + n), x = x0[i];
    y[i] = make_uint4(x.x ^ a.x ^ b.x, x.y ^ a.y ^ b.y, x.z ^ a.z ^ b.z, x.w ^ a.w ^ b.w);
  }
}

unsigned grid_for(int64_t total, int64_t cap = 1 << 16) {
  int64_t g = (total + 255) / 256;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (unsigned)g;
}

Beta g_beta;
bool g_beta_ready = false;

const Beta& beta() {
  if (!g_beta_ready) {
    gf32_cantor_basis(g_beta.v);
    g_beta_ready = true;
  }
  return g_beta;
}

// forward transform of the packed lower half (planes 0..15) into X (32 planes)
hipError_t forward(uint32_t* X, int m, hipStream_t st) {
  const int64_t M = (int64_t)1 << m;
  const int64_t Mh = M >> 1;
  for (int k = m - 2; k >= 0; k--) {
    int64_t nchunk = Mh >> (k + 1);
    hipLaunchKernelGGL(qc_hi_fix, dim3(grid_for(16 * nchunk * ((int64_t)1 << (k > 0 ? k - 1 : 0)) * 4)),
                       dim3(256), 0, st, X, M, 16, Mh, k);
    hipLaunchKernelGGL(qc_lo_update, dim3(grid_for(16 * (Mh >> 1) * 4)), dim3(256), 0, st, X, M, 16, Mh, k);
  }
  hipLaunchKernelGGL(qc_dup, dim3(grid_for(32 * Mh)), dim3(256), 0, st, (uint4*)X, M);
  for (int k = m - 2; k >= 0; k--)
    hipLaunchKernelGGL(qc_fft_layer<false>, dim3(grid_for(M * 2)), dim3(256), 0, st, X, M, k, beta());
  return hipGetLastError();
}

hipError_t inverse(uint32_t* X, int m, hipStream_t st) {
  const int64_t M = (int64_t)1 << m;
  for (int k = 0; k <= m - 1; k++)
    hipLaunchKernelGGL(qc_fft_layer<true>, dim3(grid_for(M * 2)), dim3(256), 0, st, X, M, k, beta());
  for (int k = 0; k <= m - 1; k++) {
    int64_t nchunk = M >> (k + 1);
    hipLaunchKernelGGL(qc_lo_update, dim3(grid_for(32 * (M >> 1) * 4)), dim3(256), 0, st, X, M, 32, M, k);
    hipLaunchKernelGGL(qc_hi_fix, dim3(grid_for(32 * nchunk * ((int64_t)1 << (k > 0 ? k - 1 : 0)) * 4)),
                       dim3(256), 0, st, X, M, 32, M, k);
  }
  return hipGetLastError();
}

}  // namespace

int cuot_qc_log_len(int64_t n) {
  const int64_t nc = (n + 15) / 16;
  int m = 1;
  while (((int64_t)1 << m) < 2 * nc - 1) m++;
  return m;
}

size_t cuot_qc_scratch_bytes_for(int64_t n) {
  const int m = cuot_qc_log_len(n);
  return (size_t)32 * ((size_t)1 << m) * 16;
}

hipError_t cuot_launch_qc_prepare(const uint8_t* a_bits, int64_t n, uint32_t* ahat, void* scratch, hipStream_t st) {
  const int m = cuot_qc_log_len(n);
  const int64_t M = (int64_t)1 << m;
  uint32_t* X = (uint32_t*)scratch;
  hipLaunchKernelGGL(qc_pack_a, dim3(grid_for(16 * (M >> 1))), dim3(256), 0, st, a_bits, n, (uint4*)X, M);
  hipError_t e = forward(X, m, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(qc_extract_row0, dim3(grid_for(M)), dim3(256), 0, st, X, M, ahat);
  return hipGetLastError();
}

hipError_t cuot_launch_qc_encode(const uint32_t* ahat, int64_t n, const void* x, void* y, void* scratch,
                                 hipStream_t st) {
  const int m = cuot_qc_log_len(n);
  const int64_t M = (int64_t)1 << m;
  uint32_t* X = (uint32_t*)scratch;
  const uint4* x0 = (const uint4*)x;
  hipLaunchKernelGGL(qc_pack_x, dim3(grid_for(16 * (M >> 1))), dim3(256), 0, st, x0 + n, n, (uint4*)X, M);
  hipError_t e = forward(X, m, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(qc_pointwise, dim3(grid_for(M * 4)), dim3(256), 0, st, X, M, ahat);
  e = inverse(X, m, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(qc_output, dim3(grid_for(n)), dim3(256), 0, st, (const uint4*)X, M, x0, n, (uint4*)y);
  return hipGetLastError();
}
