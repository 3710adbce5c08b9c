// bits.hip -- bit-level kernels around the QC code:
//   * sparse encode of the receiver's regular noise (choice bits):
//       b_i = e_i ^ XOR_{p in pos, p >= n} a_{(i - (p - n)) mod n},  i < n,
//     i.e. the same code applied to the weight-t vector e (the reference's
//     cyclic_mat_vec_prod + Mat::modp, silent/quasi_cyclic.cu:130-161, on the
//     workload's 2n -> n code);
//   * the 128-row bit transpose with the reference's bit order
//     (gpu/gpu_ops.cu:31-70): row k bit m = bit k of block m, bit k = byte k/8
//     bit k%8, LSB-first within bytes -- done as 32x32 in-register transposes
//     across 32 lanes, staged through LDS so both sides are coalesced.
#include "kernels.h"
#include "transpose32.h"

namespace {

// 32 bits of the cyclic n-bit vector a starting at bit `start` (bit b of the
// result = a[(start + b) mod n]).  a is padded with at least one zero word.
CUOT_DEV uint32_t cyc_window(const uint32_t* __restrict__ a, int64_t n, int64_t start) {
  uint32_t out = 0;
  int got = 0;
  while (got < 32) {
    const int64_t s = start >= n ? start - n : start;  // start < 2n on every call
    int64_t avail = n - s;  // bits before wrap
    int take = (int)((32 - got) < avail ? (32 - got) : avail);
    const int64_t wi = s >> 5;
    const int sh = (int)(s & 31);
    uint64_t two = (uint64_t)a[wi] | ((uint64_t)a[wi + 1] << 32);
    uint32_t bits = (uint32_t)(two >> sh);
    if (take < 32) bits &= (1u << take) - 1u;
    out |= bits << got;
    got += take;
    start = s + take;
  }
  return out;
}

__global__ void sparse_encode_kernel(const uint32_t* __restrict__ a, int64_t n, const int64_t* __restrict__ pos,
                                     int w, uint32_t* __restrict__ out) {
  const int64_t nw = (n + 31) / 32;
  for (int64_t W = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; W < nw; W += (int64_t)gridDim.x * blockDim.x) {
    uint32_t acc = 0;
    const int64_t lo = W * 32;
    for (int q = 0; q < w; q++) {
      const int64_t p = pos[q];
      if (p < 0 || p >= 2 * n) continue;
      if (p < n) {
        if ((p >> 5) == W) acc ^= 1u << (p & 31);
      } else {
        const int64_t s = p - n;  // out_i ^= a_{(i - s) mod n}
        int64_t start = lo - s;  // in (-n, n): lo < n and 0 <= s < n
        if (start < 0) start += n;
        acc ^= cyc_window(a, n, start);
      }
    }
    if (lo + 32 > n) acc &= (1u << (n - lo)) - 1u;
    out[W] = acc;
  }
}

// n a multiple of 32 (and at least 64 words): the window of position s starts
// at the same bit offset sb = s mod 32 in every word, so output word W takes
// (a[W - so] << sb) | (a[W - so - 1] >> (32 - sb)) with so = s / 32 (indices mod
// n / 32); SPR consecutive output words per lane share their loads.
constexpr int SPR = 2;  // measured at n = 2^24: 1 -> 0.31 ms, 2 -> 0.27 ms, 4 -> 0.37 ms
__global__ void sparse_encode_w32(const uint32_t* __restrict__ a, int64_t n, const int64_t* __restrict__ pos, int w,
                                  uint32_t* __restrict__ out) {
  const int64_t nw = n >> 5;
  for (int64_t W0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * SPR; W0 < nw;
       W0 += (int64_t)gridDim.x * blockDim.x * SPR) {
    uint32_t acc[SPR];
#pragma unroll
    for (int r = 0; r < SPR; r++) acc[r] = 0u;
    for (int q = 0; q < w; q++) {
      const int64_t p = pos[q];
      if (p < 0 || p >= 2 * n) continue;
      if (p < n) {
        const int64_t pw = (p >> 5) - W0;
#pragma unroll
        for (int r = 0; r < SPR; r++)
          if (pw == r) acc[r] ^= 1u << (p & 31);
        continue;
      }
      const int64_t sft = p - n;  // out_i ^= a_{(i - sft) mod n}
      const int64_t so = sft >> 5;
      const int sb = (int)(sft & 31);
      int64_t i = W0 - so - 1;  // in [-nw, nw)
      if (i < 0) i += nw;
      uint32_t prev = a[i];
#pragma unroll
      for (int r = 0; r < SPR; r++) {
        if (++i == nw) i = 0;
        const uint32_t cur = a[i];
        acc[r] ^= sb ? (cur << sb) | (prev >> (32 - sb)) : cur;
        prev = cur;
      }
    }
#pragma unroll
    for (int r = 0; r < SPR; r++)
      if (W0 + r < nw) out[W0 + r] = acc[r];
  }
}

// ---- transpose ---------------------------------------------------------------
CUOT_DEV uint32_t transpose32(uint32_t x, int lane) { return cuot_transpose32(x, lane); }

// blocks -> rows.  WG: 256 lanes, 1024 blocks (32 tiles of 32 blocks).
// rows: 128 rows of N bits, row stride N/32 words.  blockIdx.y selects a plane
// (input plane stride in_ps blocks, output plane stride out_ps words).
__global__ __launch_bounds__(256) void transpose_fwd(const uint4* __restrict__ in, int64_t N,
                                                     uint32_t* __restrict__ rows, int P, int64_t in_is, int64_t in_ps,
                                                     int64_t out_is, int64_t out_ps) {
  __shared__ uint32_t sm[128][33];
  const int inst = (int)blockIdx.y / P, pl = (int)blockIdx.y % P;
  in += inst * in_is + pl * in_ps;
  rows += inst * out_is + pl * out_ps;
  const int lane = threadIdx.x & 31;
  const int grp = threadIdx.x >> 5;  // 8 groups of 32 lanes
  const int64_t ntile = N / 32;
  for (int64_t t0 = (int64_t)blockIdx.x * 32; t0 < ntile; t0 += (int64_t)gridDim.x * 32) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int tl = grp * 4 + r;  // tile within the group of 32 tiles
      const int64_t tile = t0 + tl;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (tile < ntile) v = in[tile * 32 + lane];
      sm[0 * 32 + lane][tl] = transpose32(v.x, lane);
      sm[1 * 32 + lane][tl] = transpose32(v.y, lane);
      sm[2 * 32 + lane][tl] = transpose32(v.z, lane);
      sm[3 * 32 + lane][tl] = transpose32(v.w, lane);
    }
    __syncthreads();
    const int64_t rw = N / 32;  // words per row
    if ((N & 127) == 0) {
      // rows of a multiple of 4 words: 16-byte stores (4 words of one row per lane)
      for (int f = threadIdx.x; f < 128 * 8; f += 256) {
        const int k = f >> 3, tl = (f & 7) * 4;
        if (t0 + tl < ntile)
          *(uint4*)&rows[(int64_t)k * rw + t0 + tl] = make_uint4(sm[k][tl], sm[k][tl + 1], sm[k][tl + 2], sm[k][tl + 3]);
      }
    } else {
      for (int f = threadIdx.x; f < 128 * 32; f += 256) {
        const int k = f >> 5, tl = f & 31;
        if (t0 + tl < ntile) rows[(int64_t)k * rw + t0 + tl] = sm[k][tl];
      }
    }
    __syncthreads();
  }
}

// rows -> blocks
__global__ __launch_bounds__(256) void transpose_inv(const uint32_t* __restrict__ rows, int64_t N,
                                                     uint4* __restrict__ out, int P, int64_t in_is, int64_t in_ps,
                                                     int64_t out_is, int64_t out_ps) {
  __shared__ uint32_t sm[128][33];
  const int inst = (int)blockIdx.y / P, pl = (int)blockIdx.y % P;
  rows += inst * in_is + pl * in_ps;
  out += inst * out_is + pl * out_ps;
  const int lane = threadIdx.x & 31;
  const int grp = threadIdx.x >> 5;
  const int64_t ntile = N / 32;
  const int64_t rw = N / 32;
  for (int64_t t0 = (int64_t)blockIdx.x * 32; t0 < ntile; t0 += (int64_t)gridDim.x * 32) {
    if ((N & 127) == 0) {
      // rows of a multiple of 4 words: 16-byte loads (4 words of one row per lane)
      for (int f = threadIdx.x; f < 128 * 8; f += 256) {
        const int k = f >> 3, tl = (f & 7) * 4;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (t0 + tl < ntile) v = *(const uint4*)&rows[(int64_t)k * rw + t0 + tl];
        sm[k][tl] = v.x;
        sm[k][tl + 1] = v.y;
        sm[k][tl + 2] = v.z;
        sm[k][tl + 3] = v.w;
      }
    } else {
      for (int f = threadIdx.x; f < 128 * 32; f += 256) {
        const int k = f >> 5, tl = f & 31;
        sm[k][tl] = (t0 + tl < ntile) ? rows[(int64_t)k * rw + t0 + tl] : 0u;
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int tl = grp * 4 + r;
      const int64_t tile = t0 + tl;
      uint4 v;
      v.x = transpose32(sm[0 * 32 + lane][tl], lane);
      v.y = transpose32(sm[1 * 32 + lane][tl], lane);
      v.z = transpose32(sm[2 * 32 + lane][tl], lane);
      v.w = transpose32(sm[3 * 32 + lane][tl], lane);
      if (tile < ntile) out[tile * 32 + lane] = v;
    }
    __syncthreads();
  }
}

unsigned grid_for(int64_t total, int64_t cap = 1 << 16) {
  int64_t g = (total + 255) / 256;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace

hipError_t cuot_launch_sparse(const uint32_t* a_words, int64_t n, const int64_t* pos, int w, uint32_t* out,
                              hipStream_t st) {
  if ((n & 31) == 0 && (n >> 5) >= 64) {
    hipLaunchKernelGGL(sparse_encode_w32, dim3(grid_for(((n >> 5) + SPR - 1) / SPR)), dim3(256), 0, st, a_words, n,
                       pos, w, out);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(sparse_encode_kernel, dim3(grid_for((n + 31) / 32)), dim3(256), 0, st, a_words, n, pos, w,
                     out);
  return hipGetLastError();
}

hipError_t cuot_launch_transpose_planes2(const void* in, int64_t N, void* out, bool inverse, int planes, int P,
                                        int64_t in_is, int64_t in_ps, int64_t out_is, int64_t out_ps,
                                        hipStream_t st) {
  if (planes < 1 || P < 1 || planes > 65535) return hipErrorInvalidValue;
  int64_t ntile = N / 32;
  int64_t g = (ntile + 31) / 32;
  int64_t cap = planes > 1 ? 16384 / planes : 8192;
  if (cap < 1) cap = 1;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  dim3 grid((unsigned)g, (unsigned)planes);
  cuot_rec_begin(st, inverse ? "transpose_inv" : "transpose_fwd", 2.0 * 16.0 * (double)N * planes);
  if (!inverse)
    hipLaunchKernelGGL(transpose_fwd, grid, dim3(256), 0, st, (const uint4*)in, N, (uint32_t*)out, P, in_is, in_ps,
                       out_is, out_ps);
  else
    hipLaunchKernelGGL(transpose_inv, grid, dim3(256), 0, st, (const uint32_t*)in, N, (uint4*)out, P, in_is, in_ps,
                       out_is, out_ps);
  cuot_rec_end(st);
  return hipGetLastError();
}

hipError_t cuot_launch_transpose_planes(const void* in, int64_t N, void* out, bool inverse, int planes,
                                       int64_t in_ps, int64_t out_ps, hipStream_t st) {
  return cuot_launch_transpose_planes2(in, N, out, inverse, planes, planes, 0, in_ps, 0, out_ps, st);
}

hipError_t cuot_launch_transpose(const void* in, int64_t N, void* out, bool inverse, hipStream_t st) {
  return cuot_launch_transpose_planes(in, N, out, inverse, 1, 0, 0, st);
}
