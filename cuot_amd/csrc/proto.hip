// proto.hip -- small protocol-glue kernels of the SOT / Ferret instances:
// seed-derived setup fix-ups, the SOT sender's masked base-OT rows, the
// receiver's view of them, and Ferret's secret sum and bootstrap COTs.
#include "kernels.h"

namespace {

unsigned grid_for(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > 65535) g = 65535;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// Delta |= 1; choices c[t][h] = 1 - bit (h-1-d) of alpha_t; pos_t = t*2^h + alpha;
// a bits beyond n cleared (words up to a_words).
__global__ void setup_fix_kernel(blk* delta, const blk* alpha_blocks, int64_t t, int h, uint8_t* choice,
                                 int64_t* pos, uint32_t* a_words, int64_t n, int64_t a_words_total) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if (tid == 0 && delta) delta->x |= 1u;
  for (int64_t i = tid; i < t; i += stride) {
    const uint32_t alpha = alpha_blocks[i].x & (uint32_t)((1ull << h) - 1);
    for (int d = 0; d < h; d++) choice[i * h + d] = (uint8_t)(1u - ((alpha >> (h - 1 - d)) & 1u));
    if (pos) pos[i] = (i << h) + alpha;
  }
  if (a_words) {
    for (int64_t wdx = tid; wdx < a_words_total; wdx += stride) {
      const int64_t lo = wdx * 32;
      if (lo >= n)
        a_words[wdx] = 0;
      else if (lo + 32 > n)
        a_words[wdx] &= (1u << (n - lo)) - 1u;
    }
  }
}

// silent/sender.cu:80-103: M0[d] = m0b[d] ^ L_d, M1[d] = m1b[d] ^ R_d (d < h);
// M0[h] = m0b[h-1] ^ R_{h-1} ^ Delta, M1[h] = m1b[h-1] ^ L_{h-1} ^ Delta.
__global__ void sot_msgs_kernel(const blk* sums, const blk* m0b, const blk* m1b, const blk* delta, int64_t t,
                                int h, blk* M0, blk* M1) {
  const blk D = *delta;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < t * h; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t tr = i / h;
    const int d = (int)(i % h);
    const blk L = sums[(tr * h + d) * 2], R = sums[(tr * h + d) * 2 + 1];
    M0[d * t + tr] = bxor(m0b[d * t + tr], L);
    M1[d * t + tr] = bxor(m1b[d * t + tr], R);
    if (d == h - 1) {
      M0[h * t + tr] = bxor(bxor(m0b[d * t + tr], R), D);
      M1[h * t + tr] = bxor(bxor(m1b[d * t + tr], L), D);
    }
  }
}

// The receiver's view (silent/recver.cu:143-157): mc = m_c base-OT message;
// chosen sum at level d = M_c[d] ^ mc[d]; final = M_c[h] ^ mc[h-1].
__global__ void sot_recv_prep_kernel(const uint8_t* choice, const blk* M0, const blk* M1, const blk* m0b,
                                     const blk* m1b, int64_t t, int h, blk* csum, blk* fin) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < t * h; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t tr = i / h;
    const int d = (int)(i % h);
    const int c = choice[tr * h + d];
    const blk mc = c ? m1b[d * t + tr] : m0b[d * t + tr];
    csum[tr * h + d] = bxor(c ? M1[d * t + tr] : M0[d * t + tr], mc);
    if (d == h - 1) fin[tr] = bxor(c ? M1[h * t + tr] : M0[h * t + tr], mc);
  }
}

// Ferret: secret_sum = (L_h ^ R_h) & ~1 ^ Delta (spcot_sender.h:49-58) and
// the receiver's chosen sums (OT of the level sums, mpcot_reg.h:146-157).
__global__ void ferret_sums_kernel(const blk* sums, const blk* delta, const uint8_t* choice, int64_t t, int h,
                                   blk* ssum, blk* csum, blk* lsum, blk* rsum) {
  const blk D = *delta;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < t * h; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t tr = i / h;
    const int d = (int)(i % h);
    const blk L = sums[i * 2], R = sums[i * 2 + 1];
    if (csum) csum[i] = choice[i] ? R : L;
    if (lsum) lsum[i] = L;
    if (rsum) rsum[i] = R;
    if (d == h - 1) {
      blk s = bxor(L, R);
      s.x &= ~1u;
      ssum[tr] = bxor(s, D);
    }
  }
}

// Ferret bootstrap COTs: K_j &= ~1; M_j = K_j ^ c_j * Delta.
__global__ void ferret_boot_kernel(blk* K, blk* Mr, const uint8_t* cbits, const blk* delta, int64_t k) {
  const blk D = *delta;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < k; j += (int64_t)gridDim.x * blockDim.x) {
    blk v = K[j];
    v.x &= ~1u;
    K[j] = v;
    const int c = (cbits[j >> 3] >> (j & 7)) & 1;
    Mr[j] = c ? bxor(v, D) : v;
  }
}

__global__ void interleave_kernel(const blk* __restrict__ a, const blk* __restrict__ b, int64_t k,
                                  blk* __restrict__ out) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < k; j += (int64_t)gridDim.x * blockDim.x) {
    out[2 * j] = a[j];
    out[2 * j + 1] = b[j];
  }
}

}  // namespace

hipError_t cuot_launch_interleave(const blk* a, const blk* b, int64_t k, blk* out, hipStream_t st) {
  hipLaunchKernelGGL(interleave_kernel, dim3(grid_for(k)), dim3(256), 0, st, a, b, k, out);
  return hipGetLastError();
}

hipError_t cuot_launch_setup_fix(blk* delta, const blk* alpha_blocks, int64_t t, int h, uint8_t* choice,
                                 int64_t* pos, uint32_t* a_words, int64_t n, int64_t a_words_total,
                                 hipStream_t st) {
  int64_t work = t > a_words_total ? t : a_words_total;
  hipLaunchKernelGGL(setup_fix_kernel, dim3(grid_for(work)), dim3(256), 0, st, delta, alpha_blocks, t, h, choice,
                     pos, a_words, n, a_words_total);
  return hipGetLastError();
}

hipError_t cuot_launch_sot_msgs(const blk* sums, const blk* m0b, const blk* m1b, const blk* delta, int64_t t,
                                int h, blk* M0, blk* M1, hipStream_t st) {
  hipLaunchKernelGGL(sot_msgs_kernel, dim3(grid_for(t * h)), dim3(256), 0, st, sums, m0b, m1b, delta, t, h, M0, M1);
  return hipGetLastError();
}

hipError_t cuot_launch_sot_recv_prep(const uint8_t* choice, const blk* M0, const blk* M1, const blk* m0b,
                                     const blk* m1b, int64_t t, int h, blk* csum, blk* fin, hipStream_t st) {
  hipLaunchKernelGGL(sot_recv_prep_kernel, dim3(grid_for(t * h)), dim3(256), 0, st, choice, M0, M1, m0b, m1b, t, h,
                     csum, fin);
  return hipGetLastError();
}

hipError_t cuot_launch_ferret_sums(const blk* sums, const blk* delta, const uint8_t* choice, int64_t t, int h,
                                   blk* ssum, blk* csum, blk* lsum, blk* rsum, hipStream_t st) {
  hipLaunchKernelGGL(ferret_sums_kernel, dim3(grid_for(t * h)), dim3(256), 0, st, sums, delta, choice, t, h, ssum,
                     csum, lsum, rsum);
  return hipGetLastError();
}

hipError_t cuot_launch_ferret_boot(blk* K, blk* Mr, const uint8_t* cbits, const blk* delta, int64_t k,
                                   hipStream_t st) {
  hipLaunchKernelGGL(ferret_boot_kernel, dim3(grid_for(k)), dim3(256), 0, st, K, Mr, cbits, delta, k);
  return hipGetLastError();
}
