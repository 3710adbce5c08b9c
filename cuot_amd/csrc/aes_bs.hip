// aes_bs.hip -- AES-128 ECB / counter-mode streams with the bitsliced cipher
// (aes_bs.h): no tables and no LDS.  One lane encrypts 32 blocks; a wave
// covers 2048 consecutive blocks, lane l taking blocks 64 k + l (k = 0..31),
// so every load and store instruction of the wave touches 1 KiB of
// consecutive memory.  The T-table kernels (aes.hip) and these compute the
// same function; the knob aes_impl picks one (execution choice only).
#include "aes_bs.h"
#include "kernels.h"

namespace {

struct BsArgs {
  uint32_t rk[44];  // round keys with the S-box constant folded in
  const blk* in;
  blk* out;
  int64_t n;
  uint64_t stream;  // counter mode: high word of the counter block
  uint64_t start;   // counter mode: first counter
};

template <bool PRG>
__global__ __launch_bounds__(256) void aes_bs_stream_kernel(BsArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t g0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 2048;
  uint32_t S[128];
#pragma unroll
  for (int k = 0; k < 32; k++) {
    const int64_t i = g0 + 64 * k + lane;
    blk x = bzero();
    if (PRG) {
      const uint64_t c = a.start + (uint64_t)i;
      x = make_uint4((uint32_t)c, (uint32_t)(c >> 32), (uint32_t)a.stream, (uint32_t)(a.stream >> 32));
    } else if (i < a.n) {
      x = a.in[i];
    }
    S[k] = x.x;
    S[32 + k] = x.y;
    S[64 + k] = x.z;
    S[96 + k] = x.w;
  }
#pragma unroll
  for (int w = 0; w < 4; w++) bs_transpose32(S + 32 * w);
  bs_encrypt(S, a.rk);
#pragma unroll
  for (int w = 0; w < 4; w++) bs_transpose32(S + 32 * w);
#pragma unroll
  for (int k = 0; k < 32; k++) {
    const int64_t i = g0 + 64 * k + lane;
    if (i < a.n) a.out[i] = make_uint4(S[k], S[32 + k], S[64 + k], S[96 + k]);
  }
}

hipError_t launch(BsArgs& a, const uint32_t rk[44], bool prg, hipStream_t st) {
  bs_fold_sbox_constant(rk, a.rk);
  const unsigned grid = (unsigned)((a.n + 8191) / 8192);
  cuot_rec_begin(st, "aes_bs_stream", (prg ? 16.0 : 32.0) * (double)a.n);
  if (prg)
    hipLaunchKernelGGL(aes_bs_stream_kernel<true>, dim3(grid), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(aes_bs_stream_kernel<false>, dim3(grid), dim3(256), 0, st, a);
  cuot_rec_end(st);
  return hipGetLastError();
}

}  // namespace

hipError_t cuot_launch_aes_ecb_bs(const uint32_t rk[44], const blk* in, blk* out, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  BsArgs a;
  a.in = in;
  a.out = out;
  a.n = n;
  a.stream = a.start = 0;
  return launch(a, rk, false, st);
}

hipError_t cuot_launch_prg_bs(const uint32_t rk[44], uint64_t stream, uint64_t start, int64_t n, blk* out,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  BsArgs a;
  a.in = nullptr;
  a.out = out;
  a.n = n;
  a.stream = stream;
  a.start = start;
  return launch(a, rk, true, st);
}
