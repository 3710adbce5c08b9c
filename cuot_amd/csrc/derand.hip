// derand.hip -- online derandomisation of random COTs for both parties:
//   * chosen-bit COT (FerretCOT::online_sender / online_recver,
//     ferret_cot.hpp:277-295): bo_i = b_i ^ LSB(M_i) (the receiver's message),
//     K_i ^= bo_i * Delta, so that M_i = K_i ^ b_i * Delta;
//   * COT -> OT (COT::send / recv, cot.h:16-56): the sender's pads
//     pad_2i = H_i(K_i) ^ m0_i, pad_2i+1 = H_i(K_i ^ Delta) ^ m1_i and the
//     receiver's output pad_{2i + r_i} ^ H_i(M_i) = m_{r_i, i};
//   * COT -> random OT (COT::send_rot / recv_rot, cot.h:58-91):
//     H_i(K_i), H_i(K_i ^ Delta) and H_i(M_i).
// H_i is emp-tool's multi-instance tweakable hash MITCCRH (mitccrh.h of
// emp-tool @44b1dde, absent here): H_i(x) = AES_{k_i}(x) ^ x with one key per
// COT, k_i = s ^ makeBlock(g0 + i, 0) (s the sender's random block, g0 the
// number of keys earlier calls consumed).  Each lane expands its own key with
// the LDS S-box, then runs one (receiver) or two (sender) encryptions.
#include <string.h>

#include "aes_dev.h"
#include "ctx.h"
#include "kernels.h"

namespace {

// S-box byte x from the T-table (T0[x] bytes are 2s, s, s, 3s)
CUOT_DEV uint32_t sbox(const AesLds& a, uint32_t x) { return (tlook(a, x) >> 8) & 0xffu; }

CUOT_DEV uint32_t sub_word(const AesLds& a, uint32_t w) {
  return sbox(a, w & 0xffu) | (sbox(a, (w >> 8) & 0xffu) << 8) | (sbox(a, (w >> 16) & 0xffu) << 16) |
         (sbox(a, w >> 24) << 24);
}

// FIPS-197 key expansion on little-endian words (word i = bytes 4i..4i+3)
CUOT_DEV void expand_key(const AesLds& a, blk key, uint32_t rk[44]) {
  rk[0] = key.x;
  rk[1] = key.y;
  rk[2] = key.z;
  rk[3] = key.w;
  uint32_t rcon = 1;
#pragma unroll
  for (int i = 4; i < 44; i++) {
    uint32_t t = rk[i - 1];
    if ((i & 3) == 0) {
      t = sub_word(a, (t >> 8) | (t << 24)) ^ rcon;  // RotWord, SubWord, Rcon on byte 0
      rcon = (rcon << 1) ^ ((rcon & 0x80u) ? 0x11bu : 0u);
    }
    rk[i] = rk[i - 4] ^ t;
  }
}

struct DerandArgs {
  const uint32_t* T0;
  blk* ks;             // sender K (online: updated in place)
  const blk* mr;       // receiver M
  const blk* delta;    // device Delta
  const uint8_t* bits; // receiver bits (online: chosen b; OT: choice r), one byte each
  uint8_t* bo;         // online: receiver's message bits
  blk s;               // MITCCRH start point
  uint64_t g0;         // MITCCRH key counter of this call's first COT
  int64_t n;
  const blk* m0;       // OT: sender messages
  const blk* m1;
  blk* pad;            // OT: 2n sender pads
  blk* out_r;          // OT: receiver output; ROT: H(M)
  blk* rot0;           // ROT: H(K), H(K ^ Delta)
  blk* rot1;
};

__global__ __launch_bounds__(256) void cot_online_kernel(DerandArgs a) {
  const blk D = *a.delta;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * 256) {
    const uint8_t bo = (uint8_t)((a.bits[i] & 1u) ^ (a.mr[i].x & 1u));
    a.bo[i] = bo;
    if (bo) a.ks[i] = bxor(a.ks[i], D);
  }
}

template <bool ROT>
__global__ __launch_bounds__(256) void cot_hash_kernel(DerandArgs a) {
  __shared__ uint32_t tab[CUOT_AES_LDS_WORDS];
  aes_stage_table(tab, a.T0, threadIdx.x, 256);
  __syncthreads();
  const AesLds aes = aes_lds(tab);
  const blk D = *a.delta;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * 256) {
    const uint64_t g = a.g0 + (uint64_t)i;  // makeBlock(g, 0): hi = g, lo = 0
    const blk key = make_uint4(a.s.x, a.s.y, a.s.z ^ (uint32_t)g, a.s.w ^ (uint32_t)(g >> 32));
    uint32_t rk[44];
    expand_key(aes, key, rk);
    const blk k0 = a.ks[i], k1 = bxor(k0, D), m = a.mr[i];
    const blk h0 = bxor(aes_enc(aes, rk, k0), k0);
    const blk h1 = bxor(aes_enc(aes, rk, k1), k1);
    const blk hm = bxor(aes_enc(aes, rk, m), m);
    if (ROT) {
      a.rot0[i] = h0;
      a.rot1[i] = h1;
      a.out_r[i] = hm;
    } else {
      const blk p0 = bxor(h0, a.m0[i]), p1 = bxor(h1, a.m1[i]);
      a.pad[2 * i] = p0;
      a.pad[2 * i + 1] = p1;
      a.out_r[i] = bxor((a.bits[i] & 1u) ? p1 : p0, hm);
    }
  }
}

unsigned grid_of(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace

hipError_t cuot_launch_cot_online(blk* ks, const blk* mr, const blk* delta, const uint8_t* b, uint8_t* bo, int64_t n,
                                  hipStream_t st) {
  if (n <= 0) return hipSuccess;
  DerandArgs a;
  memset(&a, 0, sizeof(a));
  a.ks = ks;
  a.mr = mr;
  a.delta = delta;
  a.bits = b;
  a.bo = bo;
  a.n = n;
  hipLaunchKernelGGL(cot_online_kernel, dim3(grid_of(n)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t cuot_launch_cot_hash(const blk* ks, const blk* mr, const blk* delta, blk s, uint64_t g0, int64_t n,
                                const uint8_t* r, const blk* m0, const blk* m1, blk* pad, blk* out_r, blk* rot0,
                                blk* rot1, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  DerandArgs a;
  memset(&a, 0, sizeof(a));
  a.T0 = cuot_device_T0();
  a.ks = (blk*)ks;
  a.mr = mr;
  a.delta = delta;
  a.bits = r;
  a.s = s;
  a.g0 = g0;
  a.n = n;
  a.m0 = m0;
  a.m1 = m1;
  a.pad = pad;
  a.out_r = out_r;
  a.rot0 = rot0;
  a.rot1 = rot1;
  const bool rot = rot0 != nullptr;
  // three AES (and one key expansion) per COT; 16 B in, 48 B (OT) out per party pair
  cuot_rec_begin(st, rot ? "cot_hash_kernel<rot>" : "cot_hash_kernel<ot>", 80.0 * (double)n);
  if (rot)
    hipLaunchKernelGGL(cot_hash_kernel<true>, dim3(grid_of(n)), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(cot_hash_kernel<false>, dim3(grid_of(n)), dim3(256), 0, st, a);
  cuot_rec_end(st);
  return hipGetLastError();
}

// ---- C ABI ----------------------------------------------------------------------
extern "C" int cuot_cot_online(cuot_ctx* ctx, cuot_block* data_s_d, const cuot_block* data_r_d,
                               const uint8_t* b_d, const cuot_block* delta_d, uint8_t* bo_d, int64_t n,
                               void* stream) {
  CUOT_CHECK_ARG(ctx, "NULL context");
  CUOT_CHECK_ARG(n >= 0, "n must be >= 0");
  if (n == 0) return CUOT_OK;
  CUOT_CHECK_ARG(data_s_d && data_r_d && b_d && delta_d && bo_d, "buffer is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  CUOT_HIP_TRY(cuot_launch_cot_online((blk*)data_s_d, (const blk*)data_r_d, (const blk*)delta_d, b_d, bo_d, n,
                                      ctx->st(stream)));
  return CUOT_OK;
}

extern "C" int cuot_cot_to_ot(cuot_ctx* ctx, const cuot_block* data_s_d, const cuot_block* data_r_d,
                              const cuot_block* delta_d, cuot_block s, uint64_t key_start, const uint8_t* r_d,
                              const cuot_block* m0_d, const cuot_block* m1_d, cuot_block* pad_d,
                              cuot_block* out_r_d, int64_t n, void* stream) {
  CUOT_CHECK_ARG(ctx, "NULL context");
  CUOT_CHECK_ARG(n >= 0, "n must be >= 0");
  if (n == 0) return CUOT_OK;
  CUOT_CHECK_ARG(data_s_d && data_r_d && delta_d && r_d && m0_d && m1_d && pad_d && out_r_d, "buffer is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  blk sb;
  memcpy(&sb, &s, 16);
  CUOT_HIP_TRY(cuot_launch_cot_hash((const blk*)data_s_d, (const blk*)data_r_d, (const blk*)delta_d, sb, key_start,
                                    n, r_d, (const blk*)m0_d, (const blk*)m1_d, (blk*)pad_d, (blk*)out_r_d,
                                    nullptr, nullptr, ctx->st(stream)));
  return CUOT_OK;
}

extern "C" int cuot_cot_to_rot(cuot_ctx* ctx, const cuot_block* data_s_d, const cuot_block* data_r_d,
                               const cuot_block* delta_d, cuot_block s, uint64_t key_start, cuot_block* data0_d,
                               cuot_block* data1_d, cuot_block* out_r_d, int64_t n, void* stream) {
  CUOT_CHECK_ARG(ctx, "NULL context");
  CUOT_CHECK_ARG(n >= 0, "n must be >= 0");
  if (n == 0) return CUOT_OK;
  CUOT_CHECK_ARG(data_s_d && data_r_d && delta_d && data0_d && data1_d && out_r_d, "buffer is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  blk sb;
  memcpy(&sb, &s, 16);
  CUOT_HIP_TRY(cuot_launch_cot_hash((const blk*)data_s_d, (const blk*)data_r_d, (const blk*)delta_d, sb, key_start,
                                    n, nullptr, nullptr, nullptr, nullptr, (blk*)out_r_d, (blk*)data0_d,
                                    (blk*)data1_d, ctx->st(stream)));
  return CUOT_OK;
}
