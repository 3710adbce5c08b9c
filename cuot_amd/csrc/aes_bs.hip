// aes_bs.hip -- AES-128 ECB / counter-mode streams and a breadth-first GGM
// sender with the bitsliced cipher (aes_bs.h): no tables and no LDS.  One lane encrypts 32 blocks; a wave
// covers 2048 consecutive blocks, lane l taking blocks 64 k + l (k = 0..31),
// so every load and store instruction of the wave touches 1 KiB of
// consecutive memory.  The T-table kernels (aes.hip) and these compute the
// same function; the knob aes_impl picks one (execution choice only).
#include <string.h>

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

// ---- GGM sender on the bitsliced cipher ------------------------------------
// Breadth-first, one launch per level: lane g takes the 32 consecutive
// parents 32g .. 32g+31 of the level, encrypts them under k0 and under k1
// (32 blocks per lane, bit-sliced), adds the parent (child = AES_k(p) ^ p) and
// writes children 2j (k0) and 2j+1 (k1).  The level sums of every tree are
// XORed into sums[tree][d][side] with 64-bit atomics: per block for levels
// with fewer than 32 nodes per tree, else reduced over the lanes of one tree
// first (2^(d-5) lanes, aligned groups of a wave).
struct GgmBsArgs {
  uint32_t rk[2][44];  // k0, k1 round keys with the S-box constant folded in
  const blk* par;      // t * 2^d parents
  blk* out;            // children, tree-major (the leaves at the last level)
  blk* sums;           // [t][h][2]
  int64_t np;          // parents
  int64_t n_out;       // children with index >= n_out are not written
  int d, h;
  int clear_lsb;       // last level in Ferret mode: leaves with bit 0 cleared
};

CUOT_DEV void sums_xor(blk* s, blk v) {
  unsigned long long* q = (unsigned long long*)s;
  atomicXor(q, (unsigned long long)v.x | ((unsigned long long)v.y << 32));
  atomicXor(q + 1, (unsigned long long)v.z | ((unsigned long long)v.w << 32));
}

__global__ __launch_bounds__(256) void ggm_bs_level(GgmBsArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t j0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 32;
  for (int side = 0; side < 2; side++) {
    uint32_t S[128];
#pragma unroll
    for (int k = 0; k < 32; k++) {
      const blk p = j0 + k < a.np ? a.par[j0 + k] : bzero();
      S[k] = p.x;
      S[32 + k] = p.y;
      S[64 + k] = p.z;
      S[96 + k] = p.w;
    }
#pragma unroll
    for (int w = 0; w < 4; w++) bs_transpose32(S + 32 * w);
    bs_encrypt(S, a.rk[side]);
#pragma unroll
    for (int w = 0; w < 4; w++) bs_transpose32(S + 32 * w);
    blk acc = bzero();
#pragma unroll
    for (int k = 0; k < 32; k++) {
      const int64_t j = j0 + k;
      if (j >= a.np) continue;
      const blk c = bxor(make_uint4(S[k], S[32 + k], S[64 + k], S[96 + k]), a.par[j]);
      const int64_t ci = 2 * j + side;
      if (ci < a.n_out) stb(a.out + ci, a.clear_lsb ? make_uint4(c.x & ~1u, c.y, c.z, c.w) : c);
      if (a.d < 5)
        sums_xor(a.sums + ((j >> a.d) * a.h + a.d) * 2 + side, c);
      else
        acc = bxor(acc, c);
    }
    if (a.d >= 5) {
      // lanes of one tree: 2^(d-5), capped at the wave
      const int G = a.d - 5 >= 6 ? 64 : 1 << (a.d - 5);
      for (int o = G >> 1; o >= 1; o >>= 1) {
        acc.x ^= (uint32_t)__shfl_xor((int)acc.x, o, 64);
        acc.y ^= (uint32_t)__shfl_xor((int)acc.y, o, 64);
        acc.z ^= (uint32_t)__shfl_xor((int)acc.z, o, 64);
        acc.w ^= (uint32_t)__shfl_xor((int)acc.w, o, 64);
      }
      const int64_t tree = j0 >> a.d;
      if ((lane & (G - 1)) == 0 && j0 < a.np) sums_xor(a.sums + (tree * a.h + a.d) * 2 + side, acc);
    }
  }
}

__global__ void ggm_bs_zero(blk* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = bzero();
}

// Receiver, after level d's expansion of its own nodes (the punctured node
// holds 0, so its children are AES_k0(0), AES_k1(0), taken out of the sums):
// the chosen sum gives the sibling of the punctured path, the other child
// stays unknown (0) and becomes the punctured node of the next level; at the
// last level it takes the final rule (SOT: fin ^ opposite-side sum, Ferret:
// fin ^ the LSB-cleared xor of the other leaves).  One lane per tree.
struct GgmBsFix {
  const blk* sums;  // [t][h][2] of the receiver's nodes
  int64_t* punct;   // punctured index per tree at level d
  const uint8_t* choice;
  const blk* csum;
  const blk* fin;
  blk* out;       // level d+1 nodes, or the leaves
  int64_t n_out;  // leaves written below this index (last level)
  int64_t t;
  int d, h, last, ferret;
  blk z0, z1;  // children of a zero parent
};

// one tree per lane, grid-stride (the grid is capped at small_grid's size)
__global__ void ggm_bs_fix(GgmBsFix a) {
  for (int64_t tree = (int64_t)blockIdx.x * 256 + threadIdx.x; tree < a.t; tree += (int64_t)gridDim.x * 256) {
    const int64_t p = a.punct[tree];
    const blk L = bxor(a.sums[(tree * a.h + a.d) * 2], a.z0);
    const blk R = bxor(a.sums[(tree * a.h + a.d) * 2 + 1], a.z1);
    const int c = a.choice[tree * a.h + a.d] ? 1 : 0;
    const int64_t sib = 2 * p + c, np = 2 * p + 1 - c;
    const blk sibv = bxor(a.csum[tree * a.h + a.d], c ? R : L);
    const int64_t base = tree << (a.d + 1);
    if (!a.last) {
      a.out[base + sib] = sibv;
      a.out[base + np] = bzero();
      a.punct[tree] = np;
      continue;
    }
    blk v;
    if (!a.ferret) {
      v = bxor(a.fin[tree], c ? L : R);
    } else {
      blk o = bxor(bxor(L, R), sibv);
      o.x &= ~1u;
      v = bxor(a.fin[tree], o);
    }
    if (base + sib < a.n_out) a.out[base + sib] = a.ferret ? make_uint4(sibv.x & ~1u, sibv.y, sibv.z, sibv.w) : sibv;
    if (base + np < a.n_out) a.out[base + np] = v;
  }
}

__global__ void ggm_bs_zero64(int64_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = 0;
}

unsigned small_grid(int64_t n) { return (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096); }

}  // namespace

size_t cuot_ggm_bs_work_bytes(int64_t t, int h) { return 2 * (size_t)(t << (h - 1)) * 16; }

size_t cuot_ggm_recver_bs_work_bytes(int64_t t, int h) {
  return cuot_ggm_bs_work_bytes(t, h) + (size_t)t * 16 + (size_t)t * h * 2 * 16 + (size_t)t * 8;
}

hipError_t cuot_launch_ggm_recver_bs(const uint32_t rk[88], int64_t t, int h, int ferret, int64_t n_out,
                                     const uint8_t* choice, const blk* csum, const blk* fin, blk* leaves, void* work,
                                     hipStream_t st) {
  GgmBsArgs a;
  bs_fold_sbox_constant(rk, a.rk[0]);
  bs_fold_sbox_constant(rk + 44, a.rk[1]);
  blk* buf[2] = {(blk*)work, (blk*)work + (t << (h - 1))};
  blk* zroots = buf[1] + (t << (h - 1));
  blk* sums = zroots + t;
  int64_t* punct = (int64_t*)(sums + t * h * 2);
  hipLaunchKernelGGL(ggm_bs_zero, dim3(small_grid(t + t * h * 2)), dim3(256), 0, st, zroots, t + t * h * 2);
  hipLaunchKernelGGL(ggm_bs_zero64, dim3(small_grid(t)), dim3(256), 0, st, punct, t);
  GgmBsFix f;
  f.sums = sums;
  f.punct = punct;
  f.choice = choice;
  f.csum = csum;
  f.fin = fin;
  f.t = t;
  f.h = h;
  f.ferret = ferret;
  uint8_t zero[16] = {0}, e0[16], e1[16];
  cuot_aes_encrypt_host(rk, zero, e0);
  cuot_aes_encrypt_host(rk + 44, zero, e1);
  memcpy(&f.z0, e0, 16);
  memcpy(&f.z1, e1, 16);
  a.sums = sums;
  a.h = h;
  for (int d = 0; d < h; d++) {
    a.d = d;
    a.np = t << d;
    a.par = d == 0 ? zroots : buf[d & 1];
    const bool last = d == h - 1;
    a.out = last ? leaves : buf[(d + 1) & 1];
    a.n_out = last ? n_out : 2 * a.np;
    a.clear_lsb = last && ferret;
    const int64_t lanes = (a.np + 31) / 32;
    cuot_rec_begin(st, "ggm_bs_level", 16.0 * 3.0 * (double)a.np);
    hipLaunchKernelGGL(ggm_bs_level, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, st, a);
    cuot_rec_end(st);
    f.d = d;
    f.last = last;
    f.out = a.out;
    f.n_out = last ? n_out : 2 * a.np;
    hipLaunchKernelGGL(ggm_bs_fix, dim3(small_grid(t)), dim3(256), 0, st, f);
  }
  return hipGetLastError();
}

hipError_t cuot_launch_ggm_sender_bs(const uint32_t rk[88], int64_t t, int h, int ferret, int64_t n_out,
                                     const blk* roots, blk* leaves, blk* sums, void* work, hipStream_t st) {
  GgmBsArgs a;
  bs_fold_sbox_constant(rk, a.rk[0]);
  bs_fold_sbox_constant(rk + 44, a.rk[1]);
  a.sums = sums;
  a.h = h;
  const int64_t nsums = t * h * 2;
  hipLaunchKernelGGL(ggm_bs_zero, dim3((unsigned)((nsums + 255) / 256 < 4096 ? (nsums + 255) / 256 : 4096)),
                     dim3(256), 0, st, sums, nsums);
  blk* buf[2] = {(blk*)work, (blk*)work + (t << (h - 1))};
  for (int d = 0; d < h; d++) {
    a.d = d;
    a.np = t << d;
    a.par = d == 0 ? roots : buf[d & 1];
    const bool last = d == h - 1;
    a.out = last ? leaves : buf[(d + 1) & 1];
    a.n_out = last ? n_out : 2 * a.np;
    a.clear_lsb = last && ferret;
    const int64_t lanes = (a.np + 31) / 32;
    cuot_rec_begin(st, "ggm_bs_level", 16.0 * 3.0 * (double)a.np);
    hipLaunchKernelGGL(ggm_bs_level, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, st, a);
    cuot_rec_end(st);
  }
  return hipGetLastError();
}

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
