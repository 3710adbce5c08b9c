// consist.hip -- emp-ot's malicious-security consistency check of the Ferret
// MPCOT (f2k variant) for both parties of a session on the device:
//   per tree i: seed_i = SHA-256(secret_sum_i)[0:16]            (spcot_sender.h:91-92)
//               chi_j = seed_i^(j+1)                          (uni_hash_coeff_gen)
//               V_i = sum_j chi_j leaf_s[j], W_i = sum_j chi_j leaf_r[j]
//                                                              (vector_inn_prdt_sum_red)
//               chi_alpha_i = chi at the punctured leaf        (spcot_recver.h:104)
//   then consistency_check_f2k (mpcot_reg.h:177-220) for both parties:
//     x'_j = bit j of XOR chi_alpha ^ LSB(pre_r[j]), j < 128
//     ALICE digest = SHA-256(XOR V ^ sum_j (pre_s[j] ^ x'_j Delta) x^j)
//     BOB digest   = SHA-256(XOR W ^ sum_j pre_r[j] x^j)      (GaloisFieldPacking)
// in GF(2^128) = GF(2)[x]/(x^128 + x^7 + x^2 + x + 1), bit i of a block (bit 0
// of byte 0 first) the coefficient of x^i.  Leaves at a global index >= n are
// not stored and are left out, and so is a chi_alpha whose leaf is not stored.
//
// One workgroup per tree.  With P = min(256, 2^h) lanes, lane p folds the
// leaves j = p + P m by Horner in S = seed^P and multiplies the result by
// seed^(p+1); the lanes' terms are then XOR-reduced.  Multiplications by the
// fixed S, seed and seed^16 use 4-bit window tables in LDS (32 lookups per
// product); the few general products use a bitwise carry-less multiply.
#include "kernels.h"

namespace {

// ---- SHA-256 of at most 55 bytes (FIPS 180-4, one padded block) --------------
__device__ const uint32_t kSha[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
    0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
    0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
    0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
    0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
    0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
    0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};

CUOT_DEV uint32_t ror(uint32_t x, int r) { return (x >> r) | (x << (32 - r)); }

// digest of the 16 bytes of b (memory order), as 8 big-endian words
CUOT_DEV void sha256_block16(const blk b, uint32_t H[8]) {
  uint32_t w[64];
  const uint32_t in[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] = __builtin_bswap32(in[i]);  // bytes 4i..4i+3, big-endian
  w[4] = 0x80000000u;
#pragma unroll
  for (int i = 5; i < 15; i++) w[i] = 0;
  w[15] = 128;  // message length in bits
#pragma unroll
  for (int i = 16; i < 64; i++) {
    const uint32_t s0 = ror(w[i - 15], 7) ^ ror(w[i - 15], 18) ^ (w[i - 15] >> 3);
    const uint32_t s1 = ror(w[i - 2], 17) ^ ror(w[i - 2], 19) ^ (w[i - 2] >> 10);
    w[i] = w[i - 16] + s0 + w[i - 7] + s1;
  }
  const uint32_t h0[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                          0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
  uint32_t a = h0[0], bb = h0[1], c = h0[2], d = h0[3], e = h0[4], f = h0[5], g = h0[6], h = h0[7];
#pragma unroll
  for (int i = 0; i < 64; i++) {
    const uint32_t t1 = h + (ror(e, 6) ^ ror(e, 11) ^ ror(e, 25)) + ((e & f) ^ (~e & g)) + kSha[i] + w[i];
    const uint32_t t2 = (ror(a, 2) ^ ror(a, 13) ^ ror(a, 22)) + ((a & bb) ^ (a & c) ^ (bb & c));
    h = g;
    g = f;
    f = e;
    e = d + t1;
    d = c;
    c = bb;
    bb = a;
    a = t1 + t2;
  }
  H[0] = h0[0] + a; H[1] = h0[1] + bb; H[2] = h0[2] + c; H[3] = h0[3] + d;
  H[4] = h0[4] + e; H[5] = h0[5] + f; H[6] = h0[6] + g; H[7] = h0[7] + h;
}

// the first 16 digest bytes as a block (emp-tool's digest[0])
CUOT_DEV blk hash_seed(const blk b) {
  uint32_t H[8];
  sha256_block16(b, H);
  return make_uint4(__builtin_bswap32(H[0]), __builtin_bswap32(H[1]), __builtin_bswap32(H[2]),
                    __builtin_bswap32(H[3]));
}

// ---- GF(2^128) ------------------------------------------------------------------
CUOT_DEV uint64_t lo64(blk a) { return (uint64_t)a.x | ((uint64_t)a.y << 32); }
CUOT_DEV uint64_t hi64(blk a) { return (uint64_t)a.z | ((uint64_t)a.w << 32); }
CUOT_DEV blk mk(uint64_t lo, uint64_t hi) {
  return make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
}
CUOT_DEV blk bx(blk a, blk b) { return make_uint4(a.x ^ b.x, a.y ^ b.y, a.z ^ b.z, a.w ^ b.w); }

CUOT_DEV void clmul64(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
  uint64_t l = 0, h = 0;
  for (int i = 0; i < 64; i++) {
    const uint64_t m = 0 - ((b >> i) & 1);
    l ^= (a << i) & m;
    h ^= (i ? (a >> (64 - i)) : 0) & m;
  }
  lo = l;
  hi = h;
}

// p[0..3] (256 bits, p[0] lowest) mod x^128 + x^7 + x^2 + x + 1
CUOT_DEV blk reduce256(uint64_t p0, uint64_t p1, uint64_t p2, uint64_t p3) {
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const uint64_t h0 = p2, h1 = p3;
    p0 ^= h0 ^ (h0 << 1) ^ (h0 << 2) ^ (h0 << 7);
    p1 ^= h1 ^ (h1 << 1) ^ (h1 << 2) ^ (h1 << 7) ^ (h0 >> 63) ^ (h0 >> 62) ^ (h0 >> 57);
    p2 = (h1 >> 63) ^ (h1 >> 62) ^ (h1 >> 57);
    p3 = 0;
  }
  return mk(p0, p1);
}

CUOT_DEV blk gf_mul(blk a, blk b) {
  uint64_t l, h, p0, p1, p2, p3;
  clmul64(lo64(a), lo64(b), p0, p1);
  clmul64(hi64(a), hi64(b), p2, p3);
  clmul64(lo64(a), hi64(b), l, h);
  p1 ^= l;
  p2 ^= h;
  clmul64(hi64(a), lo64(b), l, h);
  p1 ^= l;
  p2 ^= h;
  return reduce256(p0, p1, p2, p3);
}

// bits of x at the even positions of a 64-bit word
CUOT_DEV uint64_t spread32(uint32_t v) {
  uint64_t x = v;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

// a^2: in characteristic 2 the square of sum a_i x^i is sum a_i x^(2i)
CUOT_DEV blk gf_sqr(blk a) {
  return reduce256(spread32(a.x), spread32(a.y), spread32(a.z), spread32(a.w));
}

// a * x
CUOT_DEV blk gf_mul_x(blk a) {
  const uint64_t l = lo64(a), h = hi64(a);
  return reduce256(l << 1, (h << 1) | (l >> 63), h >> 63, 0);
}

// a * x^j, 0 <= j < 128
CUOT_DEV blk gf_mul_xj(blk a, int j) {
  const uint64_t l = lo64(a), h = hi64(a);
  const int s = j & 63;
  // (h:l) << s as three words
  const uint64_t w0 = l << s;
  const uint64_t w1 = (h << s) | (s ? l >> (64 - s) : 0);
  const uint64_t w2 = s ? h >> (64 - s) : 0;
  if (j < 64) return reduce256(w0, w1, w2, 0);
  return reduce256(0, w0, w1, w2);
}

// c * K through the 4-bit window table T[k][v] = K * (v x^(4k))
CUOT_DEV blk mul_tab(blk c, const uint4 (*T)[16]) {
  blk r = make_uint4(0u, 0u, 0u, 0u);
  const uint32_t wd[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
  for (int k = 0; k < 32; k++) r = bx(r, T[k][(wd[k >> 3] >> (4 * (k & 7))) & 15u]);
  return r;
}

struct ConsistArgs {
  const blk* ls;
  const blk* lr;
  int64_t n;
  int64_t t;
  int h;
  const blk* ssum;
  const uint8_t* bits;
  blk* V;
  blk* W;
  blk* chia;
};

__global__ __launch_bounds__(256) void consist_tree_kernel(ConsistArgs a) {
  __shared__ uint4 B[3][128];     // K x^i for the three table constants
  __shared__ uint4 T[3][32][16];  // window tables of S = seed^P, seed, seed^16
  __shared__ uint4 red[2][256];
  __shared__ uint4 cs[4];
  __shared__ uint4 PW[2][17];     // seed^b and seed^(16a), a, b = 0..16
  const int64_t tr = blockIdx.x;
  const int L = 1 << a.h;
  const int P = L < 256 ? L : 256;
  const int Mc = L / P;
  const int tid = threadIdx.x;
  if (tid < 2) {
    const blk s = hash_seed(a.ssum[tr]);
    if (tid == 0) {  // S = s^P
      blk S = s;
      for (int q = 1; q < P; q <<= 1) S = gf_sqr(S);
      cs[0] = S;
      cs[1] = s;
    } else if (tid == 1) {
      blk s16 = s;
      for (int q = 0; q < 4; q++) s16 = gf_sqr(s16);
      cs[2] = s16;
    }
  }
  __syncthreads();
  if (tid < 3) {
    blk c = cs[tid];
    for (int i = 0; i < 128; i++) {
      B[tid][i] = c;
      c = gf_mul_x(c);
    }
  }
  __syncthreads();
  for (int f = tid; f < 3 * 512; f += 256) {
    const int q = f >> 9, k = (f >> 4) & 31, v = f & 15;
    blk r = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (int b = 0; b < 4; b++)
      if ((v >> b) & 1) r = bx(r, B[q][4 * k + b]);
    T[q][k][v] = r;
  }
  __syncthreads();
  // powers for the lanes' chi offsets: seed^(tid+1) = seed^(16a) * seed^b
  // with a = (tid+1) >> 4 and b = (tid+1) & 15, one multiplication per lane
  if (tid < 2) {
    blk c = make_uint4(1u, 0u, 0u, 0u);
    for (int i = 0; i < 17; i++) {
      PW[tid][i] = c;
      c = mul_tab(c, T[1 + tid]);
    }
  }
  __syncthreads();
  if (tid == 0) {
    // chi_alpha = seed^(alpha+1) from the tables: alpha+1 = P*c + r, r < P
    int64_t alpha = 0;
    for (int d = 0; d < a.h; d++) alpha = (alpha << 1) + (a.bits[tr * a.h + d] ? 0 : 1);
    blk v = make_uint4(0u, 0u, 0u, 0u);
    if (tr * L + alpha < a.n) {
      const int64_t e1 = alpha + 1, c = e1 / P, r = e1 % P;
      v = gf_mul(PW[1][r >> 4], PW[0][r & 15]);
      for (int64_t q = 0; q < c; q++) v = mul_tab(v, T[0]);
    }
    cs[3] = v;
  }
  blk hs = make_uint4(0u, 0u, 0u, 0u), hr = hs;
  if (tid < P) {
    const int64_t base = tr * L + tid;
    for (int m = Mc - 1; m >= 0; m--) {
      hs = mul_tab(hs, T[0]);
      hr = mul_tab(hr, T[0]);
      const int64_t g = base + (int64_t)P * m;
      if (g < a.n) {
        hs = bx(hs, a.ls[g]);
        hr = bx(hr, a.lr[g]);
      }
    }
    const blk e = gf_mul(PW[1][(tid + 1) >> 4], PW[0][(tid + 1) & 15]);  // seed^(tid+1)
    hs = gf_mul(e, hs);
    hr = gf_mul(e, hr);
  }
  red[0][tid] = hs;
  red[1][tid] = hr;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      red[0][tid] = bx(red[0][tid], red[0][tid + w]);
      red[1][tid] = bx(red[1][tid], red[1][tid + w]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    a.V[tr] = red[0][0];
    a.W[tr] = red[1][0];
    a.chia[tr] = cs[3];
  }
}

// out: x' (128 bytes), ALICE digest (32), BOB digest (32), ok (int32)
__global__ __launch_bounds__(256) void consist_final_kernel(const blk* __restrict__ V, const blk* __restrict__ W,
                                                            const blk* __restrict__ chia, int64_t t,
                                                            const blk* __restrict__ pre_s,
                                                            const blk* __restrict__ pre_r,
                                                            const blk* __restrict__ delta, uint8_t* __restrict__ out) {
  __shared__ uint4 red[3][256];
  __shared__ uint4 pk[2][128];
  const int tid = threadIdx.x;
  blk r1a = make_uint4(0u, 0u, 0u, 0u), r1b = r1a, r2 = r1a;
  for (int64_t i = tid; i < t; i += 256) {
    r1a = bx(r1a, V[i]);
    r1b = bx(r1b, W[i]);
    r2 = bx(r2, chia[i]);
  }
  red[0][tid] = r1a;
  red[1][tid] = r1b;
  red[2][tid] = r2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
      for (int q = 0; q < 3; q++) red[q][tid] = bx(red[q][tid], red[q][tid + w]);
    __syncthreads();
  }
  if (tid < 128) {
    const blk c = red[2][0];
    const uint32_t wd[4] = {c.x, c.y, c.z, c.w};
    const uint32_t xp = ((wd[tid >> 5] >> (tid & 31)) & 1u) ^ (pre_r[tid].x & 1u);
    out[tid] = (uint8_t)xp;
    blk s = pre_s[tid];
    if (xp) s = bx(s, *delta);
    pk[0][tid] = gf_mul_xj(s, tid);
    pk[1][tid] = gf_mul_xj(pre_r[tid], tid);
  }
  __syncthreads();
  for (int w = 64; w > 0; w >>= 1) {
    if (tid < w) {
      pk[0][tid] = bx(pk[0][tid], pk[0][tid + w]);
      pk[1][tid] = bx(pk[1][tid], pk[1][tid + w]);
    }
    __syncthreads();
  }
  if (tid < 2) {
    uint32_t H[8];
    sha256_block16(bx(red[tid][0], pk[tid][0]), H);
    uint8_t* d = out + 128 + 32 * tid;
    for (int i = 0; i < 8; i++) {
      d[4 * i] = (uint8_t)(H[i] >> 24);
      d[4 * i + 1] = (uint8_t)(H[i] >> 16);
      d[4 * i + 2] = (uint8_t)(H[i] >> 8);
      d[4 * i + 3] = (uint8_t)H[i];
    }
  }
  __syncthreads();
  if (tid == 0) {
    int ok = 1;
    for (int i = 0; i < 32; i++) ok &= out[128 + i] == out[160 + i];
    *(int32_t*)(out + 192) = ok;
  }
}

}  // namespace

hipError_t cuot_launch_consist(const blk* leaves_s, const blk* leaves_r, int64_t n, int64_t t, int h,
                               const blk* ssum, const uint8_t* bits, const blk* pre_s, const blk* pre_r,
                               const blk* delta, blk* vwc, uint8_t* out, hipStream_t st) {
  if (t < 1 || h < 1 || h > 30 || t > 0x7fffffff) return hipErrorInvalidValue;
  ConsistArgs a;
  a.ls = leaves_s;
  a.lr = leaves_r;
  a.n = n;
  a.t = t;
  a.h = h;
  a.ssum = ssum;
  a.bits = bits;
  a.V = vwc;
  a.W = vwc + t;
  a.chia = vwc + 2 * t;
  // algorithmic bytes: both parties' stored leaves read once
  cuot_rec_begin(st, "consist_tree_kernel", 32.0 * (double)(n < (t << h) ? n : (t << h)));
  hipLaunchKernelGGL(consist_tree_kernel, dim3((unsigned)t), dim3(256), 0, st, a);
  cuot_rec_end(st);
  hipLaunchKernelGGL(consist_final_kernel, dim3(1), dim3(256), 0, st, a.V, a.W, a.chia, t, pre_s, pre_r, delta, out);
  return hipGetLastError();
}
