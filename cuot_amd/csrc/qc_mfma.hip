// qc_mfma.hip -- the quasi-cyclic encoder of small codes on the matrix cores.
//
//     y_i = x_i ^ XOR_{k < n} a_{(i-k) mod n} x_{n+k},   i < n
//
// (the same map as qc.hip; the reference's float cuFFT version is
// silent/quasi_cyclic.cu:118-128).  For all 128 bit-rows r at once this is the
// dense 0/1 matrix product
//     P[i][r] = sum_k C[i][k] E[k][r],   C[i][k] = a_{(i-k) mod n},   E[k][r] = bit r of x_{n+k},
// and bit r of y_i is bit r of x_i XOR (P[i][r] mod 2).  Every sum is below
// 2^17, so v_mfma_i32_32x32x32_i8 with int32 accumulation yields each parity
// exactly.  The product costs n^2 * 128 multiply-adds, so it beats the
// additive FFT only for small n; the crossover is an execution knob.
//
// Operands.
//   B: E is unpacked once per encode into bytes, E8[k/16][r][k%16]: the B
//      operand of one lane (column r, 16 consecutive k) is one 16-byte word.
//      The four waves of a workgroup share each k-step's 4 KB of E8 through
//      LDS (three buffers, one barrier per step).  Only ceil(n/32)*32 rows of
//      k are stored (zero beyond n), which is at most the FFT's 64 planes x
//      2^m words, so the encode scratch is shared.
//   A: the circulant never exists in memory.  A workgroup holds the bytes
//      a_{(-j) mod n} for the window of j = k - i it touches in LDS; lane
//      (row i, k-group g) reads its 16 bytes a_{(i-k) mod n}, k = k0+16g .. +15,
//      as 16 consecutive window bytes (five aligned dwords and four byte
//      alignments).  Row tile rt of a wave at k-step s needs exactly row tile
//      0's operand of step s - rt, so each step reads one new A operand per
//      wave and moves the others along.
// The k range is split over gridDim.y slices so that small n still fills the
// chip; the slices XOR their parities into y (set to x by the unpack kernel)
// with 64-bit atomics, and a single slice writes y = x ^ parity directly.
#include "kernels.h"

namespace {

typedef int mf_v4 __attribute__((ext_vector_type(4)));
typedef int mf_v16 __attribute__((ext_vector_type(16)));

constexpr int MF_RT = 2;                      // 32-row tiles per wave
constexpr int MF_WAVES = 4;                   // waves per workgroup
constexpr int MF_RW = 32 * MF_RT * MF_WAVES;  // rows per workgroup
constexpr int64_t MF_NMIN = 17;               // below: the E8 rows exceed the FFT scratch (m = 1)
constexpr int64_t MF_NMAX = 1 << 16;          // bits of a are kept after the transform up to here
constexpr int64_t MF_TARGET_WGS = 1024;       // rows x slices aimed for (four per CU)
constexpr int64_t MF_KLEN_MIN = 256, MF_KLEN_MAX = 16384;
constexpr int MF_WIN_BYTES = MF_KLEN_MAX + MF_RW + 16;
constexpr int64_t MF_DEFAULT_MAX = 49152;  // built-in crossover (n at or below: matrix cores), measured

struct MfPlan {
  int64_t gx;    // row workgroups
  int64_t S;     // k slices
  int64_t klen;  // k per slice (multiple of 32)
  int64_t k32;   // ceil(n/32)*32: stored rows of E8
};

__host__ __device__ inline MfPlan mf_plan(int64_t n) {
  MfPlan p;
  p.gx = (n + MF_RW - 1) / MF_RW;
  p.k32 = (n + 31) / 32 * 32;
  int64_t S = MF_TARGET_WGS / p.gx;
  const int64_t s_hi = (n + MF_KLEN_MIN - 1) / MF_KLEN_MIN;  // slices of at least MF_KLEN_MIN
  const int64_t s_lo = (n + MF_KLEN_MAX - 1) / MF_KLEN_MAX;  // window fits the LDS reserve
  if (S > s_hi) S = s_hi;
  if (S < s_lo) S = s_lo;
  if (S < 1) S = 1;
  p.klen = ((p.k32 + S - 1) / S + 31) / 32 * 32;
  p.S = (p.k32 + p.klen - 1) / p.klen;
  return p;
}

// Thread (kb, r): bytes t = 0..15 of E8[kb][r] = bit r of x_{n + 16 kb + t}
// (0 beyond n).  With more than one slice the same grid copies y = x first.
__global__ __launch_bounds__(256) void qc_mf_unpack(const QcIo* io, QcIo io1, uint4* e8_base, int64_t e8_stride) {
  const QcIo p = io ? io[blockIdx.y] : io1;
  const int64_t n = p.n;
  const MfPlan pl = mf_plan(n);
  const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (pl.S > 1 && gid < n) p.y[gid] = p.x[gid];
  const int64_t kb = gid >> 7;
  if (kb >= pl.k32 / 16) return;
  const int r = (int)(gid & 127);
  const uint32_t* xw = (const uint32_t*)(p.x + n) + (r >> 5);
  uint32_t d[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int t = 0; t < 16; t++) {
    const int64_t k = 16 * kb + t;
    const uint32_t bit = k < n ? (xw[4 * k] >> (r & 31)) & 1u : 0u;
    d[t >> 2] |= bit << (8 * (t & 3));
  }
  e8_base[(int64_t)blockIdx.y * e8_stride + gid] = make_uint4(d[0], d[1], d[2], d[3]);
}

// A operand from the window: bytes win[o .. o+15], o = 4 * (p - win) + sh.
__device__ __forceinline__ mf_v4 mf_a(const uint32_t* p, uint32_t sh) {
  const uint32_t d0 = p[0], d1 = p[1], d2 = p[2], d3 = p[3], d4 = p[4];
  mf_v4 a;
  a.x = (int)__builtin_amdgcn_alignbyte(d1, d0, sh);
  a.y = (int)__builtin_amdgcn_alignbyte(d2, d1, sh);
  a.z = (int)__builtin_amdgcn_alignbyte(d3, d2, sh);
  a.w = (int)__builtin_amdgcn_alignbyte(d4, d3, sh);
  return a;
}

// grid (row workgroups, slices, instances); 4 waves of 2 x 32 rows each.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void qc_mf_encode(
    const QcIo* io, QcIo io1, const uint32_t* bits1, const uint32_t* const* ahats, int64_t bits_off,
    const uint4* e8_base, int64_t e8_stride) {
  __shared__ uint32_t win[MF_WIN_BYTES / 4];
  __shared__ uint4 ow[MF_WAVES][32 * MF_RT];  // parity words of each wave's rows
  __shared__ mf_v4 bsh[3][256];               // B tiles of three k-steps
  const QcIo p = io ? io[blockIdx.z] : io1;
  const int64_t n = p.n;
  const MfPlan pl = mf_plan(n);
  const int64_t i0 = (int64_t)blockIdx.x * MF_RW;
  const int64_t kl = (int64_t)blockIdx.y * pl.klen;
  if (i0 >= n || kl >= pl.k32) return;  // uniform over the workgroup
  const int64_t kh = kl + pl.klen < pl.k32 ? kl + pl.klen : pl.k32;
  const uint32_t* bits = io ? ahats[blockIdx.z] + bits_off : bits1;
  const mf_v4* e8 = (const mf_v4*)(e8_base + (int64_t)blockIdx.z * e8_stride);

  // window byte w holds a_{(-(jb + w)) mod n}, j = k - i over the workgroup's
  // rows and slice (4 bytes of slack in front for the aligned dword reads)
  const int64_t jb = kl - i0 - MF_RW + 1 - 4;
  const int wlen = (int)(kh - kl) + MF_RW + 16;
  for (int d = threadIdx.x; d < wlen / 4; d += 256) {
    int64_t s = (-(jb + 4 * d)) % n;
    if (s < 0) s += n;
    uint32_t v = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      int64_t q = s - b;
      if (q < 0) q += n;
      v |= ((bits[q >> 5] >> (q & 31)) & 1u) << (8 * b);
    }
    win[d] = v;
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 5, li = lane & 31;
  const int64_t i0w = i0 + (int64_t)wave * 32 * MF_RT;
  // window offset of row tile 0's A operand at step 0 (>= 4 + 32 (MF_RT-1))
  const int o = (int)(kl + 16 * g - (i0w + li) - jb);
  const uint32_t sh = (uint32_t)(o & 3);
  const uint32_t* wp = win + (o >> 2);

  mf_v4 A[MF_RT];
#pragma unroll
  for (int rt = 1; rt < MF_RT; rt++) A[rt] = mf_a(wp - 8 * rt, sh);
  mf_v16 acc[MF_RT][4];
#pragma unroll
  for (int rt = 0; rt < MF_RT; rt++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[rt][c] = (mf_v16){0};

  // the step count is wave-uniform (64-bit divisions above ran on the VALU)
  const int nsteps = __builtin_amdgcn_readfirstlane((int)((kh - kl) / 32));
  // B: the workgroup's four waves share each k-step's 4 KB tile of E8
  // (E8[kb..kb+1][0..127]) through three LDS buffers: during step s a wave
  // multiplies with B(s) in registers, reads B(s+1) from LDS and stores the
  // tile of step s+2, which thread t carries as word t in a register; buffer
  // (s+2) % 3 was last read in step s-1, before that step's barrier
  const mf_v4* tp = e8 + (kl / 16) * 128 + threadIdx.x;
  auto tile = [&](int st) { return tp[(int64_t)(st < nsteps ? st : nsteps - 1) * 256]; };
  bsh[0][threadIdx.x] = tile(0);
  bsh[1][threadIdx.x] = tile(1);
  mf_v4 r = tile(2);
  __syncthreads();
  mf_v4 B[4], Bn[4];
#pragma unroll
  for (int c = 0; c < 4; c++) B[c] = bsh[0][g * 128 + 32 * c + li];
  mf_v4 An = mf_a(wp, sh);
  int rb = 1, wb = 2;  // buffers of steps s + 1 and s + 2
  for (int s = 0; s < nsteps; s++) {
#pragma unroll
    for (int c = 0; c < 4; c++) Bn[c] = bsh[rb][g * 128 + 32 * c + li];
    A[0] = An;
    An = mf_a(wp + 8 * (s + 1 < nsteps ? s + 1 : s), sh);
#pragma unroll
    for (int rt = 0; rt < MF_RT; rt++)
#pragma unroll
      for (int c = 0; c < 4; c++) acc[rt][c] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[rt], B[c], acc[rt][c], 0, 0, 0);
#pragma unroll
    for (int rt = MF_RT - 1; rt >= 1; rt--) A[rt] = A[rt - 1];
    bsh[wb][threadIdx.x] = r;
    r = tile(s + 3);
#pragma unroll
    for (int c = 0; c < 4; c++) B[c] = Bn[c];
    rb = rb == 2 ? 0 : rb + 1;
    wb = wb == 2 ? 0 : wb + 1;
    __syncthreads();
  }

  // parities to rows: accumulator v of lane l holds row 8(v/4) + 4(l/32) + v%4
  // of its tile, column l%32, so one ballot per (tile, v) gives word c of two
  // rows; lane 0 parks the words in the wave's LDS rows, and each lane then
  // takes the row it writes out
#pragma unroll
  for (int rt = 0; rt < MF_RT; rt++) {
#pragma unroll
    for (int v = 0; v < 16; v++) {
      uint32_t lo[4], hi[4];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const unsigned long long b = __ballot(acc[rt][c][v] & 1);
        lo[c] = (uint32_t)b;
        hi[c] = (uint32_t)(b >> 32);
      }
      const int rl = 32 * rt + 8 * (v >> 2) + (v & 3);
      if (lane == 0) {
        ow[wave][rl] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
        ow[wave][rl + 4] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
      }
    }
    __builtin_amdgcn_sched_barrier(0);  // one row tile's accumulators live at a time
  }
  __syncthreads();
  constexpr int SL = MF_RT / 2;
  uint32_t wd[SL][4];
#pragma unroll
  for (int sl = 0; sl < SL; sl++) {
    const uint4 w = ow[wave][64 * sl + lane];
    wd[sl][0] = w.x;
    wd[sl][1] = w.y;
    wd[sl][2] = w.z;
    wd[sl][3] = w.w;
  }
#pragma unroll
  for (int sl = 0; sl < SL; sl++) {
    const int64_t i = i0w + 64 * sl + lane;
    if (i >= n) continue;
    if (pl.S == 1) {
      const blk xi = p.x[i];
      p.y[i] = make_uint4(xi.x ^ wd[sl][0], xi.y ^ wd[sl][1], xi.z ^ wd[sl][2], xi.w ^ wd[sl][3]);
    } else {
      unsigned long long* yq = (unsigned long long*)(p.y + i);
      atomicXor(yq, (unsigned long long)wd[sl][0] | ((unsigned long long)wd[sl][1] << 32));
      atomicXor(yq + 1, (unsigned long long)wd[sl][2] | ((unsigned long long)wd[sl][3] << 32));
    }
  }
}

// bits of a (n of them, from ceil(n/8) bytes) to the words after the transform
__global__ __launch_bounds__(256) void qc_mf_keep_bits(const QcPrep* pr, QcPrep pr1, int64_t off) {
  const QcPrep p = pr ? pr[blockIdx.y] : pr1;
  if (p.n < MF_NMIN || p.n > MF_NMAX) return;
  const int64_t nbytes = (p.n + 7) / 8, words = (p.n + 31) / 32;
  uint8_t* dst = (uint8_t*)(p.ahat + off);
  for (int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x; b < 4 * words; b += (int64_t)gridDim.x * 256)
    dst[b] = b < nbytes ? p.a_bits[b] : (uint8_t)0;
}

}  // namespace

bool cuot_qc_mfma_eligible(int64_t n) { return n >= MF_NMIN && n <= MF_NMAX; }

int64_t cuot_qc_ahat_words(int64_t n) {
  const int m = cuot_qc_log_len(n);
  return ((int64_t)1 << m) + (cuot_qc_mfma_eligible(n) ? (n + 31) / 32 : 0);
}

int64_t cuot_qc_mfma_limit() {
  const int k = cuot_knobs().qc_mfma_max;
  const int64_t lim = k == 0 ? MF_DEFAULT_MAX : (k < 0 ? 0 : k);
  return lim < MF_NMAX ? lim : MF_NMAX;
}

bool cuot_qc_mfma_selected(int64_t n) { return cuot_qc_mfma_eligible(n) && n <= cuot_qc_mfma_limit(); }

hipError_t cuot_launch_qc_keep_bits(const QcPrep* pr, QcPrep pr1, int B, int m, int64_t nmax, hipStream_t st) {
  const int64_t bytes = 4 * ((nmax + 31) / 32);
  hipLaunchKernelGGL(qc_mf_keep_bits, dim3((unsigned)((bytes + 255) / 256), (unsigned)B), dim3(256), 0, st, pr, pr1,
                     (int64_t)1 << m);
  return hipGetLastError();
}

hipError_t cuot_launch_qc_mfma(const QcIo* io, QcIo io1, const uint32_t* bits1, const uint32_t* const* ahats, int B,
                               int m, int64_t nmax, void* scratch, hipStream_t st) {
  if (!cuot_qc_mfma_eligible(nmax) || B < 1) return hipErrorInvalidValue;
  // per-instance E8 region: the largest n of the transform class, 2^(m+3)
  const int64_t e8_stride = ((int64_t)1 << (m + 3)) * 8;
  if (B > 1 && (size_t)B * (size_t)e8_stride * 16 > cuot_qc_batch_scratch_bytes(m, B)) return hipErrorInvalidValue;
  const MfPlan pl = mf_plan(nmax);
  // a class's smaller members may use more slices than its largest
  const MfPlan lo = mf_plan(B > 1 ? ((int64_t)1 << (m + 2)) + 1 : nmax);
  const int64_t S = lo.S > pl.S ? lo.S : pl.S;
  const int64_t uthreads = pl.k32 / 16 * 128;
  cuot_rec_begin(st, "qc_mf_unpack", 16.0 * (double)nmax + 128.0 * (double)pl.k32 * B);
  hipLaunchKernelGGL(qc_mf_unpack, dim3((unsigned)((uthreads + 255) / 256), (unsigned)B), dim3(256), 0, st, io, io1,
                     (uint4*)scratch, e8_stride);
  cuot_rec_end(st);
  // algorithmic bytes: E8 read once, y read and written
  cuot_rec_begin(st, "qc_mf_encode", (128.0 * (double)pl.k32 + 32.0 * (double)nmax) * B);
  hipLaunchKernelGGL(qc_mf_encode, dim3((unsigned)pl.gx, (unsigned)S, (unsigned)B), dim3(256), 0, st, io, io1, bits1,
                     ahats, (int64_t)1 << m, (const uint4*)scratch, e8_stride);
  cuot_rec_end(st);
  return hipGetLastError();
}
