// qc.hip -- exact quasi-cyclic encoder over GF(2) for all 128 bit-rows at once:
//     y_i = x_i ^ XOR_{j : a_j = 1} x_{n + ((i - j) mod n)},   i < n
// (the workload's "2n noisy blocks -> n through a random circulant"; the
// reference computes a float32 cuFFT convolution instead,
// silent/quasi_cyclic.cu:11-49,118-128, whose parity is inexact at this size).
//
// Algorithm: binary polynomial product a(x) * e1(x) by an additive FFT over
// GF(2^32) (Lin-Chung-Han novel-basis FFT on a Cantor basis), then the fold
// mod x^n - 1.  Bits are packed 16 per field element (chunk polynomials of
// degree < 16 multiply without reduction, so chunk products are exact), and the
// 128 rows are bitsliced: block x_{n+16c+b} IS bit-plane b of chunk c for all
// 128 rows, so no bit transpose is needed.
//
// Data layout (device scratch): X[plane][elem] of 16-byte words (128 rows),
// 32 planes x M elements, M = 2^m >= 2*ceil(n/16) - 1, plus a ping-pong buffer
// of G planes for the out-of-place conversion levels.
//
// Encode (per party):
//   pack (16 planes, lower half)
//   basis conversion, levels m-2..B: one out-of-place pass per level, run per
//     group of G planes so the group stays in the last-level cache;
//     levels B-1..0 in one LDS pass per contiguous tile of one plane
//   FFT: LDS radix passes of up to 8 layers (the first pass also performs the
//     trivial top layer and the duplication of the lower half); the lowest
//     pass does forward layers, the pointwise product by A^, and the inverse
//     layers in one LDS round trip
//   inverse FFT radix passes, inverse conversion (LDS pass, global levels),
//     and the top inverse level fused into the overlap-add / fold / xor x_0.
#include <string.h>

#include "conv_fast.h"
#include "gf32.h"
#include "kernels.h"
#include "transpose32.h"

namespace {

constexpr int CONV_B = 12;     // conversion levels done in LDS (tile 2^B elements of one plane)
constexpr int CONV_G = 4;      // planes per conversion group (ping-pong buffer size)
constexpr int FFT_TC = 128;    // elements per FFT LDS tile (2^R * C): two tiles fit one CU's LDS
constexpr int FFT_LOGTC = 7;
constexpr int FFT_STRIDE = 132;  // LDS words per element (32 planes x 4 words + pad)
#ifndef FFT_SWZ
#define FFT_SWZ 1
#endif
// LDS slot of tile element e.  With the 132-word stride an element's word w of
// plane p sits in bank group e mod 8; the butterflies of the lowest tile layers
// (s = 0..2) touch elements whose indices share their residues mod 8 four at a
// time.  Reflecting the residue of every odd group of eight (e ^ 7) spreads
// each such access over all eight bank groups.
CUOT_DEV int fslot(int e) { return FFT_SWZ ? e ^ (((e >> 3) & 1) * 7) : e; }
constexpr int FFT_THREADS = 256;  // one (butterfly, word) task per lane per layer
constexpr int FFT_LD = 16;  // 16-byte loads in flight per lane (tile staging)

constexpr int FFT_WGS = 512;  // FFT pass grid: two resident WGs on each of the 256 CUs
constexpr int CR_LD = 8;    // same for the bit-row conversion tiles
// the row-layout conversion handles strings of 2^5 .. 2^20 coefficients (128 KB, padded 144 KB, of LDS)
constexpr int ROW_MP_MIN = 5, ROW_MP_MAX = 20;

struct Beta {
  uint32_t v[32];
};

// Cantor basis beta_0..beta_31 (uploaded once per process by configure_lds)
__constant__ uint32_t c_beta[32];

// lower terms of s_k: e = 2^j for j a proper submask of k (at most 15 for k < 31)
CUOT_DEV int lower_terms(int k, uint32_t* e) {
  int c = 0;
  for (int j = 0; j < k; j++)
    if ((j & ~k) == 0) e[c++] = 1u << j;
  return c;
}

// C(g) = sum_b g_b * beta_{b+1}
CUOT_DEV uint32_t chunk_const(const uint32_t* beta, uint64_t g) {
  uint32_t c = 0;
  for (int b = 0; g; b++, g >>= 1)
    if (g & 1) c ^= beta[b + 1];
  return c;
}

// C(g) from three 256-entry tables of the low 24 bits of g (tab[256*q + v] =
// C(v << 8q)); g < 2^24 for every transform length handled here (m <= 24)
constexpr int CTAB_WORDS = 3 * 256;

CUOT_DEV void ctab_fill(uint32_t* tab, int tid, int nth) {
  for (int i = tid; i < CTAB_WORDS; i += nth) {
    const int q = i >> 8, v = i & 255;
    uint32_t c = 0;
    for (int b = 0; b < 8; b++)
      if ((v >> b) & 1) c ^= c_beta[8 * q + b + 1];
    tab[i] = c;
  }
}

CUOT_DEV uint32_t ctab_const(const uint32_t* tab, uint32_t g) {
  return tab[g & 255] ^ tab[256 + ((g >> 8) & 255)] ^ tab[512 + ((g >> 16) & 255)];
}

CUOT_DEV uint4 x4(uint4 a, uint4 b) { return make_uint4(a.x ^ b.x, a.y ^ b.y, a.z ^ b.z, a.w ^ b.w); }
CUOT_DEV uint4 z4() { return make_uint4(0, 0, 0, 0); }

// ---- pack -------------------------------------------------------------------
// X[b][c] = x1[16c + b] (zero beyond n), c < M/2, planes 0..15.  blockIdx.y
// selects the instance of a batch (X stride XS blocks, io table or io1).
__global__ void qc_pack_x(const QcIo* __restrict__ io, QcIo io1, uint4* __restrict__ X, int64_t M, int64_t XS) {
  const QcIo q = io ? io[blockIdx.y] : io1;
  const uint4* x1 = q.x + q.n;
  const int64_t n = q.n;
  X += (int64_t)blockIdx.y * XS;
  const int64_t half = M >> 1;
  const int64_t total = half * 16;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t >> 4;
    const int b = (int)(t & 15);
    const int64_t src = 16 * c + b;
    X[(int64_t)b * M + c] = src < n ? x1[src] : z4();
  }
}

// ---- basis conversion -------------------------------------------------------
// Division by s_k on chunks of 2^(k+1) (forward, monomial -> novel basis):
//   q_i = hi_i ^ XOR_{e in T_k, e > i} hi_{i + half - e}
//   r_j = lo_j ^ XOR_{e in T_k, e <= j} q_{j - e}
// Multiplication back (inverse): lo_j ^= XOR_{e <= j} hi_{j-e} (original hi),
// then hi_i ^= XOR_{e > i} hi_{i + half - e}.  Only positions i < e_max of hi
// change, and they read positions >= e_max (half >= 2 e_max), so each step is
// a data-parallel pass.

// one level, out-of-place, planes [0, P) of a polynomial of Mp elements
template <bool INV>
__global__ void qc_conv_level(const uint4* __restrict__ in, uint4* __restrict__ out, int64_t M, int P, int64_t Mp,
                              int k) {
  uint32_t e[16];
  const int ne = lower_terms(k, e);
  const uint32_t emax = ne ? e[ne - 1] : 0;
  const int64_t half = (int64_t)1 << k;
  const int64_t total = (int64_t)P * Mp;
  const int lgMp = 63 - __builtin_clzll((unsigned long long)Mp);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)(t >> lgMp);
    const int64_t i = t & (Mp - 1);
    const int64_t pos = i & (2 * half - 1);
    const uint4* lo = in + (int64_t)p * M + (i - pos);
    const uint4* hi = lo + half;
    uint4 v;
    if (pos < half) {
      const int64_t j = pos;
      v = lo[j];
      for (int q = 0; q < ne; q++) {
        if ((int64_t)e[q] > j) break;
        const int64_t s = j - e[q];
        uint4 qs = hi[s];
        if (!INV && s < (int64_t)emax) {
          for (int r = 0; r < ne; r++)
            if ((int64_t)e[r] > s) qs = x4(qs, hi[s + half - e[r]]);
        }
        v = x4(v, qs);
      }
    } else {
      const int64_t s = pos - half;
      v = hi[s];
      if (s < (int64_t)emax) {
        for (int r = 0; r < ne; r++)
          if ((int64_t)e[r] > s) v = x4(v, hi[s + half - e[r]]);
      }
    }
    out[(int64_t)p * M + i] = v;
  }
}

// levels B-1..0 (forward) or 0..B-1 (inverse) in LDS on contiguous tiles of
// 2^B elements of one plane; reads src, writes dst (may be equal).
template <bool INV>
__global__ __launch_bounds__(512) void qc_conv_lds(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t M,
                                                   int P, int64_t Mp, int B) {
  extern __shared__ uint4 sm[];
  const int64_t tile_n = (int64_t)1 << B;
  const int64_t tiles_pp = Mp >> B;
  const int64_t ntiles = (int64_t)P * tiles_pp;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int p = (int)(tile / tiles_pp);
    const int64_t off = (int64_t)p * M + (tile % tiles_pp) * tile_n;
    for (int64_t i = threadIdx.x; i < tile_n; i += blockDim.x) sm[i] = src[off + i];
    __syncthreads();
    for (int step = 0; step < B; step++) {
      const int k = INV ? step : B - 1 - step;
      uint32_t e[16];
      const int ne = lower_terms(k, e);
      if (ne == 0) continue;
      const uint32_t emax = e[ne - 1];
      const int half = 1 << k;
      for (int phase = 0; phase < 2; phase++) {
        const bool do_hi = (phase == 0) != INV;  // forward: hi_fix first
        if (do_hi) {
          const int64_t cnt = (tile_n >> (k + 1)) * emax;
          for (int64_t x = threadIdx.x; x < cnt; x += blockDim.x) {
            const int64_t g = x / emax, i = x % emax;
            uint4* h = sm + g * 2 * half + half;
            uint4 v = h[i];
            for (int r = 0; r < ne; r++)
              if ((int64_t)e[r] > i) v = x4(v, h[i + half - e[r]]);
            h[i] = v;
          }
        } else {
          const int64_t cnt = tile_n >> 1;
          for (int64_t x = threadIdx.x; x < cnt; x += blockDim.x) {
            const int64_t g = x >> k, j = x & (half - 1);
            uint4* l = sm + g * 2 * half;
            const uint4* h = l + half;
            uint4 v = l[j];
            for (int r = 0; r < ne; r++) {
              if ((int64_t)e[r] > j) break;
              v = x4(v, h[j - e[r]]);
            }
            l[j] = v;
          }
        }
        __syncthreads();
      }
    }
    for (int64_t i = threadIdx.x; i < tile_n; i += blockDim.x) dst[off + i] = sm[i];
    __syncthreads();
  }
}

// Row-string conversion: after a 128-row bit transpose of each plane, every
// (plane, row) is a string of 2^mp coefficients (the bits of one row over
// the element index); tiles are consecutive strings of tile_words words.
struct BarrierSync {
  __device__ void operator()() const { __syncthreads(); }
};

// fold (inverse mode, odd tiles = the upper half q of a row): after the
// conversion, q is replaced by q ^ XOR_{e in E} rot(q, e), E = the lower terms
// of s_mp, rotations cyclic over the 2^mp-bit string.  The lower half r then
// satisfies r ^ T(q) = (F mod X^(2^mp) + 1) for F = r + s_mp q, the product
// folded at half its length (s_mp = X^(2^mp) + sum_E X^e).
// PAD (strings of >= 2^10 coefficients): the string sits in LDS with four pad
// words after every 32, and levels 9..0 run per 1024-bit block in registers.
// MULTI (PAD, strings shorter than 2^20): a workgroup converts S = 1024 /
// (blocks per string) strings side by side, one lane per block, so that short
// strings still occupy every lane (the same LDS footprint for every length:
// 1024 padded blocks).  Its per-lane LDS base costs registers, so there the
// in-word levels (the first inverse and the last forward step) ride on the
// copies between global memory and LDS (INW); a single string keeps them in
// the register blocks, which measured faster.
CUOT_DEV uint4 inword4(uint4 v, bool inv) {
  v.x = cbt_word_all(v.x, inv);
  v.y = cbt_word_all(v.y, inv);
  v.z = cbt_word_all(v.z, inv);
  v.w = cbt_word_all(v.w, inv);
  return v;
}

template <bool PAD, bool INW>
__device__ __forceinline__ void rows_load(const uint32_t* g, uint32_t* sw, int64_t tile_words, int lt, int tps) {
  const int n4 = (int)(tile_words >> 2);
  if ((tile_words & 3) == 0) {
    // 16-byte accesses, CR_LD in flight per lane
    const uint4* g4 = (const uint4*)g;
    for (int f0 = lt; f0 < n4; f0 += tps * CR_LD) {
      uint4 v[CR_LD];
#pragma unroll
      for (int u = 0; u < CR_LD; u++) {
        const int f = f0 + u * tps;
        v[u] = z4();
        if (f < n4) v[u] = g4[f];
      }
#pragma unroll
      for (int u = 0; u < CR_LD; u++) {
        const int f = f0 + u * tps;
        if (f < n4) *(uint4*)(sw + cbp_at<PAD>(4 * (int64_t)f)) = INW ? inword4(v[u], true) : v[u];
      }
    }
  } else {
    for (int64_t i = lt; i < tile_words; i += tps) sw[i] = g[i];
  }
}

// fold (inverse, odd tiles): the terms e = 2^j (j a proper submask of mp) are
// enumerated in the loop; tile_words is a power of two and a multiple of 4
template <bool PAD, bool INW>
__device__ __forceinline__ void rows_store(uint32_t* g, const uint32_t* sw, int64_t tile_words, int mp, bool fold_q,
                                           int lt, int tps) {
  const int n4 = (int)(tile_words >> 2);
  if (fold_q) {
    const int64_t wm = tile_words - 1;
    uint4* g4 = (uint4*)g;
    auto word = [&](int64_t i) { return sw[cbp_at<PAD>(i & wm)]; };
    auto quad = [&](int64_t i) { return *(const uint4*)(sw + cbp_at<PAD>(i & wm)); };
    for (int f = lt; f < n4; f += tps) {
      const int64_t w = 4 * (int64_t)f;
      // the terms below one word (e < 32) all shift the quad at w with the
      // word before it: both are read once
      const uint4 q0 = quad(w);
      const uint32_t pv0 = word(w - 1);
      uint4 acc = q0;
      for (int j = 0; j < mp; j++) {
        if (j & ~mp) continue;
        const int64_t eo = ((int64_t)1 << j) >> 5;
        const int eb = j < 5 ? 1 << j : 0;
        uint4 c4;
        if (eo == 0) {
          c4 = q0;
        } else if ((eo & 3) == 0) {
          c4 = quad(w - eo);
        } else {
          c4 = make_uint4(word(w - eo), word(w - eo + 1), word(w - eo + 2), word(w - eo + 3));
        }
        if (eb) {
          const uint32_t pv = pv0;  // eb != 0 only for j < 5, where eo = 0
          acc.x ^= (c4.x << eb) | (pv >> (32 - eb));
          acc.y ^= (c4.y << eb) | (c4.x >> (32 - eb));
          acc.z ^= (c4.z << eb) | (c4.y >> (32 - eb));
          acc.w ^= (c4.w << eb) | (c4.z >> (32 - eb));
        } else {
          acc = x4(acc, c4);
        }
      }
      g4[f] = acc;
    }
  } else if ((tile_words & 3) == 0) {
    uint4* g4 = (uint4*)g;
    for (int f = lt; f < n4; f += tps) {
      const uint4 v = *(const uint4*)(sw + cbp_at<PAD>(4 * (int64_t)f));
      g4[f] = INW ? inword4(v, false) : v;
    }
  } else {
    for (int64_t i = lt; i < tile_words; i += tps) g[i] = sw[i];
  }
}

template <bool PAD, bool INV, bool MULTI>
__global__ __launch_bounds__(1024) void qc_conv_rows(uint32_t* __restrict__ Bw, int64_t tile_words, int64_t ntiles,
                                                     int64_t tiles_per_inst, int64_t inst_words, int mp,
                                                     int fold) {
  static_assert(PAD || !MULTI, "several strings per workgroup need the blocked form");
  extern __shared__ uint32_t sw_all[];
  if constexpr (!MULTI) {
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
      uint32_t* g = Bw + (tile / tiles_per_inst) * inst_words + (tile % tiles_per_inst) * tile_words;
      rows_load<PAD, false>(g, sw_all, tile_words, threadIdx.x, blockDim.x);
      __syncthreads();
      if constexpr (PAD)
        cbp_convert<true, INV>(sw_all, 0, mp, threadIdx.x, blockDim.x, BarrierSync());
      else
        cbt_convert(sw_all, mp, INV, threadIdx.x, blockDim.x, BarrierSync());
      rows_store<PAD, false>(g, sw_all, tile_words, mp, fold && ((tile % tiles_per_inst) & 1), threadIdx.x,
                             blockDim.x);
      __syncthreads();
    }
  } else {
    const int lgt = mp - 10;  // log2 lanes per string: one lane per 1024-bit block
    const int tps = 1 << lgt;
    const int S = (int)blockDim.x >> lgt;  // strings per workgroup
    const int lt = (int)threadIdx.x & (tps - 1);
    const int sl = (int)threadIdx.x >> lgt;
    const int base = sl * (int)cbp_words(tile_words, true);  // this lane's string in LDS
    uint32_t* sw = sw_all + base;
    // 32-bit tile arithmetic (tile counts stay far below 2^31); the global
    // pointer is recomputed after the conversion rather than held through it
    const unsigned tpi = (unsigned)tiles_per_inst;
    for (unsigned t0 = blockIdx.x * (unsigned)S; t0 < (unsigned)ntiles; t0 += gridDim.x * (unsigned)S) {
      // every string group runs the barriers of the conversion; one past the
      // last tile converts its own LDS region and stores nothing
      const unsigned tile = t0 + (unsigned)sl;
      const bool live = tile < (unsigned)ntiles;
      if (live)
        rows_load<true, INV>(Bw + (int64_t)(tile / tpi) * inst_words + (int64_t)(tile % tpi) * tile_words, sw,
                             tile_words, lt, tps);
      __syncthreads();
      cbp_convert<true, INV, false>(sw_all, base, mp, lt, tps, BarrierSync());
      if (live)
        rows_store<true, !INV>(Bw + (int64_t)(tile / tpi) * inst_words + (int64_t)(tile % tpi) * tile_words, sw,
                               tile_words, mp, fold && ((tile % tpi) & 1), lt, tps);
      __syncthreads();
    }
  }
}

// the blocked, padded form needs 1024-bit blocks: strings of >= 2^10 coefficients
bool row_pad(int mp) { return mp >= 10; }
// LDS of one workgroup: one string, or 1024 / (blocks per string) padded strings
size_t row_lds_bytes(int64_t tile_words, int mp) {
  return row_pad(mp) ? (size_t)cbp_words(32 * 1024, true) * 4 : (size_t)tile_words * 4;
}

hipError_t launch_conv_rows(unsigned cap, hipStream_t st, uint32_t* Bw, int64_t tile_words, int64_t ntiles,
                            int64_t tpi, int64_t inst_words, int mp, int inverse, int fold) {
  if (row_pad(mp) && (tile_words < 32 || tile_words > 32 * 1024)) return hipErrorInvalidValue;
  const size_t lds = row_lds_bytes(tile_words, mp);
  const bool multi = row_pad(mp) && tile_words < 32 * 1024;
  const int64_t per_wg = multi ? 1024 / (tile_words >> 5) : 1;  // strings per workgroup
  const int64_t need = (ntiles + per_wg - 1) / per_wg;
  const unsigned grid = (unsigned)(need < (int64_t)cap ? (need < 1 ? 1 : need) : cap);
#define CONV_ROWS_LAUNCH(PAD, INV, MULTI)                                                                    \
  hipLaunchKernelGGL((qc_conv_rows<PAD, INV, MULTI>), dim3(grid), dim3(1024), lds, st, Bw, tile_words, ntiles, \
                     tpi, inst_words, mp, fold)
  if (multi) {
    if (inverse)
      CONV_ROWS_LAUNCH(true, true, true);
    else
      CONV_ROWS_LAUNCH(true, false, true);
  } else if (row_pad(mp)) {
    if (inverse)
      CONV_ROWS_LAUNCH(true, true, false);
    else
      CONV_ROWS_LAUNCH(true, false, false);
  } else {
    if (inverse)
      CONV_ROWS_LAUNCH(false, true, false);
    else
      CONV_ROWS_LAUNCH(false, false, false);
  }
#undef CONV_ROWS_LAUNCH
  return hipGetLastError();
}

// ---- FFT radix pass -----------------------------------------------------------
// Tile: T = 2^R elements spaced 2^lo apart, times C consecutive low columns.
// Layers k in [klo, khi] (lo <= klo, khi < lo + R).  forward: u ^= c v; v ^= u;
// inverse: v ^= u; u ^= c v; c = C(idx_u >> (k+1)).
// NP = 32: a tile element holds all 32 planes of an element.  NP = 16 (half
// mode, blockIdx.z = the half): it holds one GF(2^16) half, which is enough
// for layers whose twiddles all lie in the subfield GF(2^16) (k >= m - 16,
// checked on the host): there the two halves transform independently, so a
// tile of the same LDS holds twice the elements (more layers per pass).
enum { PASS_FWD = 0, PASS_INV = 1, PASS_FUSED = 2 };

struct FftPassArgs {
  uint32_t* X;
  int64_t M;
  int lo, R, C, lgC;
  int mode;
  int khi;           // highest layer to apply (FWD/FUSED: start, INV: end)
  int first;         // load only planes 0..15 of the lower half (top layer input) and embed the chunks
  int last;          // completes the inverse transform: unpack the chunk products before the write-back
  const uint32_t* ahat;
  const uint32_t* const* ahats;  // per-instance A^ (batch) or NULL
  int64_t XS_words;  // instance stride of X in 32-bit words (blockIdx.y)
  // planes 0..15 are read from S0 and written to D0 (X when NULL): a first pass
  // in half mode cannot overwrite the chunk planes 0..15 that its other half
  // still reads, so it writes half 0 to D0 = the idle row buffer, from where the
  // next pass reads it (S0)
  const uint32_t* S0;
  uint32_t* D0;
  // Plane layout of what the pass reads (gin) and writes (gout): 0 = plane
  // major, element idx of plane p at block p*M + idx; 1 = planes grouped by
  // four, at block ((p/4)*M + idx)*4 + p%4, so that the C consecutive
  // columns of a tile are C*64 contiguous bytes of four planes.  Only the
  // passes exchange the grouped form: the first pass reads the plane-major
  // chunk planes, and the last pass writes plane-major into Xo.
  int gin, gout;
  uint32_t* Xo;  // destination of the planes not sent to D0 (X when NULL)
};

// 16-byte block of (plane gp, element idx) in plane-major (g = 0) or
// grouped-by-four (g = 1) layout
CUOT_DEV int64_t fft_xaddr(int g, int gp, int64_t idx, int64_t M) {
  return g ? ((((int64_t)(gp >> 2)) * M + idx) << 2) | (int64_t)(gp & 3) : (int64_t)gp * M + idx;
}

// staged word f of a tile -> (plane p, tile row t, column c); the grouped
// order keeps the four planes of a group and the columns innermost, so that
// consecutive lanes touch consecutive blocks of the grouped layout
CUOT_DEV void fft_split(int f, int g, int lgC, int R, int& c, int& t, int& p) {
  if (g) {
    const int q = f & 3;
    c = (f >> 2) & ((1 << lgC) - 1);
    const int r = f >> (2 + lgC);
    t = r & ((1 << R) - 1);
    p = ((r >> R) << 2) | q;
  } else {
    c = f & ((1 << lgC) - 1);
    const int pt = f >> lgC;
    t = pt & ((1 << R) - 1);
    p = pt >> R;
  }
}

// LDS words per tile element: NP planes x 4 words + 4 pad words (with the pad
// the element stride is 4 mod 64 banks, so 16 consecutive elements' words of
// one plane fall on distinct banks)
template <int NP>
constexpr int fstride() {
  return NP * 4 + 4;
}
template <int NP>
constexpr int ftc_max() {
  return FFT_TC * 32 / NP;  // tile elements for the same LDS footprint
}
// half-mode workgroups have twice the lanes (the 16-plane tasks need half the
// registers): 8 waves each, 4 per SIMD with two workgroups per CU
template <int NP>
constexpr int fthreads() {
  return NP == 16 ? 2 * FFT_THREADS : FFT_THREADS;
}
template <int NP>
constexpr int fld() {  // staged 16-byte words per lane: ftc_max * NP <= fld * fthreads
  return FFT_LD * FFT_THREADS / fthreads<NP>();
}

// The top forward layer (k = m - 1, tile-relative s = R - 1) of the first
// pass: the upper half of the padded input is zero, so u += c*0 leaves u and
// v += u makes v a copy of u.
template <int NP>
__device__ __forceinline__ void fft_copy_top(int R, int C, int lgC, uint32_t* lds) {
  constexpr int ST = fstride<NP>();
  const int s = R - 1;
  const int ntask = ((1 << R) >> 1) * C * 4;
  for (int tau = threadIdx.x; tau < ntask; tau += blockDim.x) {
    const int w = tau & 3;
    const int c = (tau >> 2) & (C - 1);
    const int bp = (tau >> 2) >> lgC;  // < 2^s: tu = bp, tv = bp + 2^s
    const uint32_t* pu = lds + fslot(bp * C + c) * ST + w;
    uint32_t* pv = lds + fslot((bp | (1 << s)) * C + c) * ST + w;
#pragma unroll
    for (int p = 0; p < NP; p++) pv[4 * p] = pu[4 * p];
  }
  __syncthreads();
}

// acc = C(g) * xv: wave-uniform group indices take scalar branches; a
// subfield constant (high half zero: every layer k >= m - 16) takes two
// 16-bit products instead of the Karatsuba product.
template <int NP>
__device__ __forceinline__ void fft_mul(const uint32_t* ctab, uint32_t g, const uint32_t* xv, uint32_t* acc) {
  const uint32_t g0 = __builtin_amdgcn_readfirstlane(g);
  if constexpr (NP == 16) {
    // one GF(2^16) half, twiddle in GF(2^16)
    if (__all(g == g0))
      gf16_mulc(xv, __builtin_amdgcn_readfirstlane(ctab_const(ctab, g0)), acc);
    else
      gf16_mulc(xv, ctab_const(ctab, g), acc);
  } else {
    if (__all(g == g0)) {
      const uint32_t cst = __builtin_amdgcn_readfirstlane(ctab_const(ctab, g0));
      if ((cst >> 16) == 0u)
        gf32_mulc_sub<true>(xv, cst, acc);
      else
        gf32_mulc<true>(xv, cst, acc);
    } else {
      const uint32_t cst = ctab_const(ctab, g);
      if (__all((cst >> 16) == 0u))
        gf32_mulc_sub<false>(xv, cst, acc);
      else
        gf32_mulc<false>(xv, cst, acc);
    }
  }
}

template <int NP>
__device__ __forceinline__ void fft_layer(int lo, int R, int C, int lgC, uint32_t* lds, int64_t hib, int64_t cb,
                                          int k, bool inv) {
  constexpr int ST = fstride<NP>();
  const uint32_t* ctab = lds + (C << R) * ST;
  const int s = k - lo;
  const int T = 1 << R;
  const int ntask = (T >> 1) * C * 4;
  for (int tau = threadIdx.x; tau < ntask; tau += blockDim.x) {
    const int w = tau & 3;
    const int c = (tau >> 2) & (C - 1);
    const int bp = (tau >> 2) >> lgC;
    const int tu = ((bp >> s) << (s + 1)) | (bp & ((1 << s) - 1));
    const int tv = tu | (1 << s);
    const int64_t idx = (hib << (lo + R)) | ((int64_t)tu << lo) | (cb * C + c);
    const uint64_t g = (uint64_t)idx >> (k + 1);
    uint32_t* pu = lds + fslot(tu * C + c) * ST + w;
    uint32_t* pv = lds + fslot(tv * C + c) * ST + w;
    // forward: u += c*v, v += u; inverse: v += u, u += c*v.  Only v (or the
    // new v) and the product are held in registers; u streams through LDS.
    uint32_t xv[NP], acc[NP];
    if (inv) {
#pragma unroll
      for (int p = 0; p < NP; p++) {
        xv[p] = pv[4 * p] ^ pu[4 * p];
        pv[4 * p] = xv[p];
      }
    } else {
#pragma unroll
      for (int p = 0; p < NP; p++) xv[p] = pv[4 * p];
    }
    fft_mul<NP>(ctab, (uint32_t)g, xv, acc);
    if (inv) {
#pragma unroll
      for (int p = 0; p < NP; p++) pu[4 * p] ^= acc[p];
    } else {
#pragma unroll
      for (int p = 0; p < NP; p++) {
        const uint32_t un = pu[4 * p] ^ acc[p];
        pu[4 * p] = un;
        pv[4 * p] = xv[p] ^ un;
      }
    }
  }
  __syncthreads();
}

// Chunk embedding sweeps over the staged tile (element-independent linear maps
// on the planes of each element word): the 16 chunk-bit planes become the
// coordinates of u(theta) (all 32, or the 16 of half h) before the first
// forward layer, and after the last inverse layer the 32 coordinates of a chunk
// product become its 31 coefficient planes (plane 31 zero), which the inverse
// conversion and the fold read.
template <int NP>
__device__ __forceinline__ void fft_pack_tile(int TC, int h, uint32_t* lds) {
  constexpr int ST = fstride<NP>();
  for (int tau = threadIdx.x; tau < TC * 4; tau += blockDim.x) {
    uint32_t* pe = lds + fslot(tau >> 2) * ST + (tau & 3);
    uint32_t in[16], out[NP];
#pragma unroll
    for (int p = 0; p < 16; p++) in[p] = pe[4 * p];
    if constexpr (NP == 32) {
      gf32_pack_planes(in, out);
    } else {
      if (h)
        gf32_pack_planes_half<16>(in, out);
      else
        gf32_pack_planes_half<0>(in, out);
    }
#pragma unroll
    for (int p = 0; p < NP; p++) pe[4 * p] = out[p];
  }
  __syncthreads();
}

__device__ __forceinline__ void fft_unpack_tile(int TC, uint32_t* lds) {
  for (int tau = threadIdx.x; tau < TC * 4; tau += blockDim.x) {
    uint32_t* pe = lds + fslot(tau >> 2) * FFT_STRIDE + (tau & 3);
    uint32_t in[32], out[31];
#pragma unroll
    for (int p = 0; p < 32; p++) in[p] = pe[4 * p];
    gf32_unpack_planes(in, out);
#pragma unroll
    for (int p = 0; p < 31; p++) pe[4 * p] = out[p];
    pe[4 * 31] = 0u;
  }
  __syncthreads();
}

// pointwise product by A^ (per-element constant) on the tile
__device__ __forceinline__ void fft_pointwise(int lo, int R, int C, int lgC,
                                                      const uint32_t* __restrict__ ahat, uint32_t* lds, int64_t hib,
                                                      int64_t cb) {
  const int TC = (1 << R) * C;
  for (int tau = threadIdx.x; tau < TC * 4; tau += blockDim.x) {
    const int w = tau & 3;
    const int el = tau >> 2;  // = t*C + c
    const int t = el >> lgC, c = el & (C - 1);
    const int64_t idx = (hib << (lo + R)) | ((int64_t)t << lo) | (cb * C + c);
    uint32_t* pe = lds + fslot(el) * FFT_STRIDE + w;
    uint32_t x[32], y[32];
#pragma unroll
    for (int p = 0; p < 32; p++) x[p] = pe[4 * p];
    gf32_mulc<false>(x, ahat[idx], y);
#pragma unroll
    for (int p = 0; p < 32; p++) pe[4 * p] = y[p];
  }
  __syncthreads();
}

template <int NP>
__global__ __launch_bounds__(fthreads<NP>()) __attribute__((amdgpu_waves_per_eu(NP == 16 ? 4 : 2))) void qc_fft_pass(
    FftPassArgs a) {
  extern __shared__ uint32_t lds[];
  constexpr int ST = fstride<NP>();
  constexpr int NT = fthreads<NP>();
  constexpr int LD = fld<NP>();
  const int64_t ioff = (int64_t)blockIdx.y * a.XS_words;
  a.X += ioff;
  const uint4* S0 = (const uint4*)(a.S0 ? a.S0 + ioff : a.X);
  uint4* Y4 = (uint4*)(a.Xo ? a.Xo + ioff : a.X);
  uint4* D0 = a.D0 ? (uint4*)(a.D0 + ioff) : Y4;
  if (a.ahats) a.ahat = a.ahats[blockIdx.y];
  const int h = NP == 16 ? (int)blockIdx.z : 0;  // half of a half-mode pass
  const int pbase = 16 * h;
  const int T = 1 << a.R;
  const int TC = T * a.C;
  const int64_t cblocks = ((int64_t)1 << a.lo) >> a.lgC;
  const int64_t ntiles = a.M / TC;
  const int64_t half = a.M >> 1;
  ctab_fill(lds + TC * ST, threadIdx.x, blockDim.x);  // read after the first barrier
  // Each lane stages up to FFT_LD 16-byte plane words of a tile (f = (p*T +
  // t)*C + c: c, t fastest, contiguous runs; TC * NP <= FFT_LD * blockDim).
  // Two WGs share a CU, so one's loads and stores overlap the other's layers.
  const int n4 = TC * NP;
  const uint4* X4 = (const uint4*)a.X;
  uint4 v[LD];
  // staging addresses come from an opaque copy of the lane index, so that the
  // compiler recomputes them per phase instead of holding them through the
  // layers
  auto opaque_tid = [] {
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
  };
  // the first pass reads the 16 chunk planes (lower half only) whatever the half
  auto load_tile = [&](int64_t tl) {
    const int64_t lcb = tl & (cblocks - 1), lhib = tl >> (a.lo - a.lgC);
    const int tid = opaque_tid();
#pragma unroll
    for (int u = 0; u < LD; u++) {
      const int f = tid + u * NT;
      v[u] = z4();
      if (f < n4) {
        int c, t, p;
        fft_split(f, a.gin, a.lgC, a.R, c, t, p);
        const int64_t idx = (lhib << (a.lo + a.R)) | ((int64_t)t << a.lo) | (lcb * a.C + c);
        const int gp = pbase + p;  // global plane
        if (!a.first)
          v[u] = (gp < 16 ? S0 : X4)[fft_xaddr(a.gin, gp, idx, a.M)];
        else if (idx < half && p < 16)
          v[u] = X4[(int64_t)p * a.M + idx];  // the chunk planes (plane major)
      }
    }
  };
  // workgroups are dealt to the 8 XCDs in turn: renumber them so that one
  // XCD's workgroups take neighbouring tiles, whose column blocks share
  // 128-byte lines in that XCD's L2
  const int64_t G = gridDim.x, bx = blockIdx.x;
  const int64_t slot = (G & 7) ? bx : (bx & 7) * (G >> 3) + (bx >> 3);
  // A tile's loads are issued in the previous tile's write-back phase (after
  // its results have left the LDS for registers), so that the HBM reads and
  // writes of consecutive tiles are in flight together.
  if (slot < ntiles) load_tile(slot);
  for (int64_t tile = slot; tile < ntiles; tile += G) {
    const int64_t cb = tile & (cblocks - 1), hib = tile >> (a.lo - a.lgC);
#pragma unroll
    for (int u = 0; u < LD; u++) {
      const int f = (int)threadIdx.x + u * (int)blockDim.x;
      if (f < n4) {
        int c, t, p;
        fft_split(f, a.gin, a.lgC, a.R, c, t, p);
        *(uint4*)&lds[fslot(t * a.C + c) * ST + p * 4] = v[u];
      }
    }
    __syncthreads();
    if (a.first) fft_pack_tile<NP>(TC, h, lds);
    // one step loop so that the layer body is instantiated once: forward
    // layers khi..lo, the pointwise product (fused pass), inverse layers lo..khi
    const int nl = a.khi - a.lo + 1;
    const int s_fwd = (a.mode == PASS_FWD || a.mode == PASS_FUSED) ? nl : 0;
    const int s_pw = a.mode == PASS_FUSED ? 1 : 0;
    const int s_inv = (a.mode == PASS_INV || a.mode == PASS_FUSED) ? nl : 0;
    for (int step = 0; step < s_fwd + s_pw + s_inv; step++) {
      if (step == s_fwd && s_pw) {
        if constexpr (NP == 32) fft_pointwise(a.lo, a.R, a.C, a.lgC, a.ahat, lds, hib, cb);
      } else {
        const bool inv = step >= s_fwd;
        const int k = inv ? a.lo + (step - s_fwd - s_pw) : a.khi - step;
        if (a.first && step == 0)
          fft_copy_top<NP>(a.R, a.C, a.lgC, lds);  // top layer on a zero upper half: v = u
        else
          fft_layer<NP>(a.lo, a.R, a.C, a.lgC, lds, hib, cb, k, inv);
      }
    }
    if constexpr (NP == 32) {
      if (a.last) fft_unpack_tile(TC, lds);
    }
    uint4 w[LD];
    const int tid = opaque_tid();
#pragma unroll
    for (int u = 0; u < LD; u++) {
      const int f = tid + u * NT;
      w[u] = z4();
      if (f < n4) {
        int c, t, p;
        fft_split(f, a.gout, a.lgC, a.R, c, t, p);
        w[u] = *(const uint4*)&lds[fslot(t * a.C + c) * ST + p * 4];
      }
    }
    if (tile + G < ntiles) load_tile(tile + G);
#pragma unroll
    for (int u = 0; u < LD; u++) {
      const int f = tid + u * NT;
      if (f < n4) {
        int c, t, p;
        fft_split(f, a.gout, a.lgC, a.R, c, t, p);
        const int64_t idx = (hib << (a.lo + a.R)) | ((int64_t)t << a.lo) | (cb * a.C + c);
        const int gp = pbase + p;
        (gp < 16 ? D0 : Y4)[fft_xaddr(a.gout, gp, idx, a.M)] = w[u];
      }
    }
    __syncthreads();
  }
}


// ---- scalar transform of the single polynomial a (A^ = its evaluations) ------
// Same conversion + FFT as the bitsliced path, on one 32-bit field element per
// index instead of 128 bitsliced rows (A^ is computed once per code).
CUOT_DEV uint32_t gf32_mul_dev(uint32_t a, uint32_t b) { return gf32_mul(a, b); }

__global__ void qcs_pack(const QcPrep* __restrict__ pr, QcPrep pr1, int64_t M) {
  const QcPrep q = pr ? pr[blockIdx.y] : pr1;
  const uint8_t* a = q.a_bits;
  const int64_t n = q.n;
  uint32_t* A = q.ahat;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = 0;
    if (c < (M >> 1)) {
      for (int b = 0; b < 16; b++) {
        const int64_t src = 16 * c + b;
        if (src < n) v |= (uint32_t)((a[src >> 3] >> (src & 7)) & 1u) << b;
      }
    }
    A[c] = gf32_pack_chunk(v);
  }
}

__global__ void qcs_hi_fix(const QcPrep* __restrict__ pr, QcPrep pr1, int64_t Mp, int k) {
  uint32_t* A = (pr ? pr[blockIdx.y] : pr1).ahat;
  uint32_t e[16];
  const int ne = lower_terms(k, e);
  if (ne == 0) return;
  const int64_t emax = e[ne - 1], half = (int64_t)1 << k;
  const int le = __builtin_ctz((uint32_t)emax);
  const int64_t total = (Mp >> (k + 1)) << le;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = t >> le, i = t & (emax - 1);
    uint32_t* h = A + g * 2 * half + half;
    uint32_t v = h[i];
    for (int r = 0; r < ne; r++)
      if ((int64_t)e[r] > i) v ^= h[i + half - e[r]];
    h[i] = v;
  }
}

__global__ void qcs_lo_update(const QcPrep* __restrict__ pr, QcPrep pr1, int64_t Mp, int k) {
  uint32_t* A = (pr ? pr[blockIdx.y] : pr1).ahat;
  uint32_t e[16];
  const int ne = lower_terms(k, e);
  if (ne == 0) return;
  const int64_t half = (int64_t)1 << k;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (Mp >> 1); t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = t >> k, j = t & (half - 1);
    uint32_t* l = A + g * 2 * half;
    uint32_t v = l[j];
    for (int r = 0; r < ne; r++) {
      if ((int64_t)e[r] > j) break;
      v ^= l[half + j - e[r]];
    }
    l[j] = v;
  }
}

__global__ void qcs_fft_layer(const QcPrep* __restrict__ pr, QcPrep pr1, int64_t M, int k) {
  uint32_t* A = (pr ? pr[blockIdx.y] : pr1).ahat;
  const int64_t half = (int64_t)1 << k;
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < (M >> 1); b += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = b >> k, j = b & (half - 1);
    const int64_t u = g * 2 * half + j, v = u + half;
    const uint32_t c = chunk_const(c_beta, (uint64_t)g);
    const uint32_t xu = A[u] ^ gf32_mul_dev(c, A[v]);
    A[u] = xu;
    A[v] = A[v] ^ xu;
  }
}

// The whole preparation of one polynomial in one workgroup's LDS when its
// 2^m words fit (small codes): the same pack, conversion levels and FFT layers
// as the kernels above, separated by barriers instead of launches.
constexpr int PREP_LDS_MMAX = 15;  // 2^15 words = 128 KB

__global__ __launch_bounds__(1024) void qcs_prepare_lds(const QcPrep* __restrict__ pr, QcPrep pr1, int m) {
  extern __shared__ uint32_t L[];
  const QcPrep q = pr ? pr[blockIdx.x] : pr1;
  const int64_t M = (int64_t)1 << m, Mh = M >> 1;
  const int tid = threadIdx.x, nth = blockDim.x;
  for (int64_t c = tid; c < M; c += nth) {
    uint32_t v = 0;
    if (c < Mh) {
      for (int b = 0; b < 16; b++) {
        const int64_t src = 16 * c + b;
        if (src < q.n) v |= (uint32_t)((q.a_bits[src >> 3] >> (src & 7)) & 1u) << b;
      }
    }
    L[c] = gf32_pack_chunk(v);
  }
  __syncthreads();
  // monomial -> novel basis on the lower half (division by s_k, k = m-2 .. 0)
  for (int k = m - 2; k >= 0; k--) {
    uint32_t e[16];
    const int ne = lower_terms(k, e);
    if (ne == 0) continue;
    const int64_t half = (int64_t)1 << k;
    const int64_t emax = e[ne - 1];
    const int le = __builtin_ctz((uint32_t)emax);
    for (int64_t t = tid; t < ((Mh >> (k + 1)) << le); t += nth) {
      const int64_t g = t >> le, i = t & (emax - 1);
      uint32_t* h = L + g * 2 * half + half;
      uint32_t v = h[i];
      for (int r = 0; r < ne; r++)
        if ((int64_t)e[r] > i) v ^= h[i + half - e[r]];
      h[i] = v;
    }
    __syncthreads();
    for (int64_t t = tid; t < (Mh >> 1); t += nth) {
      const int64_t g = t >> k, j = t & (half - 1);
      uint32_t* l = L + g * 2 * half;
      uint32_t v = l[j];
      for (int r = 0; r < ne; r++) {
        if ((int64_t)e[r] > j) break;
        v ^= l[half + j - e[r]];
      }
      l[j] = v;
    }
    __syncthreads();
  }
  // FFT layers m-1 .. 0 (the top one on a zero upper half copies the lower)
  for (int k = m - 1; k >= 0; k--) {
    const int64_t half = (int64_t)1 << k;
    for (int64_t b = tid; b < Mh; b += nth) {
      const int64_t g = b >> k, j = b & (half - 1);
      const int64_t u = g * 2 * half + j, v = u + half;
      const uint32_t c = chunk_const(c_beta, (uint64_t)g);
      const uint32_t xu = L[u] ^ gf32_mul_dev(c, L[v]);
      L[u] = xu;
      L[v] = L[v] ^ xu;
    }
    __syncthreads();
  }
  for (int64_t c = tid; c < M; c += nth) q.ahat[c] = L[c];
}

// Tiled forms for long polynomials: the conversion levels below QCS_LB and the
// FFT layers run in LDS tiles of 2^QCS_LB words, one launch each instead of
// one per level / layer (the upper conversion levels keep their launches).
constexpr int QCS_LB = 12;

// conversion levels kmax..0 of the lower half on consecutive tiles
__global__ __launch_bounds__(1024) void qcs_conv_tile(const QcPrep* __restrict__ pr, QcPrep pr1, int kmax) {
  __shared__ uint32_t L[1 << QCS_LB];
  constexpr int TW = 1 << QCS_LB;
  uint32_t* A = (pr ? pr[blockIdx.y] : pr1).ahat + ((int64_t)blockIdx.x << QCS_LB);
  for (int i = threadIdx.x; i < TW; i += blockDim.x) L[i] = A[i];
  __syncthreads();
  for (int k = kmax; k >= 0; k--) {
    uint32_t e[16];
    const int ne = lower_terms(k, e);
    if (ne == 0) continue;
    const int half = 1 << k;
    const int emax = (int)e[ne - 1];
    const int le = __builtin_ctz((uint32_t)emax);
    for (int t = threadIdx.x; t < ((TW >> (k + 1)) << le); t += blockDim.x) {
      const int g = t >> le, i = t & (emax - 1);
      uint32_t* h = L + g * 2 * half + half;
      uint32_t v = h[i];
      for (int r = 0; r < ne; r++)
        if ((int)e[r] > i) v ^= h[i + half - (int)e[r]];
      h[i] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < (TW >> 1); t += blockDim.x) {
      const int g = t >> k, j = t & (half - 1);
      uint32_t* l = L + g * 2 * half;
      uint32_t v = l[j];
      for (int r = 0; r < ne; r++) {
        if ((int)e[r] > j) break;
        v ^= l[half + j - (int)e[r]];
      }
      l[j] = v;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < TW; i += blockDim.x) A[i] = L[i];
}

// FFT layers khi..lo on tiles of 2^R elements spaced 2^lo apart times 2^lgC
// consecutive columns (2^(R + lgC) = 2^QCS_LB words, lgC <= lo)
__global__ __launch_bounds__(1024) void qcs_fft_tile(const QcPrep* __restrict__ pr, QcPrep pr1, int lo, int R, int lgC,
                                                    int khi) {
  __shared__ uint32_t L[1 << QCS_LB];
  constexpr int TW = 1 << QCS_LB;
  uint32_t* A = (pr ? pr[blockIdx.y] : pr1).ahat;
  const int C = 1 << lgC;
  const int64_t tile = blockIdx.x;
  const int64_t cb = tile & ((((int64_t)1 << lo) >> lgC) - 1), hib = tile >> (lo - lgC);
  auto gidx = [&](int t, int c) { return (hib << (lo + R)) | ((int64_t)t << lo) | (cb * C + c); };
  for (int i = threadIdx.x; i < TW; i += blockDim.x) L[i] = A[gidx(i >> lgC, i & (C - 1))];
  __syncthreads();
  for (int k = khi; k >= lo; k--) {
    const int s = k - lo;
    for (int b = threadIdx.x; b < (TW >> 1); b += blockDim.x) {
      const int c = b & (C - 1), bp = b >> lgC;
      const int tu = ((bp >> s) << (s + 1)) | (bp & ((1 << s) - 1)), tv = tu | (1 << s);
      const uint32_t cst = chunk_const(c_beta, (uint64_t)gidx(tu, c) >> (k + 1));
      const int u = tu * C + c, v = tv * C + c;
      const uint32_t xu = L[u] ^ gf32_mul_dev(cst, L[v]);
      L[u] = xu;
      L[v] ^= xu;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < TW; i += blockDim.x) A[gidx(i >> lgC, i & (C - 1))] = L[i];
}

// ---- output: top inverse level + overlap-add + fold + x0 ---------------------
// H_c (plane p) after the top inverse level k = m-1 on X (r = X[0..M/2), q = X[M/2..M)):
//   c <  M/2: r_c ^ XOR_{e in T, e <= c} q_{c-e}
//   c >= M/2: q_s ^ XOR_{e in T, e > s} q_{s + M/2 - e},  s = c - M/2
struct TopTerms {
  uint32_t e[16];
  int ne;
};

CUOT_DEV uint4 h_top(const uint4* __restrict__ X, int64_t M, int p, int64_t c, const TopTerms& tt) {
  const int64_t half = M >> 1;
  const uint4* r = X + (int64_t)p * M;
  const uint4* q = r + half;
  uint4 v;
  if (c < half) {
    v = r[c];
    for (int i = 0; i < tt.ne; i++) {
      if ((int64_t)tt.e[i] > c) break;
      v = x4(v, q[c - tt.e[i]]);
    }
  } else {
    const int64_t s = c - half;
    v = q[s];
    for (int i = 0; i < tt.ne; i++)
      if ((int64_t)tt.e[i] > s) v = x4(v, q[s + half - tt.e[i]]);
  }
  return v;
}

// H(s) = plane (s & 15) of chunk s>>4  ^  plane (s & 15) + 16 of chunk (s>>4) - 1
CUOT_DEV uint4 hcontrib(const uint4* __restrict__ X, int64_t M, int64_t s, const TopTerms& tt) {
  const int64_t c = s >> 4;
  const int b = (int)(s & 15);
  uint4 v = z4();
  if (c < M) v = h_top(X, M, b, c, tt);
  if (c >= 1 && c - 1 < M) v = x4(v, h_top(X, M, b + 16, c - 1, tt));
  return v;
}

__global__ void qc_output(const uint4* __restrict__ X, int64_t M, int m, int64_t XS, const QcIo* __restrict__ io,
                          QcIo io1) {
  const QcIo q = io ? io[blockIdx.y] : io1;
  const uint4* x0 = q.x;
  const int64_t n = q.n;
  uint4* y = q.y;
  X += (int64_t)blockIdx.y * XS;
  TopTerms tt;
  tt.ne = lower_terms(m - 1, tt.e);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    y[i] = x4(x0[i], x4(hcontrib(X, M, i, tt), hcontrib(X, M, i + n, tt)));
  }
}

// Output from the row layout after the folded inverse conversion (n = 16 h
// chunks, h = M/2): G_p = r_p ^ T(q_p) is the product folded mod x^n - 1 for
// coefficient plane p, and
//   y[16c + b] = x0[16c + b] ^ G_b[c] ^ G_{b+16}[(c - 1) mod h]   (G_31 = 0),
// the second term being the overlap of chunk c-1's upper 15 planes.
// Step 1 (qc_rows_fold): per plane pair b and a tile of 32 words (1024
// chunks) of every row, combine the four row strings in LDS and bit-transpose
// them into the element layout D[b][c] (16 planes of h elements).
// Step 2 (qc_output_planes): y[16c + b] = x0[16c + b] ^ D[b][c], interleaving
// the 16 planes through LDS so that both sides are contiguous.
__global__ __launch_bounds__(256) void qc_rows_fold(const uint32_t* __restrict__ Bw, int64_t hw, int64_t inst_words,
                                                    uint4* __restrict__ D, int64_t dps) {
  __shared__ uint32_t sm[128][33];
  const int b = (int)blockIdx.y;
  const uint32_t* rb = Bw + (int64_t)b * 128 * (2 * hw);
  const uint32_t* rh = Bw + (int64_t)(b + 16) * 128 * (2 * hw);
  D += (int64_t)b * dps;
  const int lane = threadIdx.x & 31;
  const int grp = threadIdx.x >> 5;
  const int64_t ntile = hw / 32;
  for (int64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
    const int64_t w0 = t * 32;
    // 128 rows x 32 words: thread f -> row f >> 3, words 4 (f & 7) .. + 3
    for (int f = threadIdx.x; f < 128 * 8; f += 256) {
      const int row = f >> 3, wq = (f & 7) * 4;
      const int64_t ro = (int64_t)row * (2 * hw) + w0 + wq;
      const uint4 a = x4(*(const uint4*)(rb + ro), *(const uint4*)(rb + ro + hw));
      uint4 h = z4();
      uint32_t hp = 0;
      if (b < 15) {
        h = x4(*(const uint4*)(rh + ro), *(const uint4*)(rh + ro + hw));
        const int64_t rp = (int64_t)row * (2 * hw) + ((w0 + wq - 1) & (hw - 1));
        hp = rh[rp] ^ rh[rp + hw];
      }
      uint32_t* S = &sm[row][wq];
      S[0] = a.x ^ (h.x << 1) ^ (hp >> 31);
      S[1] = a.y ^ (h.y << 1) ^ (h.x >> 31);
      S[2] = a.z ^ (h.z << 1) ^ (h.y >> 31);
      S[3] = a.w ^ (h.w << 1) ^ (h.z >> 31);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int tl = grp * 4 + r;
      uint4 v;
      v.x = cuot_transpose32(sm[0 * 32 + lane][tl], lane);
      v.y = cuot_transpose32(sm[1 * 32 + lane][tl], lane);
      v.z = cuot_transpose32(sm[2 * 32 + lane][tl], lane);
      v.w = cuot_transpose32(sm[3 * 32 + lane][tl], lane);
      D[(w0 + tl) * 32 + lane] = v;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void qc_output_planes(const uint4* __restrict__ D, int64_t dps, int64_t h,
                                                        const QcIo* __restrict__ io, QcIo io1) {
  __shared__ uint4 sm[16][65];
  const QcIo q = io ? io[blockIdx.y] : io1;
  const int64_t ntile = h / 64;
  for (int64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
    const int64_t c0 = t * 64;
    // D[b][c0 .. c0 + 64) for 16 planes: 1024 elements, 4 per thread
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int f = (int)threadIdx.x + u * 256;
      const int bb = f >> 6, cc = f & 63;
      sm[bb][cc] = D[(int64_t)bb * dps + c0 + cc];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int f = (int)threadIdx.x + u * 256;  // block 16 c0 + f: chunk f >> 4, plane f & 15
      const int64_t i = 16 * c0 + f;
      q.y[i] = x4(q.x[i], sm[f & 15][f >> 4]);
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

Beta g_beta;
bool g_beta_ready = false;

const Beta& beta() {
  if (!g_beta_ready) {
    gf32_cantor_basis(g_beta.v);
    g_beta_ready = true;
  }
  return g_beta;
}

int g_sm_configured = 0;

hipError_t configure_lds() {
  if (g_sm_configured) return hipSuccess;
  {
    hipError_t e0 = hipMemcpyToSymbol(HIP_SYMBOL(c_beta), beta().v, sizeof(uint32_t) * 32);
    if (e0 != hipSuccess) return e0;
  }
  hipError_t e = hipFuncSetAttribute((const void*)qc_fft_pass<32>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (ftc_max<32>() * fstride<32>() + CTAB_WORDS) * 4);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void*)qc_fft_pass<16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (ftc_max<16>() * fstride<16>() + CTAB_WORDS) * 4);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void*)qcs_prepare_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                          ((int64_t)1 << PREP_LDS_MMAX) * 4);
  if (e != hipSuccess) return e;
  {
    const void* fns[6] = {(const void*)qc_conv_rows<false, false, false>,
                          (const void*)qc_conv_rows<false, true, false>,
                          (const void*)qc_conv_rows<true, false, false>,
                          (const void*)qc_conv_rows<true, true, false>,
                          (const void*)qc_conv_rows<true, false, true>,
                          (const void*)qc_conv_rows<true, true, true>};
    for (int i = 0; i < 6; i++) {
      e = hipFuncSetAttribute(fns[i], hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)row_lds_bytes(((int64_t)1 << ROW_MP_MAX) / 32, ROW_MP_MAX));
      if (e != hipSuccess) return e;
    }
  }
  e = hipFuncSetAttribute((const void*)qc_conv_lds<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (1 << CONV_B) * 16);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void*)qc_conv_lds<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (1 << CONV_B) * 16);
  if (e != hipSuccess) return e;
  g_sm_configured = 1;
  return hipSuccess;
}

struct Bufs {
  uint32_t* X;   // B x 32 planes x M x 16 B (element layout), instance stride XS blocks
  uint4* Y;      // CONV_G planes x M x 16 B (ping-pong of the element-level path, B == 1)
  uint32_t* Bw;  // B x 32 planes x 128 rows x M bits (row layout)
  int B;
  int64_t XS;    // blocks per instance in X (and 16-byte units per instance in Bw)
};

Bufs bufs(void* scratch, int64_t M, int B) {
  Bufs b;
  b.B = B;
  b.XS = 32 * M;
  b.X = (uint32_t*)scratch;
  b.Bw = (uint32_t*)((char*)scratch + (size_t)B * 32 * M * 16);
  b.Y = (uint4*)((char*)scratch + (size_t)B * 64 * M * 16);
  return b;
}

int conv_b(int mpoly) { return mpoly < CONV_B ? mpoly : CONV_B; }

unsigned capg(int64_t v, int64_t cap) { return (unsigned)(v < 1 ? 1 : (v > cap ? cap : v)); }

// forward conversion of planes [0, P), element-level path (single instance)
hipError_t conv_forward_elem(Bufs b, int64_t M, int P, int mp, hipStream_t st) {
  const int64_t Mp = (int64_t)1 << mp;
  const int B = conv_b(mp);
  for (int p0 = 0; p0 < P; p0 += CONV_G) {
    const int G = P - p0 < CONV_G ? P - p0 : CONV_G;
    uint4* x = (uint4*)b.X + (int64_t)p0 * M;
    const uint4* cur = x;
    for (int k = mp - 1; k >= B; k--) {
      uint4* nxt = (cur == x) ? b.Y : x;
      hipLaunchKernelGGL(qc_conv_level<false>, dim3(grid_for((int64_t)G * Mp, 1 << 14)), dim3(256), 0, st, cur, nxt,
                         M, G, Mp, k);
      cur = nxt;
    }
    if (B > 0)
      hipLaunchKernelGGL(qc_conv_lds<false>, dim3(capg((int64_t)G * (Mp >> B), 8192)), dim3(512),
                         (size_t)(1 << B) * 16, st, cur, x, M, G, Mp, B);
  }
  return hipGetLastError();
}

hipError_t conv_inverse_elem(Bufs b, int64_t M, int P, int mtop, hipStream_t st) {
  const int64_t Mp = M;
  const int B = conv_b(mtop);
  for (int p0 = 0; p0 < P; p0 += CONV_G) {
    const int G = P - p0 < CONV_G ? P - p0 : CONV_G;
    uint4* x = (uint4*)b.X + (int64_t)p0 * M;
    const int nglob = mtop - B;
    uint4* first_dst = (nglob & 1) ? b.Y : x;
    if (B > 0)
      hipLaunchKernelGGL(qc_conv_lds<true>, dim3(capg((int64_t)G * (Mp >> B), 8192)), dim3(512),
                         (size_t)(1 << B) * 16, st, x, first_dst, M, G, Mp, B);
    const uint4* cur = (B > 0) ? first_dst : x;
    for (int k = B; k < mtop; k++) {
      uint4* nxt = (cur == x) ? b.Y : x;
      hipLaunchKernelGGL(qc_conv_level<true>, dim3(grid_for((int64_t)G * Mp, 1 << 14)), dim3(256), 0, st, cur, nxt,
                         M, G, Mp, k);
      cur = nxt;
    }
  }
  return hipGetLastError();
}

bool row_path(int mp) { return mp >= ROW_MP_MIN && mp <= ROW_MP_MAX; }

// forward conversion of planes [0, P) of the lower-half polynomial (2^mp
// elements) of every instance, through the row layout
hipError_t conv_forward(Bufs b, int64_t M, int P, int mp, hipStream_t st) {
  if (!row_path(mp)) return conv_forward_elem(b, M, P, mp, st);
  const int64_t Mp = (int64_t)1 << mp;
  const int64_t row_words = Mp / 32;
  const int64_t inst_words = b.XS * 4;  // Bw instance stride (same bytes as X)
  hipError_t e = cuot_launch_transpose_planes2(b.X, Mp, b.Bw, false, b.B * P, P, b.XS, M, inst_words,
                                               128 * row_words, st);
  if (e != hipSuccess) return e;
  const int64_t tpi = (int64_t)P * 128;
  cuot_rec_begin(st, "qc_conv_rows", 2.0 * (double)tpi * b.B * row_words * 4);
  e = launch_conv_rows(4096, st, b.Bw, row_words, tpi * b.B, tpi, inst_words, mp, 0, 0);
  cuot_rec_end(st);
  if (e != hipSuccess) return e;
  return cuot_launch_transpose_planes2(b.Bw, Mp, b.X, true, b.B * P, P, inst_words, 128 * row_words, b.XS, M, st);
}

// inverse conversion levels 0..mtop-1 of planes [0, P) of the full polynomial
// (2^(mtop+1) elements): each half string is converted on its own
hipError_t conv_inverse(Bufs b, int64_t M, int P, int mtop, bool fold, hipStream_t st) {
  if (!row_path(mtop)) return conv_inverse_elem(b, M, P, mtop, st);
  const int64_t row_words = M / 32;
  const int64_t inst_words = b.XS * 4;
  hipError_t e = cuot_launch_transpose_planes2(b.X, M, b.Bw, false, b.B * P, P, b.XS, M, inst_words,
                                               128 * row_words, st);
  if (e != hipSuccess) return e;
  const int64_t tile_words = row_words / 2;
  const int64_t tpi = (int64_t)P * 128 * 2;
  cuot_rec_begin(st, "qc_conv_rows", 2.0 * (double)tpi * b.B * tile_words * 4);
  e = launch_conv_rows(8192, st, b.Bw, tile_words, tpi * b.B, tpi, inst_words, mtop, 1, fold ? 1 : 0);
  cuot_rec_end(st);
  if (e != hipSuccess || fold) return e;  // folded: the output is read from the row layout
  return cuot_launch_transpose_planes2(b.Bw, M, b.X, true, b.B * P, P, inst_words, 128 * row_words, b.XS, M, st);
}

// two resident WGs per CU; every WG walks many tiles
void pass_grid(int64_t ntiles, int B, int Z, unsigned& g) {
  int64_t cap = FFT_WGS / ((B < 1 ? 1 : B) * Z);
  if (cap < 1) cap = 1;
  g = (unsigned)(ntiles < cap ? ntiles : cap);
}

FftPassArgs pass_args(const Bufs& b, int64_t M, int lo, int R, int C, int mode, int khi, int first, int last,
                      const uint32_t* ahat1, const uint32_t* const* ahats) {
  FftPassArgs a;
  a.X = b.X;
  a.M = M;
  a.lo = lo;
  a.R = R;
  a.C = C;
  a.lgC = __builtin_ctz(C);
  a.mode = mode;
  a.khi = khi;
  a.first = first;
  a.last = last;
  a.ahat = ahat1;
  a.ahats = ahats;
  a.XS_words = b.XS * 4;
  a.S0 = nullptr;
  a.D0 = nullptr;
  a.gin = 0;
  a.gout = 0;
  a.Xo = nullptr;
  return a;
}

template <int NP>
void launch_pass(const FftPassArgs& a, int B, hipStream_t st) {
  constexpr int Z = 32 / NP;  // halves per element handled by separate workgroups
  unsigned g;
  pass_grid(a.M / ((1 << a.R) * a.C), B, Z, g);
  // algorithmic bytes: every element's 32 planes read and written once (the
  // first pass reads only planes 0..15 of the lower half, once per half)
  const double per = (double)a.M * 512.0 * B;
  cuot_rec_begin(st, "qc_fft_pass", a.first ? per * (NP == 16 ? 1.5 : 1.25) : per * 2.0);
  hipLaunchKernelGGL(qc_fft_pass<NP>, dim3(g, (unsigned)B, (unsigned)Z), dim3(fthreads<NP>()),
                     ((size_t)(1 << a.R) * a.C * fstride<NP>() + CTAB_WORDS) * 4, st, a);
  cuot_rec_end(st);
}

// Shape of an upper pass over layers [lo, lo + R): half mode when every layer
// has its twiddles in GF(2^16) (lo >= m - 16); R layers (knob qc_radix, else 7
// in half mode and 5 otherwise), C columns (knob qc_cols, else as many as the
// tile holds, at most the column block 2^lo).
struct PassShape {
  int lo, R, C;
  bool half;
};

PassShape upper_shape(int m, int kfrom, int kto, bool down) {
  // kfrom..kto: remaining layers (down: kfrom >= kto, taken from the top)
  const int nrem = down ? kfrom - kto + 1 : kto - kfrom + 1;
  const int knobR = cuot_knobs().qc_radix;
  PassShape ps;
  for (int pass_half = 1; pass_half >= 0; pass_half--) {
    int R = knobR > 0 ? knobR : (pass_half ? 7 : 5);
    if (R > nrem) R = nrem;
    const int lo = down ? kfrom - R + 1 : kfrom;
    if (pass_half && lo < m - 16) continue;
    const int tcmax = pass_half ? ftc_max<16>() : ftc_max<32>();
    while ((1 << R) > tcmax) R--;
    ps.R = R;
    ps.lo = down ? kfrom - R + 1 : kfrom;
    ps.half = pass_half != 0;
    int C = cuot_knobs().qc_cols > 0 ? cuot_knobs().qc_cols : FFT_TC;
    while (C > 1 && (((int64_t)1 << ps.lo) % C)) C >>= 1;
    while (((1 << R) * C) > tcmax) C >>= 1;
    ps.C = C;
    return ps;
  }
  return ps;
}

// unpack of the chunk products after a last inverse pass in half mode: the 32
// coordinates of every element word become its 31 coefficient planes (plane 31
// zero).  Plane p of instance i is words [(i * 32 + p) * 4M, ... + 4M).
// G: the source holds the planes grouped by four (block ((p/4)*M + e)*4 + p%4),
// read from src and written plane-major to dst (a different buffer); else
// src == dst, plane-major, in place.
template <bool G>
__global__ __launch_bounds__(256) void qc_unpack_x(const uint32_t* src, uint32_t* dst, int64_t M, int64_t XS_words) {
  const uint32_t* Si = src + (int64_t)blockIdx.y * XS_words;
  uint32_t* Di = dst + (int64_t)blockIdx.y * XS_words;
  const int64_t pw = M * 4;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < pw; t += (int64_t)gridDim.x * blockDim.x) {
    uint32_t in[32], out[31];
#pragma unroll
    for (int p = 0; p < 32; p++)
      in[p] = G ? Si[((((int64_t)(p >> 2)) * M + (t >> 2)) * 4 + (p & 3)) * 4 + (t & 3)] : Si[p * pw + t];
    gf32_unpack_planes(in, out);
#pragma unroll
    for (int p = 0; p < 31; p++) Di[p * pw + t] = out[p];
    Di[31 * pw + t] = 0u;
  }
}

// forward FFT layers m-1..0 (top layer from the packed lower half); with A^,
// the lowest pass also multiplies by A^ and runs the inverse layers of its
// range.  Returns the highest layer of the lowest pass in *low_khi.
// The passes exchange X in the grouped layout when there are upper passes and
// the first is in half mode: that pass reads the plane-major chunk planes
// 0..15 while writing half 0 to the row buffer and half 1 to planes 16..31,
// so no write lands on a block another tile still reads.  (A first pass in
// full mode writes over its own input in place and keeps the plane-major
// layout throughout.)  The last pass then writes plane-major to the idle row
// buffer, and the caller swaps the two buffers' roles.
bool fft_grouped(int m) {
  const int R0 = m < FFT_LOGTC ? m : FFT_LOGTC;
  if (m - 1 < R0) return false;  // one fused pass, in place
  return upper_shape(m, m - 1, R0, true).half;
}

hipError_t fft_forward(const Bufs& b, int m, const uint32_t* ahat1, const uint32_t* const* ahats, bool fused,
                       bool grouped, int* low_khi, hipStream_t st) {
  const int64_t M = (int64_t)1 << m;
  const int R0 = m < FFT_LOGTC ? m : FFT_LOGTC;
  int khi = m - 1;
  bool first = true;
  bool h0_in_rows = false;  // half 0 of the previous (first, half-mode) pass sits in the row buffer
  while (khi >= R0) {
    const PassShape ps = upper_shape(m, khi, R0, true);
    FftPassArgs a = pass_args(b, M, ps.lo, ps.R, ps.C, PASS_FWD, khi, first ? 1 : 0, 0, nullptr, nullptr);
    a.gin = (grouped && !first) ? 1 : 0;
    a.gout = grouped ? 1 : 0;
    if (h0_in_rows) a.S0 = b.Bw;
    h0_in_rows = first && ps.half;
    if (h0_in_rows) a.D0 = b.Bw;
    if (ps.half)
      launch_pass<16>(a, b.B, st);
    else
      launch_pass<32>(a, b.B, st);
    first = false;
    khi = ps.lo - 1;
  }
  // the fused pass completes the inverse transform when no upper layers exist
  FftPassArgs a = pass_args(b, M, 0, R0, 1, fused ? PASS_FUSED : PASS_FWD, R0 - 1, first ? 1 : 0,
                            (fused && R0 == m) ? 1 : 0, ahat1, ahats);
  a.gin = a.gout = grouped ? 1 : 0;  // grouped implies upper passes before and after this one
  if (h0_in_rows) a.S0 = b.Bw;
  launch_pass<32>(a, b.B, st);
  *low_khi = R0 - 1;
  return hipGetLastError();
}

// grouped: the last pass writes to the row buffer (idle since the pass after
// the first).  In half mode it writes the grouped form, and the unpack reads
// it from there and writes the plane-major coefficient planes back to X; in
// full mode it unpacks in the tile and writes plane-major, and *swapped tells
// the caller that the result is in the row buffer.
hipError_t fft_inverse_upper(const Bufs& b, int m, int from, bool grouped, bool* swapped, hipStream_t st) {
  *swapped = false;
  const int64_t M = (int64_t)1 << m;
  int klo = from;
  while (klo <= m - 1) {
    const PassShape ps = upper_shape(m, klo, m - 1, false);
    const bool last = ps.lo + ps.R == m;
    FftPassArgs a = pass_args(b, M, ps.lo, ps.R, ps.C, PASS_INV, ps.lo + ps.R - 1, 0, (last && !ps.half) ? 1 : 0,
                              nullptr, nullptr);
    a.gin = grouped ? 1 : 0;
    a.gout = (grouped && (!last || ps.half)) ? 1 : 0;
    if (grouped && last) a.Xo = b.Bw;
    if (ps.half) {
      launch_pass<16>(a, b.B, st);
      if (last) {
        cuot_rec_begin(st, "qc_unpack_x", 2.0 * 512.0 * (double)M * b.B);
        if (grouped)
          hipLaunchKernelGGL(qc_unpack_x<true>, dim3(capg(M * 4 / 256, 8192), (unsigned)b.B), dim3(256), 0, st,
                             b.Bw, b.X, M, b.XS * 4);
        else
          hipLaunchKernelGGL(qc_unpack_x<false>, dim3(capg(M * 4 / 256, 8192), (unsigned)b.B), dim3(256), 0, st,
                             b.X, b.X, M, b.XS * 4);
        cuot_rec_end(st);
      }
    } else {
      launch_pass<32>(a, b.B, st);
      if (grouped && last) *swapped = true;
    }
    klo += ps.R;
  }
  return hipGetLastError();
}

hipError_t encode_batch(const QcIo* io, QcIo io1, const uint32_t* ahat1, const uint32_t* const* ahats, int B, int m,
                        void* scratch, hipStream_t st) {
  hipError_t e = configure_lds();
  if (e != hipSuccess) return e;
  if (B > 1 && !row_path(m - 1)) return hipErrorInvalidValue;
  const int64_t M = (int64_t)1 << m;
  Bufs b = bufs(scratch, M, B);
  cuot_rec_begin(st, "qc_pack_x", 2.0 * 16.0 * (double)(M >> 1) * 16 * B);
    hipLaunchKernelGGL(qc_pack_x, dim3(grid_for(16 * (M >> 1)), (unsigned)B), dim3(256), 0, st, io, io1, (uint4*)b.X, M,
                     b.XS);
  cuot_rec_end(st);
  if ((e = conv_forward(b, M, 16, m - 1, st))) return e;
  int khi;
  const bool grouped = fft_grouped(m);
  if ((e = fft_forward(b, m, ahat1, ahats, true, grouped, &khi, st))) return e;
  bool swapped;
  if ((e = fft_inverse_upper(b, m, khi + 1, grouped, &swapped, st))) return e;
  if (swapped) {
    // the transform's result sits in the row buffer: the two buffers swap
    // roles for the inverse conversion and the output
    uint32_t* t = b.X;
    b.X = b.Bw;
    b.Bw = t;
  }
  // n = 16 * M/2 (n a power of two): fold in the row layout and write y from it
  const bool fold = !io && io1.n == 8 * M && m >= 11 && row_path(m - 1) && B == 1;
  if ((e = conv_inverse(b, M, 31, m - 1, fold, st))) return e;
  if (fold) {
    const int64_t h = M >> 1, hw = h / 32;
    cuot_rec_begin(st, "qc_rows_fold", 31.0 * 16.0 * (double)M + 16.0 * 16.0 * (double)h);
    hipLaunchKernelGGL(qc_rows_fold, dim3(capg(hw / 32, 1024), 16), dim3(256), 0, st, (const uint32_t*)b.Bw, hw,
                       b.XS * 4, (uint4*)b.X, M);
    cuot_rec_end(st);
    cuot_rec_begin(st, "qc_output_planes", 48.0 * (double)io1.n);
    hipLaunchKernelGGL(qc_output_planes, dim3(capg(h / 64, 8192), 1), dim3(256), 0, st, (const uint4*)b.X, M, h, io,
                       io1);
    cuot_rec_end(st);
    return hipGetLastError();
  }
  const int64_t nmax = io ? (M / 2) * 16 : io1.n;  // grid covers the largest n of the class
  cuot_rec_begin(st, "qc_output", 32.0 * (double)nmax * B);
    hipLaunchKernelGGL(qc_output, dim3(grid_for(nmax), (unsigned)B), dim3(256), 0, st, (const uint4*)b.X, M, m, b.XS,
                     io, io1);
  cuot_rec_end(st);
  return hipGetLastError();
}

hipError_t prepare_batch(const QcPrep* pr, QcPrep pr1, int B, int m, hipStream_t st) {
  hipError_t e = configure_lds();
  if (e != hipSuccess) return e;
  const int64_t M = (int64_t)1 << m;
  const int64_t Mh = M >> 1;
  // several polynomials: one workgroup each; a single one keeps the
  // grid-wide launches (one workgroup would be its critical path)
  if (m <= PREP_LDS_MMAX && B > 1) {
    hipLaunchKernelGGL(qcs_prepare_lds, dim3((unsigned)B), dim3(1024), (size_t)M * 4, st, pr, pr1, m);
    return hipGetLastError();
  }
  const dim3 blk(256);
  hipLaunchKernelGGL(qcs_pack, dim3(grid_for(M), (unsigned)B), blk, 0, st, pr, pr1, M);
  if (m - 1 >= QCS_LB && m - QCS_LB <= QCS_LB) {
    // upper conversion levels per launch, the rest in tiles; the FFT in two
    // tiled launches (layers m-1..QCS_LB strided, QCS_LB-1..0 contiguous)
    for (int k = m - 2; k >= QCS_LB; k--) {
      hipLaunchKernelGGL(qcs_hi_fix, dim3(grid_for(Mh >> 1, 4096), (unsigned)B), blk, 0, st, pr, pr1, Mh, k);
      hipLaunchKernelGGL(qcs_lo_update, dim3(grid_for(Mh >> 1, 4096), (unsigned)B), blk, 0, st, pr, pr1, Mh, k);
    }
    // 1024 lanes per 4096-word tile: the butterflies' scalar multiplies are
    // serial chains, so the tiles need many waves to hide them
    const dim3 big(1024);
    hipLaunchKernelGGL(qcs_conv_tile, dim3((unsigned)(Mh >> QCS_LB), (unsigned)B), big, 0, st, pr, pr1,
                       QCS_LB - 1);
    const int Ru = m - QCS_LB;
    hipLaunchKernelGGL(qcs_fft_tile, dim3((unsigned)(M >> QCS_LB), (unsigned)B), big, 0, st, pr, pr1, QCS_LB, Ru,
                       QCS_LB - Ru, m - 1);
    hipLaunchKernelGGL(qcs_fft_tile, dim3((unsigned)(M >> QCS_LB), (unsigned)B), big, 0, st, pr, pr1, 0, QCS_LB, 0,
                       QCS_LB - 1);
    return hipGetLastError();
  }
  for (int k = m - 2; k >= 0; k--) {
    hipLaunchKernelGGL(qcs_hi_fix, dim3(grid_for(Mh >> 1, 4096), (unsigned)B), blk, 0, st, pr, pr1, Mh, k);
    hipLaunchKernelGGL(qcs_lo_update, dim3(grid_for(Mh >> 1, 4096), (unsigned)B), blk, 0, st, pr, pr1, Mh, k);
  }
  // top layer (constant 0, empty upper half): duplicate the lower half
  hipLaunchKernelGGL(qcs_fft_layer, dim3(grid_for(Mh, 4096), (unsigned)B), blk, 0, st, pr, pr1, M, m - 1);
  for (int k = m - 2; k >= 0; k--)
    hipLaunchKernelGGL(qcs_fft_layer, dim3(grid_for(Mh, 4096), (unsigned)B), blk, 0, st, pr, pr1, M, k);
  return hipGetLastError();
}

}  // namespace

int cuot_qc_log_len(int64_t n) {
  const int64_t nc = (n + 15) / 16;
  int m = 1;
  while (((int64_t)1 << m) < 2 * nc - 1) m++;
  return m;
}

size_t cuot_qc_scratch_bytes_for(int64_t n) { return cuot_qc_batch_scratch_bytes(cuot_qc_log_len(n), 1); }

bool cuot_qc_batchable(int m) { return row_path(m - 1); }

hipError_t cuot_qc_configure() { return configure_lds(); }

size_t cuot_qc_batch_scratch_bytes(int m, int B) {
  return ((size_t)B * 64 + CONV_G) * ((size_t)1 << m) * 16;
}

hipError_t cuot_launch_qc_prepare(const uint8_t* a_bits, int64_t n, uint32_t* ahat, void* scratch, hipStream_t st) {
  (void)scratch;
  QcPrep p1;
  p1.a_bits = a_bits;
  p1.ahat = ahat;
  p1.n = n;
  const int m = cuot_qc_log_len(n);
  hipError_t e = prepare_batch(nullptr, p1, 1, m, st);
  if (e == hipSuccess && cuot_qc_mfma_eligible(n)) e = cuot_launch_qc_keep_bits(nullptr, p1, 1, m, n, st);
  return e;
}

hipError_t cuot_launch_qc_prepare_batch(const QcPrep* prep_d, int B, int m, hipStream_t st) {
  QcPrep p1;
  memset(&p1, 0, sizeof(p1));
  hipError_t e = prepare_batch(prep_d, p1, B, m, st);
  // members small enough for the matrix-core encoder also keep a's bits
  const int64_t nlo = ((int64_t)1 << (m + 2)) + 1, nhi = (int64_t)1 << (m + 3);
  if (e == hipSuccess && cuot_qc_mfma_eligible(nlo))
    e = cuot_launch_qc_keep_bits(prep_d, p1, B, m, cuot_qc_mfma_eligible(nhi) ? nhi : nlo, st);
  return e;
}

hipError_t cuot_launch_qc_encode(const uint32_t* ahat, int64_t n, const void* x, void* y, void* scratch,
                                 hipStream_t st) {
  QcIo io1;
  io1.x = (const uint4*)x;
  io1.y = (uint4*)y;
  io1.n = n;
  const int m = cuot_qc_log_len(n);
  if (cuot_qc_mfma_selected(n))
    return cuot_launch_qc_mfma(nullptr, io1, ahat + ((int64_t)1 << m), nullptr, 1, m, n, scratch, st);
  return encode_batch(nullptr, io1, ahat, nullptr, 1, m, scratch, st);
}

hipError_t cuot_launch_qc_encode_batch(const QcIo* io_d, const uint32_t* const* ahats_d, int B, int m, int64_t nmax,
                                       void* scratch, hipStream_t st) {
  QcIo io1;
  memset(&io1, 0, sizeof(io1));
  if (m >= 2 && cuot_qc_mfma_selected(nmax))
    return cuot_launch_qc_mfma(io_d, io1, nullptr, ahats_d, B, m, nmax, scratch, st);
  return encode_batch(io_d, io1, nullptr, ahats_d, B, m, scratch, st);
}
