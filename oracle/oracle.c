/*
 * oracle.c -- CPU restatement of the silent-OT correlation path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See oracle.h for the pinning
 * status of each function.  Reference citations are relative to the
 * reference repository root.
 */
#include "oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

#if defined(__x86_64__)
#include <immintrin.h>
#endif

/* ======================================================================== */
/* AES-128, FIPS-197.  S-box computed from its definition (multiplicative   */
/* inverse in GF(2^8) mod x^8+x^4+x^3+x+1 followed by the affine map).       */
/* ======================================================================== */

static uint8_t SBOX[256];

static uint8_t gf8_mul(uint8_t a, uint8_t b) {
  uint8_t r = 0;
  while (b) {
    if (b & 1) r ^= a;
    a = (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1b : 0));
    b >>= 1;
  }
  return r;
}

static uint8_t rotl8(uint8_t x, int s) { return (uint8_t)((x << s) | (x >> (8 - s))); }

__attribute__((constructor)) static void or_init_sbox(void) {
  for (int x = 0; x < 256; x++) {
    uint8_t inv = 0;
    if (x) {
      /* x^254 = x^-1 */
      uint8_t p = (uint8_t)x, r = 1;
      int e = 254;
      while (e) {
        if (e & 1) r = gf8_mul(r, p);
        p = gf8_mul(p, p);
        e >>= 1;
      }
      inv = r;
    }
    SBOX[x] = (uint8_t)(inv ^ rotl8(inv, 1) ^ rotl8(inv, 2) ^ rotl8(inv, 3) ^ rotl8(inv, 4) ^ 0x63);
  }
}

/* FIPS-197 sec. 5.2 key expansion.  (The reference's gpu/aes_expand.cu:59
 * indexes Rcon one step late; this is the standard schedule.) */
void or_aes_expand(const orb *key, uint8_t rk[176]) {
  memcpy(rk, key, 16);
  uint8_t rcon = 1;
  for (int i = 4; i < 44; i++) {
    uint8_t t[4];
    memcpy(t, rk + 4 * (i - 1), 4);
    if (i % 4 == 0) {
      uint8_t t0 = t[0];
      t[0] = (uint8_t)(SBOX[t[1]] ^ rcon);
      t[1] = SBOX[t[2]];
      t[2] = SBOX[t[3]];
      t[3] = SBOX[t0];
      rcon = gf8_mul(rcon, 2);
    }
    for (int j = 0; j < 4; j++) rk[4 * i + j] = (uint8_t)(rk[4 * (i - 4) + j] ^ t[j]);
  }
}

static void aes_block_portable(const uint8_t rk[176], const uint8_t in[16], uint8_t out[16]) {
  uint8_t s[16], u[16];
  for (int i = 0; i < 16; i++) s[i] = (uint8_t)(in[i] ^ rk[i]);
  for (int r = 1; r <= 10; r++) {
    /* SubBytes + ShiftRows: state byte (row, col) sits at index row + 4*col */
    for (int c = 0; c < 4; c++)
      for (int row = 0; row < 4; row++) u[row + 4 * c] = SBOX[s[row + 4 * ((c + row) & 3)]];
    if (r < 10) {
      for (int c = 0; c < 4; c++) {
        uint8_t *col = u + 4 * c;
        uint8_t a0 = col[0], a1 = col[1], a2 = col[2], a3 = col[3];
        s[4 * c + 0] = (uint8_t)(gf8_mul(a0, 2) ^ gf8_mul(a1, 3) ^ a2 ^ a3);
        s[4 * c + 1] = (uint8_t)(a0 ^ gf8_mul(a1, 2) ^ gf8_mul(a2, 3) ^ a3);
        s[4 * c + 2] = (uint8_t)(a0 ^ a1 ^ gf8_mul(a2, 2) ^ gf8_mul(a3, 3));
        s[4 * c + 3] = (uint8_t)(gf8_mul(a0, 3) ^ a1 ^ a2 ^ gf8_mul(a3, 2));
      }
    } else {
      memcpy(s, u, 16);
    }
    for (int i = 0; i < 16; i++) s[i] ^= rk[16 * r + i];
  }
  memcpy(out, s, 16);
}

int or_have_aesni(void) {
#if defined(__x86_64__)
  __builtin_cpu_init();
  return __builtin_cpu_supports("aes");
#else
  return 0;
#endif
}

#if defined(__x86_64__)
__attribute__((target("aes,sse4.1"))) static void aes_ecb_ni(const uint8_t rk[176], const orb *in,
                                                             orb *out, size_t n) {
  __m128i k[11];
  for (int r = 0; r < 11; r++) k[r] = _mm_loadu_si128((const __m128i *)(rk + 16 * r));
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    __m128i x[8];
    for (int j = 0; j < 8; j++) x[j] = _mm_xor_si128(_mm_loadu_si128((const __m128i *)&in[i + j]), k[0]);
    for (int r = 1; r < 10; r++)
      for (int j = 0; j < 8; j++) x[j] = _mm_aesenc_si128(x[j], k[r]);
    for (int j = 0; j < 8; j++) _mm_storeu_si128((__m128i *)&out[i + j], _mm_aesenclast_si128(x[j], k[10]));
  }
  for (; i < n; i++) {
    __m128i x = _mm_xor_si128(_mm_loadu_si128((const __m128i *)&in[i]), k[0]);
    for (int r = 1; r < 10; r++) x = _mm_aesenc_si128(x, k[r]);
    _mm_storeu_si128((__m128i *)&out[i], _mm_aesenclast_si128(x, k[10]));
  }
}
#endif

static int g_aesni = -1;

void or_aes_ecb(const uint8_t rk[176], const orb *in, orb *out, size_t n, int impl) {
  if (g_aesni < 0) g_aesni = or_have_aesni();
#if defined(__x86_64__)
  if (impl == 1 && g_aesni) {
    aes_ecb_ni(rk, in, out, n);
    return;
  }
#endif
  for (size_t i = 0; i < n; i++) aes_block_portable(rk, (const uint8_t *)&in[i], (uint8_t *)&out[i]);
}

#define AES_FAST 1

/* ======================================================================== */
/* Seed derivation: AES-CTR keyed by the 128-bit seed, counter block        */
/* {lo = index, hi = stream}.  (Workload definition; the reference draws    */
/* from rand(), silent/sender.cu:110-119, which is not reproducible.)       */
/* ======================================================================== */
void or_prg(const orb *seed, uint64_t stream, uint64_t start, size_t count, orb *out) {
  uint8_t rk[176];
  or_aes_expand(seed, rk);
  for (size_t i = 0; i < count; i++) {
    out[i].lo = start + i;
    out[i].hi = stream;
  }
  or_aes_ecb(rk, out, out, count, AES_FAST);
}

/* ======================================================================== */
/* GGM tree with the two-key PRG G(s) = AES_k0(s)^s || AES_k1(s)^s          */
/* (ferret/emp-ot/emp-ot/ferret/twokeyprp.h:7-63: ParaEnc<2,N> encrypts     */
/* blocks [0,N) under key 0 and [N,2N) under key 1, children[2p] = left).   */
/* ======================================================================== */

static inline orb bx(orb a, orb b) {
  orb r = {a.lo ^ b.lo, a.hi ^ b.hi};
  return r;
}

/* expand level in place: nodes[0..w) -> nodes[0..2w), iterating from the
 * top as node_expand_4to8 does (spcot_sender.h:70-84). */
static void ggm_level_inplace(const uint8_t rk0[176], const uint8_t rk1[176], orb *nodes, int64_t w) {
  orb p[8], l[8], r[8];
  int64_t i = w;
  while (i > 0) {
    int64_t cnt = i >= 8 ? 8 : i;
    int64_t base = i - cnt;
    memcpy(p, nodes + base, (size_t)cnt * sizeof(orb));
    or_aes_ecb(rk0, p, l, (size_t)cnt, AES_FAST);
    or_aes_ecb(rk1, p, r, (size_t)cnt, AES_FAST);
    for (int64_t j = cnt - 1; j >= 0; j--) {
      nodes[2 * (base + j)] = bx(l[j], p[j]);
      nodes[2 * (base + j) + 1] = bx(r[j], p[j]);
    }
    i = base;
  }
}

static void level_sums(const orb *nodes, int64_t w, orb *ls, orb *rs) {
  orb L = {0, 0}, R = {0, 0};
  for (int64_t i = 0; i < w; i += 2) {
    L = bx(L, nodes[i]);
    R = bx(R, nodes[i + 1]);
  }
  *ls = L;
  *rs = R;
}

void or_ggm_sender_tree(const uint8_t rk0[176], const uint8_t rk1[176], orb root, int h, orb *leaves,
                        orb *lsum, orb *rsum) {
  leaves[0] = root;
  for (int d = 0; d < h; d++) {
    ggm_level_inplace(rk0, rk1, leaves, (int64_t)1 << d);
    level_sums(leaves, (int64_t)2 << d, &lsum[d], &rsum[d]);
  }
}

static const orb ORB_ZERO = {0, 0};
static const orb ORB_LSB_CLEAR = {0xFFFFFFFFFFFFFFFEull, 0xFFFFFFFFFFFFFFFFull};

static void ferret_keys(uint8_t rk0[176], uint8_t rk1[176]) {
  /* TwoKeyPRP(zero_block, makeBlock(0, 1))  (spcot_sender.h:63) */
  orb k0 = {0, 0}, k1 = {1, 0};
  or_aes_expand(&k0, rk0);
  or_aes_expand(&k1, rk1);
}

/* spcot_sender.h:49-86 (ggm_tree_gen) per tree; mpcot_reg.h:130-144
 * concatenates trees; leaves beyond n_out are dropped (truncation). */
void or_ferret_mpcot_sender(const orb *roots, int64_t t, int h, orb delta, int64_t n_out, orb *leaves,
                            orb *lsum, orb *rsum, orb *secret_sum, int nthreads) {
  uint8_t rk0[176], rk1[176];
  ferret_keys(rk0, rk1);
  const int64_t W = (int64_t)1 << h;
#pragma omp parallel num_threads(nthreads)
  {
    orb *tree = (orb *)malloc((size_t)W * sizeof(orb));
#pragma omp for schedule(dynamic)
    for (int64_t tr = 0; tr < t; tr++) {
      or_ggm_sender_tree(rk0, rk1, roots[tr], h, tree, lsum + tr * h, rsum + tr * h);
      orb s = ORB_ZERO;
      for (int64_t i = 0; i < W; i++) {
        tree[i].lo &= ORB_LSB_CLEAR.lo;
        s = bx(s, tree[i]);
      }
      secret_sum[tr] = bx(s, delta);
      int64_t base = tr * W;
      int64_t cnt = n_out - base;
      if (cnt > W) cnt = W;
      if (cnt > 0) memcpy(leaves + base, tree, (size_t)cnt * sizeof(orb));
    }
    free(tree);
  }
}

/* spcot_recver.h:53-95 (compute, ggm_tree_reconstruction, layer_recover). */
void or_ferret_mpcot_recver(int64_t t, int h, const uint8_t *b, const orb *csum, const orb *secret_sum,
                            int64_t n_out, orb *leaves, int nthreads) {
  uint8_t rk0[176], rk1[176];
  ferret_keys(rk0, rk1);
  const int64_t W = (int64_t)1 << h;
#pragma omp parallel num_threads(nthreads)
  {
    orb *tree = (orb *)malloc((size_t)W * sizeof(orb));
#pragma omp for schedule(dynamic)
    for (int64_t tr = 0; tr < t; tr++) {
      const uint8_t *bb = b + tr * h;
      const orb *m = csum + tr * h;
      int64_t to_fill = 0;
      for (int i = 1; i <= h; i++) {
        to_fill *= 2;
        tree[to_fill] = tree[to_fill + 1] = ORB_ZERO;
        int lr = bb[i - 1] ? 1 : 0;
        int64_t fill = lr ? to_fill + 1 : to_fill;
        /* layer_recover */
        int64_t item_n = (int64_t)1 << i;
        orb s = ORB_ZERO;
        for (int64_t j = lr; j < item_n; j += 2) s = bx(s, tree[j]);
        tree[fill] = bx(s, m[i - 1]);
        if (i != h) ggm_level_inplace(rk0, rk1, tree, item_n);
        if (!lr) to_fill += 1;
      }
      /* get_index (spcot_recver.h:33-41) */
      int64_t pos = 0;
      for (int i = 0; i < h; i++) pos = (pos << 1) + (bb[i] ? 0 : 1);
      tree[pos] = ORB_ZERO;
      orb s = ORB_ZERO;
      for (int64_t i = 0; i < W; i++) {
        tree[i].lo &= ORB_LSB_CLEAR.lo;
        s = bx(s, tree[i]);
      }
      tree[pos] = bx(s, secret_sum[tr]);
      int64_t base = tr * W;
      int64_t cnt = n_out - base;
      if (cnt > W) cnt = W;
      if (cnt > 0) memcpy(leaves + base, tree, (size_t)cnt * sizeof(orb));
    }
    free(tree);
  }
}

/* silent/sender.cu:76-108: masks base-OT rows with the level sums; the
 * extra row h is base row h-1 xor the opposite sum xor Delta. */
void or_sot_sender_expand(const uint8_t rk0[176], const uint8_t rk1[176], const orb *roots, int64_t t,
                          int h, orb delta, const orb *m0b, const orb *m1b, int64_t n_leaves, orb *leaves,
                          orb *M0, orb *M1, int nthreads) {
  const int64_t W = (int64_t)1 << h;
#pragma omp parallel num_threads(nthreads)
  {
    orb *tree = (orb *)malloc((size_t)W * sizeof(orb));
    orb *ls = (orb *)malloc((size_t)h * sizeof(orb));
    orb *rs = (orb *)malloc((size_t)h * sizeof(orb));
#pragma omp for schedule(dynamic)
    for (int64_t tr = 0; tr < t; tr++) {
      or_ggm_sender_tree(rk0, rk1, roots[tr], h, tree, ls, rs);
      for (int d = 0; d < h; d++) {
        M0[d * t + tr] = bx(m0b[d * t + tr], ls[d]);
        M1[d * t + tr] = bx(m1b[d * t + tr], rs[d]);
      }
      M0[h * t + tr] = bx(bx(m0b[(h - 1) * t + tr], rs[h - 1]), delta);
      M1[h * t + tr] = bx(bx(m1b[(h - 1) * t + tr], ls[h - 1]), delta);
      int64_t base = tr * W;
      int64_t cnt = n_leaves - base;
      if (cnt > W) cnt = W;
      if (cnt > 0) memcpy(leaves + base, tree, (size_t)cnt * sizeof(orb));
    }
    free(tree);
    free(ls);
    free(rs);
  }
}

/* silent/recver.cu:98-163 (fill_tree + seed_expand), with the receiver's
 * own unknown root taken as zero: the garbage on the punctured path cancels
 * in the fill rule val ^= fullSum ^ ownSideSum. */
void or_sot_recver_expand(const uint8_t rk0[176], const uint8_t rk1[176], int64_t t, int h,
                          const uint8_t *c, const orb *M0, const orb *M1, const orb *m0b, const orb *m1b,
                          int64_t n_leaves, orb *leaves, int nthreads) {
  const int64_t W = (int64_t)1 << h;
#pragma omp parallel num_threads(nthreads)
  {
    orb *tree = (orb *)malloc((size_t)W * sizeof(orb));
#pragma omp for schedule(dynamic)
    for (int64_t tr = 0; tr < t; tr++) {
      int64_t active = 0;
      tree[0] = ORB_ZERO;
      for (int d = 0; d < h; d++) {
        int64_t w = (int64_t)1 << d;
        ggm_level_inplace(rk0, rk1, tree, w);
        orb own[2];
        level_sums(tree, 2 * w, &own[0], &own[1]);
        int cc = c[d * t + tr];
        orb mc = cc ? m1b[d * t + tr] : m0b[d * t + tr];
        orb full = bx(cc ? M1[d * t + tr] : M0[d * t + tr], mc);
        int64_t fill = 2 * active + cc;
        tree[fill] = bx(bx(tree[fill], full), own[cc]);
        active = 2 * active + (1 - cc);
        if (d == h - 1) {
          orb full2 = bx(cc ? M1[h * t + tr] : M0[h * t + tr], mc);
          tree[active] = bx(bx(tree[active], full2), own[1 - cc]);
        }
      }
      int64_t base = tr * W;
      int64_t cnt = n_leaves - base;
      if (cnt > W) cnt = W;
      if (cnt > 0) memcpy(leaves + base, tree, (size_t)cnt * sizeof(orb));
    }
    free(tree);
  }
}

/* ======================================================================== */
/* Quasi-cyclic encoder over GF(2), exact.                                   */
/* ======================================================================== */

static inline int getbit(const uint8_t *v, int64_t i) { return (v[i >> 3] >> (i & 7)) & 1; }
static inline void flipbit(uint8_t *v, int64_t i) { v[i >> 3] ^= (uint8_t)(1u << (i & 7)); }

void or_qc_encode_brute(const orb *x, const uint8_t *a, int64_t n, orb *y) {
  for (int64_t i = 0; i < n; i++) y[i] = x[i];
  for (int64_t j = 0; j < n; j++) {
    if (!getbit(a, j)) continue;
    /* y_i ^= x_{n + ((i - j) mod n)} */
    for (int64_t i = 0; i < n; i++) {
      int64_t s = i - j;
      if (s < 0) s += n;
      y[i] = bx(y[i], x[n + s]);
    }
  }
}

void or_qc_encode_bits_brute(const uint8_t *e, const uint8_t *a, int64_t n, uint8_t *out) {
  memset(out, 0, (size_t)((n + 7) / 8));
  for (int64_t i = 0; i < n; i++)
    if (getbit(e, i)) flipbit(out, i);
  for (int64_t j = 0; j < n; j++) {
    if (!getbit(a, j)) continue;
    for (int64_t i = 0; i < n; i++) {
      int64_t s = i - j;
      if (s < 0) s += n;
      if (getbit(e, n + s)) flipbit(out, i);
    }
  }
}

void or_qc_sparse(const int64_t *pos, int64_t w, const uint8_t *a, int64_t n, uint8_t *out) {
  memset(out, 0, (size_t)((n + 7) / 8));
  for (int64_t q = 0; q < w; q++) {
    int64_t p = pos[q];
    if (p < 0 || p >= 2 * n) continue;
    if (p < n) {
      flipbit(out, p);
      continue;
    }
    int64_t s = p - n; /* out_i ^= a_{(i - s) mod n} */
    for (int64_t i = 0; i < n; i++) {
      int64_t j = i - s;
      if (j < 0) j += n;
      if (getbit(a, j)) flipbit(out, i);
    }
  }
}

/* ---- carry-less products ------------------------------------------------ */
static inline void clmul_soft(uint64_t a, uint64_t b, uint64_t *lo, uint64_t *hi) {
  uint64_t l = 0, h = 0;
  for (int i = 0; i < 64; i++) {
    if ((b >> i) & 1) {
      l ^= a << i;
      if (i) h ^= a >> (64 - i);
    }
  }
  *lo = l;
  *hi = h;
}

static int g_clmul = -1;

#if defined(__x86_64__)
__attribute__((target("pclmul,sse4.1"))) static void school_clmul(const uint64_t *a, const uint64_t *b,
                                                                  int64_t m, uint64_t *out) {
  memset(out, 0, (size_t)(2 * m) * 8);
  for (int64_t i = 0; i < m; i++) {
    __m128i ai = _mm_set_epi64x(0, (long long)a[i]);
    for (int64_t j = 0; j < m; j++) {
      __m128i p = _mm_clmulepi64_si128(ai, _mm_set_epi64x(0, (long long)b[j]), 0x00);
      out[i + j] ^= (uint64_t)_mm_cvtsi128_si64(p);
      out[i + j + 1] ^= (uint64_t)_mm_extract_epi64(p, 1);
    }
  }
}
#endif

static void school_soft(const uint64_t *a, const uint64_t *b, int64_t m, uint64_t *out) {
  memset(out, 0, (size_t)(2 * m) * 8);
  for (int64_t i = 0; i < m; i++)
    for (int64_t j = 0; j < m; j++) {
      uint64_t lo, hi;
      clmul_soft(a[i], b[j], &lo, &hi);
      out[i + j] ^= lo;
      out[i + j + 1] ^= hi;
    }
}

static void bpmul_rec(const uint64_t *a, const uint64_t *b, int64_t m, uint64_t *out, uint64_t *scr) {
  if (m <= 16) {
#if defined(__x86_64__)
    if (g_clmul) {
      school_clmul(a, b, m, out);
      return;
    }
#endif
    school_soft(a, b, m, out);
    return;
  }
  int64_t hh = (m + 1) / 2, m1 = m - hh;
  /* scratch layout: sa[hh] sb[hh] p1[2hh] then child scratch */
  uint64_t *sa = scr, *sb = scr + hh, *p1 = scr + 2 * hh, *child = scr + 4 * hh;
  for (int64_t i = 0; i < hh; i++) {
    sa[i] = a[i] ^ (i < m1 ? a[hh + i] : 0);
    sb[i] = b[i] ^ (i < m1 ? b[hh + i] : 0);
  }
  bpmul_rec(sa, sb, hh, p1, child);
  /* p0 -> out[0, 2hh), p2 -> out[2hh, 2hh + 2m1) */
  bpmul_rec(a, b, hh, out, child);
  bpmul_rec(a + hh, b + hh, m1, out + 2 * hh, child);
  /* middle = p1 ^ p0 ^ p2, added at word offset hh */
  for (int64_t i = 0; i < 2 * hh; i++) {
    uint64_t v = p1[i] ^ out[i] ^ (i < 2 * m1 ? out[2 * hh + i] : 0);
    p1[i] = v;
  }
  for (int64_t i = 0; i < 2 * hh; i++) out[hh + i] ^= p1[i];
}

void or_bpmul(const uint64_t *a, const uint64_t *b, int64_t m, uint64_t *out) {
  if (g_clmul < 0) {
#if defined(__x86_64__)
    __builtin_cpu_init();
    g_clmul = __builtin_cpu_supports("pclmul");
#else
    g_clmul = 0;
#endif
  }
  uint64_t *scr = (uint64_t *)calloc((size_t)(8 * m + 256), 8);
  bpmul_rec(a, b, m, out, scr);
  free(scr);
}

/* Fast exact encoder: per bit-row k, y_row = e0_row ^ (a * e1_row mod x^n-1)
 * computed with a Karatsuba/CLMUL binary-polynomial product. */
void or_qc_encode_fast(const orb *x, const uint8_t *a, int64_t n, orb *y, int nthreads) {
  int64_t m = (n + 63) / 64;
  uint64_t *aw = (uint64_t *)calloc((size_t)m, 8);
  for (int64_t i = 0; i < n; i++)
    if (getbit(a, i)) aw[i >> 6] |= 1ull << (i & 63);
  if (g_clmul < 0) {
    uint64_t dummy[2];
    uint64_t z = 0;
    or_bpmul(&z, &z, 1, dummy);
  }
  uint64_t *rows_all = (uint64_t *)calloc((size_t)(128 * m), 8);
#pragma omp parallel num_threads(nthreads)
  {
    uint64_t *ew = (uint64_t *)malloc((size_t)m * 8);
    uint64_t *pw = (uint64_t *)malloc((size_t)(2 * m) * 8);
    uint64_t *scr = (uint64_t *)calloc((size_t)(8 * m + 256), 8);
#pragma omp for schedule(dynamic)
    for (int k = 0; k < 128; k++) {
      memset(ew, 0, (size_t)m * 8);
      for (int64_t i = 0; i < n; i++) {
        const orb *bl = &x[n + i];
        uint64_t bit = k < 64 ? (bl->lo >> k) & 1 : (bl->hi >> (k - 64)) & 1;
        ew[i >> 6] |= bit << (i & 63);
      }
      bpmul_rec(aw, ew, m, pw, scr);
      /* fold mod x^n - 1: bit i of the result = p_i ^ p_{i+n} */
      uint64_t *rowbits = rows_all + (size_t)k * m;
      for (int64_t i = 0; i < 2 * n - 1; i++) {
        uint64_t bit = (pw[i >> 6] >> (i & 63)) & 1;
        if (bit) {
          int64_t r = i < n ? i : i - n;
          rowbits[r >> 6] ^= 1ull << (r & 63);
        }
      }
    }
    free(ew);
    free(pw);
    free(scr);
  }
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < n; i++) {
    uint64_t lo = 0, hi = 0;
    for (int k = 0; k < 64; k++) lo |= ((rows_all[(size_t)k * m + (i >> 6)] >> (i & 63)) & 1) << k;
    for (int k = 64; k < 128; k++) hi |= ((rows_all[(size_t)k * m + (i >> 6)] >> (i & 63)) & 1) << (k - 64);
    y[i].lo = x[i].lo ^ lo;
    y[i].hi = x[i].hi ^ hi;
  }
  free(rows_all);
  free(aw);
}

/* ======================================================================== */
/* 128-row bit transpose, bit order of gpu/gpu_ops.cu:31-70.                 */
/* ======================================================================== */
void or_bit_transpose(const orb *in, int64_t N, uint8_t *rows) {
  int64_t rb = N / 8;
  memset(rows, 0, (size_t)(128 * rb));
  for (int64_t mm = 0; mm < N; mm++) {
    const uint8_t *b = (const uint8_t *)&in[mm];
    for (int k = 0; k < 128; k++)
      if ((b[k >> 3] >> (k & 7)) & 1) rows[k * rb + (mm >> 3)] |= (uint8_t)(1u << (mm & 7));
  }
}

void or_bit_untranspose(const uint8_t *rows, int64_t N, orb *out) {
  int64_t rb = N / 8;
  memset(out, 0, (size_t)N * sizeof(orb));
  for (int64_t mm = 0; mm < N; mm++) {
    uint8_t *b = (uint8_t *)&out[mm];
    for (int k = 0; k < 128; k++)
      if ((rows[k * rb + (mm >> 3)] >> (mm & 7)) & 1) b[k >> 3] |= (uint8_t)(1u << (k & 7));
  }
}

/* ======================================================================== */
/* Ferret primal LPN, lpn_f2.h:46-79 (__compute4, __compute1, task).       */
/* ======================================================================== */
static int64_t lpn_mask(int64_t k) {
  int64_t mask = 1; /* lpn_f2.h:27-31 */
  while (mask < k) {
    mask <<= 1;
    mask |= 1;
  }
  return mask;
}

static void lpn_rows(int64_t start, int64_t end, int64_t k, int64_t mask, const uint8_t rk[176], int d,
                     void (*emit)(void *, int64_t, int, uint32_t), void *ctx, int nthreads) {
  /* __compute4 groups: j = start, start+4, ... while j < end - 4 */
  int64_t ngroups = 0;
  if (end - 4 > start) ngroups = (end - 4 - start + 3) / 4;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t g = 0; g < ngroups; g++) {
    int64_t i = start + 4 * g;
    orb tmp[16];
    for (int m = 0; m < d; m++) {
      tmp[m].lo = (uint64_t)m;
      tmp[m].hi = (uint64_t)i; /* makeBlock(i, m) */
    }
    or_aes_ecb(rk, tmp, tmp, (size_t)d, AES_FAST);
    const uint32_t *r = (const uint32_t *)tmp;
    for (int m = 0; m < 4; m++)
      for (int j = 0; j < d; j++) {
        int64_t index = (int64_t)((*r) & (uint32_t)mask);
        ++r;
        index = index >= k ? index - k : index;
        emit(ctx, i + m, j, (uint32_t)index);
      }
  }
  /* __compute1 tail */
  for (int64_t i = start + 4 * ngroups; i < end; i++) {
    int nb = d / 4 + (d % 4 != 0);
    orb tmp[8];
    for (int m = 0; m < nb; m++) {
      tmp[m].lo = (uint64_t)m;
      tmp[m].hi = (uint64_t)i;
    }
    or_aes_ecb(rk, tmp, tmp, (size_t)nb, AES_FAST);
    const uint32_t *r = (const uint32_t *)tmp;
    for (int j = 0; j < d; j++) emit(ctx, i, j, (uint32_t)(r[j] % (uint32_t)k));
  }
}

typedef struct {
  orb *nn;
  const orb *kk;
  uint32_t *idx;
  int d;
} lpn_ctx;

static void emit_xor(void *c, int64_t i, int j, uint32_t index) {
  lpn_ctx *x = (lpn_ctx *)c;
  (void)j;
  x->nn[i] = bx(x->nn[i], x->kk[index]);
}

static void emit_idx(void *c, int64_t i, int j, uint32_t index) {
  lpn_ctx *x = (lpn_ctx *)c;
  x->idx[i * x->d + j] = index;
}

static void lpn_drive(int64_t n, int64_t k, orb seed, int d, int ntasks, lpn_ctx *ctx,
                      void (*emit)(void *, int64_t, int, uint32_t), int nthreads) {
  uint8_t rk[176];
  or_aes_expand(&seed, rk); /* PRP prp(seed) */
  int64_t mask = lpn_mask(k);
  if (ntasks < 1) ntasks = 1;
  int64_t width = n / ntasks;
  for (int tk = 0; tk < ntasks; tk++) {
    int64_t start = tk * width;
    int64_t end = (tk == ntasks - 1) ? n : (tk + 1) * width;
    lpn_rows(start, end, k, mask, rk, d, emit, ctx, nthreads);
  }
}

void or_ferret_lpn(orb *nn, int64_t n, const orb *kk, int64_t k, orb seed, int d, int ntasks, int nthreads) {
  lpn_ctx c = {nn, kk, NULL, d};
  lpn_drive(n, k, seed, d, ntasks, &c, emit_xor, nthreads);
}

void or_ferret_lpn_indices(int64_t n, int64_t k, orb seed, int d, int ntasks, uint32_t *idx) {
  lpn_ctx c = {NULL, NULL, idx, d};
  lpn_drive(n, k, seed, d, ntasks, &c, emit_idx, 1);
}

/* ======================================================================== */
/* Whole instances from one seed.                                            */
/* ======================================================================== */
static orb derive_delta(const orb *seed) {
  orb d;
  or_prg(seed, 0, 0, 1, &d);
  d.lo |= 1;
  return d;
}

static uint8_t *derive_choices_th(const orb *seed, int64_t t, int h, int64_t *pos_out) {
  /* returns c[t][h]; pos_out[t] = t * 2^h + alpha_t */
  orb *al = (orb *)malloc((size_t)t * sizeof(orb));
  or_prg(seed, 2, 0, (size_t)t, al);
  uint8_t *c = (uint8_t *)malloc((size_t)(t * h));
  for (int64_t tr = 0; tr < t; tr++) {
    uint32_t alpha = (uint32_t)al[tr].lo & (uint32_t)((1ull << h) - 1);
    for (int d = 0; d < h; d++) c[tr * h + d] = (uint8_t)(1 - ((alpha >> (h - 1 - d)) & 1));
    if (pos_out) pos_out[tr] = (tr << h) + alpha;
  }
  free(al);
  return c;
}

void or_sot_run(const or_sot_cfg *cfg, int nthreads, orb *v, orb *w, uint8_t *b, orb *delta, orb *leaves_s,
                orb *leaves_r) {
  const int64_t n = cfg->n, t = cfg->t, N = 2 * n;
  const int h = cfg->h;
  uint8_t rk0[176], rk1[176];
  or_aes_expand(&cfg->k0, rk0);
  or_aes_expand(&cfg->k1, rk1);
  orb D = derive_delta(&cfg->seed);
  orb *roots = (orb *)malloc((size_t)t * sizeof(orb));
  or_prg(&cfg->seed, 1, 0, (size_t)t, roots);
  int64_t *pos = (int64_t *)malloc((size_t)t * sizeof(int64_t));
  uint8_t *cth = derive_choices_th(&cfg->seed, t, h, pos);
  uint8_t *c = (uint8_t *)malloc((size_t)(t * h)); /* [h][t] layout of recver.cu */
  for (int64_t tr = 0; tr < t; tr++)
    for (int d = 0; d < h; d++) c[d * t + tr] = cth[tr * h + d];
  int64_t ablocks = (n + 127) / 128;
  orb *ab = (orb *)malloc((size_t)ablocks * sizeof(orb));
  or_prg(&cfg->seed, 3, 0, (size_t)ablocks, ab);
  uint8_t *a = (uint8_t *)calloc((size_t)((n + 7) / 8), 1);
  memcpy(a, ab, (size_t)((n + 7) / 8));
  if (n % 8) a[(n - 1) / 8] &= (uint8_t)((1u << (n % 8)) - 1);
  orb *m0b = (orb *)malloc((size_t)(h * t) * sizeof(orb));
  orb *m1b = (orb *)malloc((size_t)(h * t) * sizeof(orb));
  or_prg(&cfg->seed, 4, 0, (size_t)(h * t), m0b);
  or_prg(&cfg->seed, 5, 0, (size_t)(h * t), m1b);
  orb *M0 = (orb *)malloc((size_t)((h + 1) * t) * sizeof(orb));
  orb *M1 = (orb *)malloc((size_t)((h + 1) * t) * sizeof(orb));
  orb *ls = leaves_s ? leaves_s : (orb *)malloc((size_t)N * sizeof(orb));
  orb *lr = leaves_r ? leaves_r : (orb *)malloc((size_t)N * sizeof(orb));
  or_sot_sender_expand(rk0, rk1, roots, t, h, D, m0b, m1b, N, ls, M0, M1, nthreads);
  or_sot_recver_expand(rk0, rk1, t, h, c, M0, M1, m0b, m1b, N, lr, nthreads);
  or_qc_encode_fast(ls, a, n, v, nthreads);
  or_qc_encode_fast(lr, a, n, w, nthreads);
  or_qc_sparse(pos, t, a, n, b);
  *delta = D;
  free(roots);
  free(pos);
  free(cth);
  free(c);
  free(ab);
  free(a);
  free(m0b);
  free(m1b);
  free(M0);
  free(M1);
  if (!leaves_s) free(ls);
  if (!leaves_r) free(lr);
}

void or_ferret_run(const or_ferret_cfg *cfg, int nthreads, orb *out_s, orb *out_r, orb *delta) {
  const int64_t n = cfg->n, t = cfg->t, k = cfg->k;
  const int h = cfg->h;
  orb D = derive_delta(&cfg->seed);
  orb *roots = (orb *)malloc((size_t)t * sizeof(orb));
  or_prg(&cfg->seed, 1, 0, (size_t)t, roots);
  uint8_t *cth = derive_choices_th(&cfg->seed, t, h, NULL);
  orb *ls = (orb *)malloc((size_t)(t * h) * sizeof(orb));
  orb *rs = (orb *)malloc((size_t)(t * h) * sizeof(orb));
  orb *ss = (orb *)malloc((size_t)t * sizeof(orb));
  or_ferret_mpcot_sender(roots, t, h, D, n, out_s, ls, rs, ss, nthreads);
  orb *csum = (orb *)malloc((size_t)(t * h) * sizeof(orb));
  for (int64_t i = 0; i < t * h; i++) csum[i] = cth[i] ? rs[i] : ls[i];
  or_ferret_mpcot_recver(t, h, cth, csum, ss, n, out_r, nthreads);
  orb seed_lpn;
  or_prg(&cfg->seed, 6, 0, 1, &seed_lpn);
  orb *K = (orb *)malloc((size_t)k * sizeof(orb));
  orb *Mb = (orb *)malloc((size_t)k * sizeof(orb));
  or_prg(&cfg->seed, 7, 0, (size_t)k, K);
  int64_t cb = (k + 127) / 128;
  orb *cbits = (orb *)malloc((size_t)cb * sizeof(orb));
  or_prg(&cfg->seed, 8, 0, (size_t)cb, cbits);
  for (int64_t j = 0; j < k; j++) {
    K[j].lo &= ~1ull;
    const uint8_t *cbb = (const uint8_t *)cbits;
    int cj = (cbb[j >> 3] >> (j & 7)) & 1;
    Mb[j] = K[j];
    if (cj) Mb[j] = bx(Mb[j], D);
  }
  or_ferret_lpn(out_s, n, K, k, seed_lpn, 10, cfg->ntasks, nthreads);
  or_ferret_lpn(out_r, n, Mb, k, seed_lpn, 10, cfg->ntasks, nthreads);
  *delta = D;
  free(roots);
  free(cth);
  free(ls);
  free(rs);
  free(ss);
  free(csum);
  free(K);
  free(Mb);
  free(cbits);
}

/* ---- iterating Ferret -------------------------------------------------------- */
void or_ccrh(const orb *in, orb *out, size_t n) {
  uint8_t rk[176];
  orb zero = {0, 0};
  or_aes_expand(&zero, rk);
  orb *sg = (orb *)malloc((n ? n : 1) * sizeof(orb));
  for (size_t i = 0; i < n; i++) {
    sg[i].lo = in[i].hi;
    sg[i].hi = in[i].hi ^ in[i].lo;
  }
  or_aes_ecb(rk, sg, out, n, 1);
  for (size_t i = 0; i < n; i++) out[i] = bx(out[i], sg[i]);
  free(sg);
}

int or_ferret_session_init(or_ferret_session *s, const or_ferret_cfg *cfg) {
  memset(s, 0, sizeof(*s));
  s->cfg = *cfg;
  const int64_t npre = (int64_t)cfg->h * cfg->t;
  s->M = cfg->k + npre + 128;
  s->ot_limit = cfg->n - s->M;
  if (s->ot_limit <= 0) return -1;
  s->delta = derive_delta(&cfg->seed);
  s->boot_s = (orb *)malloc((size_t)s->M * sizeof(orb));
  s->boot_r = (orb *)malloc((size_t)s->M * sizeof(orb));
  or_prg(&cfg->seed, 7, 0, (size_t)s->M, s->boot_s);
  const int64_t cb = (s->M + 127) / 128;
  orb *cbits = (orb *)malloc((size_t)cb * sizeof(orb));
  or_prg(&cfg->seed, 8, 0, (size_t)cb, cbits);
  const uint8_t *cbb = (const uint8_t *)cbits;
  for (int64_t j = 0; j < s->M; j++) {
    s->boot_s[j].lo &= ~1ull;
    s->boot_r[j] = s->boot_s[j];
    if ((cbb[j >> 3] >> (j & 7)) & 1) s->boot_r[j] = bx(s->boot_r[j], s->delta);
  }
  free(cbits);
  return 0;
}

void or_ferret_session_extend(or_ferret_session *s, int nthreads, orb *out_s, orb *out_r) {
  const or_ferret_cfg *cfg = &s->cfg;
  const int64_t n = cfg->n, t = cfg->t, k = cfg->k;
  const int h = cfg->h;
  const int64_t npre = (int64_t)h * t;
  orb D = s->delta;
  /* pre-OT */
  orb *P0 = (orb *)malloc((size_t)npre * sizeof(orb));
  orb *P1 = (orb *)malloc((size_t)npre * sizeof(orb));
  orb *Pr = (orb *)malloc((size_t)npre * sizeof(orb));
  orb *kd = (orb *)malloc((size_t)npre * sizeof(orb));
  uint8_t *bits = (uint8_t *)malloc((size_t)npre);
  or_ccrh(s->boot_s, P0, (size_t)npre);
  for (int64_t i = 0; i < npre; i++) kd[i] = bx(s->boot_s[i], D);
  or_ccrh(kd, P1, (size_t)npre);
  for (int64_t i = 0; i < npre; i++) bits[i] = (uint8_t)(s->boot_r[i].lo & 1);
  or_ccrh(s->boot_r, Pr, (size_t)npre);
  /* sender MPCOT */
  orb *roots = (orb *)malloc((size_t)t * sizeof(orb));
  or_prg(&cfg->seed, 1, (uint64_t)(s->round * t), (size_t)t, roots);
  orb *ls = (orb *)malloc((size_t)npre * sizeof(orb));
  orb *rs = (orb *)malloc((size_t)npre * sizeof(orb));
  orb *ss = (orb *)malloc((size_t)t * sizeof(orb));
  or_ferret_mpcot_sender(roots, t, h, D, n, out_s, ls, rs, ss, nthreads);
  /* level-sum OTs through the pre-OT pads; the receiver unmasks its choice */
  orb *csum = (orb *)malloc((size_t)npre * sizeof(orb));
  for (int64_t i = 0; i < npre; i++) {
    const orb pad0 = bx(ls[i], P0[i]), pad1 = bx(rs[i], P1[i]);
    csum[i] = bx(bits[i] ? pad1 : pad0, Pr[i]);
  }
  or_ferret_mpcot_recver(t, h, bits, csum, ss, n, out_r, nthreads);
  /* malicious mode (MpcotReg::set_malicious): the f2k consistency check on
   * the trees' leaves and the first 128 bootstrap COTs */
  if (s->malicious) {
    orb *V = (orb *)malloc((size_t)t * sizeof(orb));
    orb *Wt = (orb *)malloc((size_t)t * sizeof(orb));
    orb *ca = (orb *)malloc((size_t)t * sizeof(orb));
    or_ferret_consist_trees(out_s, out_r, n, t, h, ss, bits, V, Wt, ca, nthreads);
    s->check_ok = or_ferret_consist_final(V, Wt, ca, t, s->boot_s, s->boot_r, D, s->xprime, s->dig_a, s->dig_b);
    free(V);
    free(Wt);
    free(ca);
  }
  /* LPN with kk = bootstrap + 128 */
  orb seed_lpn;
  or_prg(&cfg->seed, 6, 0, 1, &seed_lpn);
  or_ferret_lpn(out_s, n, s->boot_s + 128, k, seed_lpn, 10, cfg->ntasks, nthreads);
  or_ferret_lpn(out_r, n, s->boot_r + 128, k, seed_lpn, 10, cfg->ntasks, nthreads);
  /* refresh */
  memcpy(s->boot_s, out_s + s->ot_limit, (size_t)s->M * sizeof(orb));
  memcpy(s->boot_r, out_r + s->ot_limit, (size_t)s->M * sizeof(orb));
  s->round++;
  free(P0);
  free(P1);
  free(Pr);
  free(kd);
  free(bits);
  free(roots);
  free(ls);
  free(rs);
  free(ss);
  free(csum);
}

void or_ferret_session_free(or_ferret_session *s) {
  free(s->boot_s);
  free(s->boot_r);
  s->boot_s = s->boot_r = NULL;
}


/* ==== malicious consistency check (emp-ot's f2k check; emp-tool arithmetic) ====
 * emp-tool @44b1dde is absent: its Hash (SHA-256), gfmul, uni_hash_coeff_gen,
 * vector_inn_prdt_sum_red and GaloisFieldPacking are restated from their
 * published definitions and the call sites spcot_sender.h:88-98,
 * spcot_recver.h:97-107, mpcot_reg.h:177-220. */

/* SHA-256, FIPS 180-4 section 6.2 */
static const uint32_t SHA_K[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
    0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
    0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
    0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
    0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
    0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
    0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};

static inline uint32_t ror32(uint32_t x, int r) { return (x >> r) | (x << (32 - r)); }

static void sha256_block(uint32_t H[8], const uint8_t blk[64]) {
  uint32_t w[64];
  for (int i = 0; i < 16; i++)
    w[i] = (uint32_t)blk[4 * i] << 24 | (uint32_t)blk[4 * i + 1] << 16 | (uint32_t)blk[4 * i + 2] << 8 | blk[4 * i + 3];
  for (int i = 16; i < 64; i++) {
    const uint32_t s0 = ror32(w[i - 15], 7) ^ ror32(w[i - 15], 18) ^ (w[i - 15] >> 3);
    const uint32_t s1 = ror32(w[i - 2], 17) ^ ror32(w[i - 2], 19) ^ (w[i - 2] >> 10);
    w[i] = w[i - 16] + s0 + w[i - 7] + s1;
  }
  uint32_t a = H[0], b = H[1], c = H[2], d = H[3], e = H[4], f = H[5], g = H[6], h = H[7];
  for (int i = 0; i < 64; i++) {
    const uint32_t S1 = ror32(e, 6) ^ ror32(e, 11) ^ ror32(e, 25);
    const uint32_t ch = (e & f) ^ (~e & g);
    const uint32_t t1 = h + S1 + ch + SHA_K[i] + w[i];
    const uint32_t S0 = ror32(a, 2) ^ ror32(a, 13) ^ ror32(a, 22);
    const uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
    const uint32_t t2 = S0 + mj;
    h = g;
    g = f;
    f = e;
    e = d + t1;
    d = c;
    c = b;
    b = a;
    a = t1 + t2;
  }
  H[0] += a; H[1] += b; H[2] += c; H[3] += d; H[4] += e; H[5] += f; H[6] += g; H[7] += h;
}

void or_sha256(const uint8_t *msg, size_t len, uint8_t out[32]) {
  uint32_t H[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                   0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
  uint8_t blk[64];
  size_t off = 0;
  while (len - off >= 64) {
    sha256_block(H, msg + off);
    off += 64;
  }
  const size_t rem = len - off;
  memset(blk, 0, 64);
  memcpy(blk, msg + off, rem);
  blk[rem] = 0x80;
  if (rem >= 56) {
    sha256_block(H, blk);
    memset(blk, 0, 64);
  }
  const uint64_t bits = (uint64_t)len * 8;
  for (int i = 0; i < 8; i++) blk[63 - i] = (uint8_t)(bits >> (8 * i));
  sha256_block(H, blk);
  for (int i = 0; i < 8; i++) {
    out[4 * i] = (uint8_t)(H[i] >> 24);
    out[4 * i + 1] = (uint8_t)(H[i] >> 16);
    out[4 * i + 2] = (uint8_t)(H[i] >> 8);
    out[4 * i + 3] = (uint8_t)H[i];
  }
}

/* GF(2^128) = GF(2)[x]/(x^128 + x^7 + x^2 + x + 1); bit i of the block
 * (lo bit 0 = byte 0 bit 0) is the coefficient of x^i (emp-tool gfmul). */
orb or_gf128_mul(orb a, orb b) {
  uint64_t p[4] = {0, 0, 0, 0}; /* 256-bit carry-less product */
  uint64_t l, h;
  clmul_soft(a.lo, b.lo, &l, &h);
  p[0] ^= l; p[1] ^= h;
  clmul_soft(a.lo, b.hi, &l, &h);
  p[1] ^= l; p[2] ^= h;
  clmul_soft(a.hi, b.lo, &l, &h);
  p[1] ^= l; p[2] ^= h;
  clmul_soft(a.hi, b.hi, &l, &h);
  p[2] ^= l; p[3] ^= h;
  /* x^128 = x^7 + x^2 + x + 1: fold the high half twice */
  for (int r = 0; r < 2; r++) {
    const uint64_t h0 = p[2], h1 = p[3];
    p[2] = p[3] = 0;
    /* H * (1 + x + x^2 + x^7), 135 bits: low 128 into p[0..1], the rest into p[2] */
    p[0] ^= h0 ^ (h0 << 1) ^ (h0 << 2) ^ (h0 << 7);
    p[1] ^= h1 ^ (h1 << 1) ^ (h1 << 2) ^ (h1 << 7) ^ (h0 >> 63) ^ (h0 >> 62) ^ (h0 >> 57);
    p[2] ^= (h1 >> 63) ^ (h1 >> 62) ^ (h1 >> 57);
  }
  orb o = {p[0], p[1]};
  return o;
}

static orb gf128_pow(orb s, uint64_t e) {
  orb r = {1, 0};
  while (e) {
    if (e & 1) r = or_gf128_mul(r, s);
    s = or_gf128_mul(s, s);
    e >>= 1;
  }
  return r;
}

static orb hash_seed(orb x) {
  uint8_t d[32];
  or_sha256((const uint8_t *)&x, 16, d);
  orb s;
  memcpy(&s, d, 16);
  return s;
}

void or_ferret_consist_trees(const orb *leaves_s, const orb *leaves_r, int64_t n, int64_t t, int h,
                             const orb *ssum, const uint8_t *bits, orb *V, orb *W, orb *chi_alpha, int nthreads) {
  const int64_t L = (int64_t)1 << h;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic)
  for (int64_t tr = 0; tr < t; tr++) {
    const orb s = hash_seed(ssum[tr]);
    int64_t alpha = 0;
    for (int d = 0; d < h; d++) alpha = (alpha << 1) + (bits[tr * h + d] ? 0 : 1);
    orb chi = s, v = {0, 0}, w = {0, 0};
    for (int64_t j = 0; j < L; j++) {
      const int64_t g = tr * L + j;
      if (g >= n) break;
      v = bx(v, or_gf128_mul(chi, leaves_s[g]));
      w = bx(w, or_gf128_mul(chi, leaves_r[g]));
      chi = or_gf128_mul(chi, s);
    }
    V[tr] = v;
    W[tr] = w;
    const orb zero = {0, 0};
    chi_alpha[tr] = (tr * L + alpha < n) ? gf128_pow(s, (uint64_t)alpha + 1) : zero;
  }
}

int or_ferret_consist_final(const orb *V, const orb *W, const orb *chi_alpha, int64_t t, const orb *pre_s,
                            const orb *pre_r, orb delta, uint8_t xprime[128], uint8_t dig_a[32],
                            uint8_t dig_b[32]) {
  orb r1a = {0, 0}, r1b = {0, 0}, r2 = {0, 0};
  for (int64_t i = 0; i < t; i++) {
    r1a = bx(r1a, V[i]);
    r1b = bx(r1b, W[i]);
    r2 = bx(r2, chi_alpha[i]);
  }
  orb pa = {0, 0}, pb = {0, 0};
  for (int j = 0; j < 128; j++) {
    const int rb = j < 64 ? (int)((r2.lo >> j) & 1) : (int)((r2.hi >> (j - 64)) & 1);
    xprime[j] = (uint8_t)(rb ^ (int)(pre_r[j].lo & 1));
    orb a = pre_s[j];
    if (xprime[j]) a = bx(a, delta);
    orb xj = {j < 64 ? 1ull << j : 0, j >= 64 ? 1ull << (j - 64) : 0};
    pa = bx(pa, or_gf128_mul(a, xj));
    pb = bx(pb, or_gf128_mul(pre_r[j], xj));
  }
  r1a = bx(r1a, pa);
  r1b = bx(r1b, pb);
  or_sha256((const uint8_t *)&r1a, 16, dig_a);
  or_sha256((const uint8_t *)&r1b, 16, dig_b);
  return memcmp(dig_a, dig_b, 32) == 0;
}

/* ==== online derandomisation (ferret_cot.hpp:277-295, cot.h:16-91) ==== */
void or_cot_online(orb *data_s, const orb *data_r, const uint8_t *b, orb delta, uint8_t *bo, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    bo[i] = (uint8_t)((b[i] & 1) ^ (data_r[i].lo & 1));
    if (bo[i]) data_s[i] = bx(data_s[i], delta);
  }
}

/* MITCCRH (emp-tool mitccrh.h): H_i(x) = AES_{s ^ makeBlock(g0 + i, 0)}(x) ^ x */
static orb mitccrh(orb s, uint64_t g, orb x) {
  orb key = {s.lo, s.hi ^ g};
  uint8_t rk[176];
  or_aes_expand(&key, rk);
  orb y;
  or_aes_ecb(rk, &x, &y, 1, 0);
  return bx(y, x);
}

void or_cot_to_ot(const orb *data_s, const orb *data_r, orb delta, orb s, uint64_t g0, const uint8_t *r,
                  const orb *m0, const orb *m1, orb *pad, orb *out_r, int64_t n, int nthreads) {
#pragma omp parallel for num_threads(nthreads)
  for (int64_t i = 0; i < n; i++) {
    const uint64_t g = g0 + (uint64_t)i;
    pad[2 * i] = bx(mitccrh(s, g, data_s[i]), m0[i]);
    pad[2 * i + 1] = bx(mitccrh(s, g, bx(data_s[i], delta)), m1[i]);
    out_r[i] = bx(pad[2 * i + (r[i] & 1)], mitccrh(s, g, data_r[i]));
  }
}

void or_cot_to_rot(const orb *data_s, const orb *data_r, orb delta, orb s, uint64_t g0, orb *data0, orb *data1,
                   orb *out_r, int64_t n, int nthreads) {
#pragma omp parallel for num_threads(nthreads)
  for (int64_t i = 0; i < n; i++) {
    const uint64_t g = g0 + (uint64_t)i;
    data0[i] = mitccrh(s, g, data_s[i]);
    data1[i] = mitccrh(s, g, bx(data_s[i], delta));
    out_r[i] = mitccrh(s, g, data_r[i]);
  }
}
