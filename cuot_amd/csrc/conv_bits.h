// conv_bits.h -- basis conversion between the monomial and the Lin-Chung-Han
// novel polynomial basis (Cantor basis) on ONE bit-row: a string of 2^mp
// coefficients in GF(2), packed 32 per little-endian word (coefficient j = bit
// j%32 of word j/32).  The conversion is GF(2)-linear and acts identically on
// every (bit-plane, row) of the bitsliced data, so after a 128-row bit
// transpose each row string can be converted on its own, with all levels of a
// 2^20-coefficient string (128 KB) inside one workgroup's LDS.
//
// Level k works on chunks of 2^(k+1) coefficients, half = 2^k, with the
// subspace polynomial s_k(x) = x^half + sum_{e in T_k} x^e, T_k = {2^j : j a
// proper submask of k}:
//   division (forward):   hi_i ^= XOR_{e > i} hi_{i + half - e}  (i < e_max)
//                         lo_j ^= XOR_{e <= j} hi_{j - e}
//   multiplication (inv): lo_j ^= XOR_{e <= j} hi_{j - e}  (original hi)
//                         hi_i ^= XOR_{e > i} hi_{i + half - e}
// Levels k <= 4 live inside one 32-bit word and are done with masks and
// shifts; levels k >= 5 are word-array operations.
//
// The functions take a thread index and a thread count so that the same code
// runs on the device (one workgroup, barriers between phases supplied by the
// caller through `sync`) and on the host (tid = 0, nth = 1, sync = no-op).
#pragma once

#include <stdint.h>

#ifndef __HIPCC__
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#endif

#define CB_HD __host__ __device__ inline

CB_HD int cb_lower_terms(int k, uint32_t* e) {
  int c = 0;
  for (int j = 0; j < k; j++)
    if ((j & ~k) == 0) e[c++] = 1u << j;
  return c;
}

// positions p (mod 2^(k+1)) with bit b clear, for b < 5
CB_HD uint32_t cb_mask_lo(int b) {
  return b == 0 ? 0x55555555u : b == 1 ? 0x33333333u : b == 2 ? 0x0F0F0F0Fu : b == 3 ? 0x00FF00FFu : 0x0000FFFFu;
}

// positions with (p mod 2^(k+1)) < 2^j  (bits j..k of p clear), j <= k <= 4
CB_HD uint32_t cb_mask_below(int j, int k) {
  uint32_t m = 0xFFFFFFFFu;
  for (int b = j; b <= k; b++) m &= cb_mask_lo(b);
  return m;
}

// one in-word level (k <= 4) on word x
CB_HD uint32_t cb_word_level(uint32_t x, int k, bool inverse) {
  uint32_t e[8];
  const int ne = cb_lower_terms(k, e);
  if (ne == 0) return x;
  const int half = 1 << k;
  const uint32_t mlo = cb_mask_lo(k);
  uint32_t h = (x >> half) & mlo;  // hi halves moved to the lo positions
  uint32_t l = x & mlo;
  if (!inverse) {
    uint32_t corr = 0;
    for (int q = 0; q < ne; q++) {
      const int j = 31 - __builtin_clz(e[q]);
      corr ^= (h >> (half - e[q])) & cb_mask_below(j, k);
    }
    h ^= corr;
    for (int q = 0; q < ne; q++) l ^= (h << e[q]) & mlo;
  } else {
    uint32_t corr = 0;
    for (int q = 0; q < ne; q++) {
      const int j = 31 - __builtin_clz(e[q]);
      l ^= (h << e[q]) & mlo;
      corr ^= (h >> (half - e[q])) & cb_mask_below(j, k);
    }
    h ^= corr;
  }
  return l | (h << half);
}

// all in-word levels (forward: 4..0, inverse: 0..4) on words [tid, nw) step nth
CB_HD void cb_word_levels(uint32_t* w, int64_t nw, int kmax, bool inverse, int64_t tid, int64_t nth) {
  for (int64_t i = tid; i < nw; i += nth) {
    uint32_t x = w[i];
    if (!inverse) {
      for (int k = kmax; k >= 0; k--) x = cb_word_level(x, k, false);
    } else {
      for (int k = 0; k <= kmax; k++) x = cb_word_level(x, k, true);
    }
    w[i] = x;
  }
}

// word-level phases for k >= 5 on a string of nw words (chunks of 2*hw words).
// hi_fix: hi bits i < e_max get hi bits i + half - e (e > i).
CB_HD void cb_hi_fix(uint32_t* w, int64_t nw, int k, int64_t tid, int64_t nth) {
  uint32_t e[16];
  const int ne = cb_lower_terms(k, e);
  if (ne == 0) return;
  const int lhw = k - 5;  // log2 words per half
  const int64_t hw = (int64_t)1 << lhw;
  const uint32_t emax = e[ne - 1];
  // written words per chunk: emax / 32 (a power of two) or 1
  const int lnwr = emax >= 32 ? (31 - __builtin_clz(emax)) - 5 : 0;
  const int64_t total = (nw >> (lhw + 1)) << lnwr;
  for (int64_t t = tid; t < total; t += nth) {
    const int64_t g = t >> lnwr, wi = t & (((int64_t)1 << lnwr) - 1);
    uint32_t* hi = w + (g << (lhw + 1)) + hw;
    uint32_t v = hi[wi];
    uint32_t corr = 0;
    for (int q = 0; q < ne; q++) {
      const uint32_t ee = e[q];
      if (ee >= 32) {
        if (wi < (int64_t)(ee >> 5)) corr ^= hi[wi + hw - (ee >> 5)];
      } else if (wi == 0) {
        corr ^= hi[hw - 1] >> (32 - ee);
      }
    }
    hi[wi] = v ^ corr;
  }
}

// lo_update: lo bits j ^= hi bits j - e (e <= j)
CB_HD void cb_lo_update(uint32_t* w, int64_t nw, int k, int64_t tid, int64_t nth) {
  uint32_t e[16];
  const int ne = cb_lower_terms(k, e);
  if (ne == 0) return;
  const int lhw = k - 5;
  const int64_t hw = (int64_t)1 << lhw;
  const int64_t total = nw >> 1;  // lo words of all chunks
  for (int64_t t = tid; t < total; t += nth) {
    const int64_t g = t >> lhw, wj = t & (hw - 1);
    uint32_t* lo = w + (g << (lhw + 1));
    const uint32_t* hi = lo + hw;
    uint32_t v = lo[wj];
    for (int q = 0; q < ne; q++) {
      const int64_t qw = e[q] >> 5;
      const int r = (int)(e[q] & 31);
      const uint32_t a = (wj - qw >= 0) ? hi[wj - qw] : 0u;
      if (r == 0) {
        v ^= a;
      } else {
        const uint32_t b = (wj - qw - 1 >= 0) ? hi[wj - qw - 1] : 0u;
        v ^= (a << r) | (b >> (32 - r));
      }
    }
    lo[wj] = v;
  }
}

// Whole conversion of a 2^mp-coefficient string held in w (nw = 2^mp / 32
// words, mp >= 5).  `sync` is called between phases that depend on each other.
template <typename Sync>
CB_HD void cb_convert(uint32_t* w, int mp, bool inverse, int64_t tid, int64_t nth, Sync sync) {
  const int64_t nw = ((int64_t)1 << mp) >> 5;
  const int kw = mp - 1 < 4 ? mp - 1 : 4;  // highest in-word level
  if (!inverse) {
    for (int k = mp - 1; k >= 5; k--) {
      cb_hi_fix(w, nw, k, tid, nth);
      sync();
      cb_lo_update(w, nw, k, tid, nth);
      sync();
    }
    cb_word_levels(w, nw, kw, false, tid, nth);
    sync();
  } else {
    cb_word_levels(w, nw, kw, true, tid, nth);
    sync();
    for (int k = 5; k <= mp - 1; k++) {
      cb_lo_update(w, nw, k, tid, nth);
      sync();
      cb_hi_fix(w, nw, k, tid, nth);
      sync();
    }
  }
}
