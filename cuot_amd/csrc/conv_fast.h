// conv_fast.h -- the bit-row basis conversion of conv_bits.h with every level
// specialised at compile time: the subspace-polynomial terms of level k
// (e = 2^j, j a proper submask of k) become immediate word offsets and shift
// counts, so the per-word loops carry no term tables.  Same arguments and
// results as cb_convert (the host test compares the two on every level).
#pragma once

#include "conv_bits.h"

struct alignas(16) cbt_w4 {
  uint32_t x, y, z, w;
};

template <int J>
CB_HD uint32_t cbt_lo_term(const uint32_t* hi, int64_t wj) {
  if constexpr (J >= 5) {
    constexpr int64_t qw = (int64_t)1 << (J - 5);
    return wj >= qw ? hi[wj - qw] : 0u;
  } else {
    constexpr int r = 1 << J;
    const uint32_t a = hi[wj];
    const uint32_t b = wj >= 1 ? hi[wj - 1] : 0u;
    return (a << r) | (b >> (32 - r));
  }
}

template <int K, int J = 0>
CB_HD void cbt_lo_terms(uint32_t& v, const uint32_t* hi, int64_t wj) {
  if constexpr (J < K) {
    if constexpr ((J & ~K) == 0) v ^= cbt_lo_term<J>(hi, wj);
    cbt_lo_terms<K, J + 1>(v, hi, wj);
  }
}

// largest proper submask exponent of K (-1 if none)
template <int K>
constexpr int cbt_jmax() {
  int m = -1;
  for (int j = 0; j < K; j++)
    if ((j & ~K) == 0) m = j;
  return m;
}

template <int K, int J = 0>
CB_HD void cbt_hi_terms(uint32_t& corr, const uint32_t* hi, int64_t hw, int64_t wi) {
  if constexpr (J < K) {
    if constexpr ((J & ~K) == 0) {
      if constexpr (J >= 5) {
        constexpr int64_t ew = (int64_t)1 << (J - 5);
        if (wi < ew) corr ^= hi[wi + hw - ew];
      } else {
        if (wi == 0) corr ^= hi[hw - 1] >> (32 - (1 << J));
      }
    }
    cbt_hi_terms<K, J + 1>(corr, hi, hw, wi);
  }
}

template <int K>
CB_HD void cbt_hi_fix(uint32_t* w, int64_t nw, int64_t tid, int64_t nth) {
  constexpr int JM = cbt_jmax<K>();
  if constexpr (JM < 0) {
    return;
  } else {
    constexpr int lhw = K - 5;
    constexpr int64_t hw = (int64_t)1 << lhw;
    constexpr int lnwr = JM >= 5 ? JM - 5 : 0;  // written words per chunk (log2)
    const int64_t total = (nw >> (lhw + 1)) << lnwr;
    for (int64_t t = tid; t < total; t += nth) {
      const int64_t g = t >> lnwr, wi = t & (((int64_t)1 << lnwr) - 1);
      uint32_t* hi = w + (g << (lhw + 1)) + hw;
      uint32_t corr = 0;
      cbt_hi_terms<K>(corr, hi, hw, wi);
      hi[wi] ^= corr;
    }
  }
}

template <int K>
CB_HD void cbt_lo_update(uint32_t* w, int64_t nw, int64_t tid, int64_t nth) {
  if constexpr (cbt_jmax<K>() < 0) {
    return;
  } else {
    constexpr int lhw = K - 5;
    constexpr int64_t hw = (int64_t)1 << lhw;
    const int64_t total = nw >> 1;
    for (int64_t t = tid; t < total; t += nth) {
      const int64_t g = t >> lhw, wj = t & (hw - 1);
      uint32_t* lo = w + (g << (lhw + 1));
      uint32_t v = lo[wj];
      cbt_lo_terms<K>(v, lo + hw, wj);
      lo[wj] = v;
    }
  }
}

// ---- in-word levels (k <= 4) with constant masks and shifts ----------------
constexpr uint32_t cbt_mask_lo(int b) {
  return b == 0 ? 0x55555555u : b == 1 ? 0x33333333u : b == 2 ? 0x0F0F0F0Fu : b == 3 ? 0x00FF00FFu : 0x0000FFFFu;
}

constexpr uint32_t cbt_mask_below(int j, int k) {
  uint32_t m = 0xFFFFFFFFu;
  for (int b = j; b <= k; b++) m &= cbt_mask_lo(b);
  return m;
}

// One level in place, one shift and one masked xor per term: with
// half = 2^K and e = 2^J, lo bit j (e <= j < half) takes hi bit j - e, i.e.
// word bit j + half - e, and hi bit i (i < e) takes hi bit i + half - e, word
// bit i + 2 half - e: both are x >> (half - e) under a mask.  The lo terms
// read hi bits only; the hi terms write hi bits below half + e <= 3 half / 2
// and read hi bits at or above 2 half - e >= 3 half / 2; so every term reads
// bits no term of its group writes, and the terms of a group commute.
template <int K, int J = 0>
CB_HD uint32_t cbt_word_lo_ip(uint32_t x) {
  if constexpr (J >= K) {
    return x;
  } else {
    if constexpr ((J & ~K) == 0) {
      constexpr uint32_t m = cbt_mask_lo(K) & ~cbt_mask_below(J, K);
      x ^= (x >> ((1 << K) - (1 << J))) & m;
    }
    return cbt_word_lo_ip<K, J + 1>(x);
  }
}

template <int K, int J = 0>
CB_HD uint32_t cbt_word_hi_ip(uint32_t x) {
  if constexpr (J >= K) {
    return x;
  } else {
    if constexpr ((J & ~K) == 0) {
      constexpr uint32_t m = cbt_mask_below(J, K) << (1 << K);
      x ^= (x >> ((1 << K) - (1 << J))) & m;
    }
    return cbt_word_hi_ip<K, J + 1>(x);
  }
}

template <int K, bool INV>
CB_HD uint32_t cbt_word_level(uint32_t x) {
  if constexpr (K == 0) {
    return x;  // s_0(x) = x^1: no lower terms
  } else if constexpr (!INV) {
    // division: fix the hi half first, then the lo half from the fixed hi
    return cbt_word_lo_ip<K>(cbt_word_hi_ip<K>(x));
  } else {
    // multiplication: the lo half from the original hi, then the hi fix
    return cbt_word_hi_ip<K>(cbt_word_lo_ip<K>(x));
  }
}

// in-word levels 4..0 (forward) or 0..4 (inverse) of every word; needs mp >= 5
CB_HD uint32_t cbt_word_all(uint32_t x, bool inverse) {
  if (!inverse) {
    x = cbt_word_level<4, false>(x);
    x = cbt_word_level<3, false>(x);
    x = cbt_word_level<2, false>(x);
    x = cbt_word_level<1, false>(x);
  } else {
    x = cbt_word_level<1, true>(x);
    x = cbt_word_level<2, true>(x);
    x = cbt_word_level<3, true>(x);
    x = cbt_word_level<4, true>(x);
  }
  return x;
}

CB_HD void cbt_word_levels(uint32_t* w, int64_t nw, bool inverse, int64_t tid, int64_t nth) {
  if ((nw & 3) == 0) {
    const int n4 = (int)(nw >> 2);
    for (int i = (int)tid; i < n4; i += (int)nth) {
      cbt_w4 v = ((const cbt_w4*)w)[i];
      v.x = cbt_word_all(v.x, inverse);
      v.y = cbt_word_all(v.y, inverse);
      v.z = cbt_word_all(v.z, inverse);
      v.w = cbt_word_all(v.w, inverse);
      ((cbt_w4*)w)[i] = v;
    }
  } else {
    for (int64_t i = tid; i < nw; i += nth) w[i] = cbt_word_all(w[i], inverse);
  }
}

// ---- four-word form of lo_update for levels with >= 4 words per half --------
CB_HD void cbt_x4(cbt_w4& a, uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3) {
  a.x ^= b0;
  a.y ^= b1;
  a.z ^= b2;
  a.w ^= b3;
}

// L ^= hi[wj - e .. wj - e + 3] for e = 2^J bits; H = hi[wj .. wj + 3],
// P = hi[wj - 4 .. wj - 1] (zero below the half)
template <int J>
CB_HD void cbt_lo_term4(cbt_w4& L, const cbt_w4& H, const cbt_w4& P, const uint32_t* hi, int wj) {
  if constexpr (J >= 7) {
    constexpr int qw = 1 << (J - 5);
    if (wj >= qw) {
      const cbt_w4 Q = *(const cbt_w4*)(hi + wj - qw);
      cbt_x4(L, Q.x, Q.y, Q.z, Q.w);
    }
  } else if constexpr (J == 6) {
    cbt_x4(L, P.z, P.w, H.x, H.y);
  } else if constexpr (J == 5) {
    cbt_x4(L, P.w, H.x, H.y, H.z);
  } else {
    constexpr int r = 1 << J;
    cbt_x4(L, (H.x << r) | (P.w >> (32 - r)), (H.y << r) | (H.x >> (32 - r)), (H.z << r) | (H.y >> (32 - r)),
           (H.w << r) | (H.z >> (32 - r)));
  }
}

template <int K, int J = 0>
CB_HD void cbt_lo_terms4(cbt_w4& L, const cbt_w4& H, const cbt_w4& P, const uint32_t* hi, int wj) {
  if constexpr (J < K) {
    if constexpr ((J & ~K) == 0) cbt_lo_term4<J>(L, H, P, hi, wj);
    cbt_lo_terms4<K, J + 1>(L, H, P, hi, wj);
  }
}

template <int K>
CB_HD void cbt_lo_update4(uint32_t* w, int nw, int tid, int nth) {
  static_assert(K >= 7, "four-word form needs >= 4 words per half");
  constexpr int lhw = K - 5;
  constexpr int hw = 1 << lhw;
  const int total4 = nw >> 3;  // groups of four lo words
  for (int t = tid; t < total4; t += nth) {
    const int t4 = t << 2;
    const int g = t4 >> lhw, wj = t4 & (hw - 1);
    uint32_t* lo = w + (g << (lhw + 1));
    const uint32_t* hi = lo + hw;
    cbt_w4 L = *(const cbt_w4*)(lo + wj);
    const cbt_w4 H = *(const cbt_w4*)(hi + wj);
    cbt_w4 P = {0u, 0u, 0u, 0u};
    if (wj >= 4) P = *(const cbt_w4*)(hi + wj - 4);
    cbt_lo_terms4<K>(L, H, P, hi, wj);
    *(cbt_w4*)(lo + wj) = L;
  }
}

template <typename Sync>
CB_HD void cbt_level(uint32_t* w, int64_t nw, int k, bool inverse, int64_t tid, int64_t nth, Sync sync) {
#define CBT_LO(K)                                                      \
  if constexpr (K >= 7)                                                \
    cbt_lo_update4<K>(w, (int)nw, (int)tid, (int)nth);                 \
  else                                                                 \
    cbt_lo_update<K>(w, nw, tid, nth);
#define CBT_CASE(K)                          \
  case K:                                    \
    if (!inverse) {                          \
      cbt_hi_fix<K>(w, nw, tid, nth);        \
      sync();                                \
      CBT_LO(K)                              \
    } else {                                 \
      CBT_LO(K)                              \
      sync();                                \
      cbt_hi_fix<K>(w, nw, tid, nth);        \
    }                                        \
    break;
  switch (k) {
    CBT_CASE(5)
    CBT_CASE(6)
    CBT_CASE(7)
    CBT_CASE(8)
    CBT_CASE(9)
    CBT_CASE(10)
    CBT_CASE(11)
    CBT_CASE(12)
    CBT_CASE(13)
    CBT_CASE(14)
    CBT_CASE(15)
    CBT_CASE(16)
    CBT_CASE(17)
    CBT_CASE(18)
    CBT_CASE(19)
    default:
      break;
  }
#undef CBT_CASE
#undef CBT_LO
  sync();
}

// whole conversion of a 2^mp-coefficient string (5 <= mp <= 20)
template <typename Sync>
CB_HD void cbt_convert(uint32_t* w, int mp, bool inverse, int64_t tid, int64_t nth, Sync sync) {
  const int64_t nw = ((int64_t)1 << mp) >> 5;
  if (!inverse) {
    for (int k = mp - 1; k >= 5; k--) cbt_level(w, nw, k, false, tid, nth, sync);
    cbt_word_levels(w, nw, false, tid, nth);
    sync();
  } else {
    cbt_word_levels(w, nw, true, tid, nth);
    sync();
    for (int k = 5; k <= mp - 1; k++) cbt_level(w, nw, k, true, tid, nth, sync);
  }
}

// ---- blocked form for strings of >= 2^10 coefficients -------------------------
// Levels k >= 10 run on the shared string as above; levels 9..5 and the in-word
// levels 4..0 act inside aligned 1024-bit blocks (32 words), so each block is
// converted by one lane in registers, with compile-time word indices and no
// barriers.  With PAD, word i of the shared string is stored at i + 4 * (i / 32):
// the eight 16-byte quads of a block then start in different banks for eight
// neighbouring lanes, and every quad stays 16-byte aligned.
template <bool PAD>
CB_HD int64_t cbp_at(int64_t i) {
  return PAD ? i + ((i >> 5) << 2) : i;
}

// shared words of a padded string of nw words
CB_HD int64_t cbp_words(int64_t nw, bool pad) { return pad ? nw + ((nw >> 5) << 2) : nw; }

template <bool PAD>
CB_HD cbt_w4* cbp_quad(uint32_t* w, int64_t i) {
  return (cbt_w4*)(w + cbp_at<PAD>(i));
}

template <int K, bool PAD, int J = 0>
CB_HD void cbp_hi_terms(uint32_t& corr, const uint32_t* w, int base, int hb, int hw, int wi) {
  if constexpr (J < K) {
    if constexpr ((J & ~K) == 0) {
      if constexpr (J >= 5) {
        constexpr int ew = 1 << (J - 5);
        if (wi < ew) corr ^= w[base + cbp_at<PAD>(hb + wi + hw - ew)];
      } else {
        if (wi == 0) corr ^= w[base + cbp_at<PAD>(hb + hw - 1)] >> (32 - (1 << J));
      }
    }
    cbp_hi_terms<K, PAD, J + 1>(corr, w, base, hb, hw, wi);
  }
}

// hi_fix of level K (>= 10) on the shared string, 32-bit indices, one word per
// task (the reference form for cbp_hi_fix4)
template <int K, bool PAD>
CB_HD void cbp_hi_fix(uint32_t* w, int base, int nw, int tid, int nth) {
  constexpr int JM = cbt_jmax<K>();
  constexpr int lhw = K - 5;
  constexpr int hw = 1 << lhw;
  constexpr int lnwr = JM >= 5 ? JM - 5 : 0;
  const int total = (nw >> (lhw + 1)) << lnwr;
  for (int t = tid; t < total; t += nth) {
    const int g = t >> lnwr, wi = t & ((1 << lnwr) - 1);
    const int hb = (g << (lhw + 1)) + hw;
    uint32_t corr = 0;
    cbp_hi_terms<K, PAD>(corr, w, base, hb, hw, wi);
    w[base + cbp_at<PAD>(hb + wi)] ^= corr;
  }
}

// Correction quad for hi words 4q..4q+3 of the chunk whose hi half starts at
// word hb: terms J >= 7 read the aligned quad ew words below the half's end
// (all four words or none, since ew is a multiple of 4); terms J <= 6 touch
// only words 0 and 1 and read the half's last quad T.
template <int K, bool PAD, int J = 0>
CB_HD void cbp_hi_terms4(cbt_w4& c, const uint32_t* w, int base, int hb, int hw, int q, const cbt_w4& T) {
  if constexpr (J < K) {
    if constexpr ((J & ~K) == 0) {
      if constexpr (J >= 7) {
        constexpr int ew = 1 << (J - 5);
        if (4 * q < ew) {
          const cbt_w4 Q = *(const cbt_w4*)(w + base + cbp_at<PAD>(hb + 4 * q + hw - ew));
          cbt_x4(c, Q.x, Q.y, Q.z, Q.w);
        }
      } else if constexpr (J == 6) {
        if (q == 0) {
          c.x ^= T.z;
          c.y ^= T.w;
        }
      } else if constexpr (J == 5) {
        if (q == 0) c.x ^= T.w;
      } else {
        if (q == 0) c.x ^= T.w >> (32 - (1 << J));
      }
    }
    cbp_hi_terms4<K, PAD, J + 1>(c, w, base, hb, hw, q, T);
  }
}

// hi_fix of level K (>= 10) in 16-byte quads, two tasks per lane in flight:
// every read comes from hi words >= hw / 2, which the fix never writes (the
// fixed words are those below e_max <= hw / 2), so the loads of both tasks
// precede both stores.  Same result as cbp_hi_fix.
template <int K, bool PAD>
CB_HD void cbp_hi_fix4(uint32_t* w, int base, int nw, int tid, int nth) {
  constexpr int JM = cbt_jmax<K>();
  static_assert(K >= 10 && JM >= 0, "levels >= 10 have lower terms");
  constexpr int lhw = K - 5;
  constexpr int hw = 1 << lhw;
  constexpr int lnq = JM >= 7 ? JM - 7 : 0;  // log2 quads per chunk
  const int total = (nw >> (lhw + 1)) << lnq;
  for (int t0 = tid; t0 < total; t0 += 2 * nth) {
    int hb[2], q[2];
    bool live[2];
    cbt_w4 c[2], H[2];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int t = t0 + u * nth;
      live[u] = t < total;
      const int tt = live[u] ? t : t0;  // an in-range task either way
      const int g = tt >> lnq;
      q[u] = tt & ((1 << lnq) - 1);
      hb[u] = (g << (lhw + 1)) + hw;
      const cbt_w4 T = *(const cbt_w4*)(w + base + cbp_at<PAD>(hb[u] + hw - 4));
      c[u] = cbt_w4{0u, 0u, 0u, 0u};
      cbp_hi_terms4<K, PAD>(c[u], w, base, hb[u], hw, q[u], T);
      H[u] = *(const cbt_w4*)(w + base + cbp_at<PAD>(hb[u] + 4 * q[u]));
    }
#pragma unroll
    for (int u = 0; u < 2; u++) {
      if (!live[u]) continue;
      cbt_x4(H[u], c[u].x, c[u].y, c[u].z, c[u].w);
      *(cbt_w4*)(w + base + cbp_at<PAD>(hb[u] + 4 * q[u])) = H[u];
    }
  }
}

// lo_update of level K (10 <= K <= 19) in runs of four quads (16 lo words) per
// task, 32-bit indices.  A run never crosses a half (hw >= 32 words); the hi
// quads 8 and 4 words before the run come along, so the terms with shifts
// below 16 words (J <= 8) are formed in registers, and those with J >= 9
// read one quad each, under a condition that is the same for the whole run.
template <int K, bool PAD, int J = 9>
CB_HD void cbq_far_terms(cbt_w4* L, const uint32_t* w, int base, int hb, int wj) {
  if constexpr (J < K) {
    if constexpr ((J & ~K) == 0) {
      constexpr int qw = 1 << (J - 5);
      if (wj >= qw) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const cbt_w4 Q = *(const cbt_w4*)(w + base + cbp_at<PAD>(hb + wj + 4 * i - qw));
          cbt_x4(L[i], Q.x, Q.y, Q.z, Q.w);
        }
      }
    }
    cbq_far_terms<K, PAD, J + 1>(L, w, base, hb, wj);
  }
}

template <int K, int J = 0>
CB_HD void cbq_near_terms(cbt_w4& L, const cbt_w4& H, const cbt_w4& P, const cbt_w4& P2) {
  if constexpr (J < K && J <= 8) {
    if constexpr ((J & ~K) == 0) {
      if constexpr (J == 8)
        cbt_x4(L, P2.x, P2.y, P2.z, P2.w);
      else if constexpr (J == 7)
        cbt_x4(L, P.x, P.y, P.z, P.w);
      else
        cbt_lo_term4<J>(L, H, P, nullptr, 0);  // J < 7 reads only H and P
    }
    cbq_near_terms<K, J + 1>(L, H, P, P2);
  }
}

template <int K, bool PAD>
CB_HD void cbq_lo_update(uint32_t* w, int base, int nw, int tid, int nth) {
  static_assert(K >= 10, "runs of 16 words need >= 32 words per half");
  constexpr int lhw = K - 5;
  constexpr int hw = 1 << lhw;
  const int nrun = nw >> 5;
  for (int r = tid; r < nrun; r += nth) {
    const int o = r << 4;
    const int g = o >> lhw, wj = o & (hw - 1);
    const int lb = g << (lhw + 1), hb = lb + hw;
    cbt_w4 L[4], H[6];  // H[0], H[1]: hi quads 8 and 4 words before the run (zero at the half's start)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      L[i] = *(const cbt_w4*)(w + base + cbp_at<PAD>(lb + wj + 4 * i));
      H[i + 2] = *(const cbt_w4*)(w + base + cbp_at<PAD>(hb + wj + 4 * i));
    }
    const int pb = wj > 0 ? hb + wj - 8 : hb;  // an in-range address either way
    H[0] = *(const cbt_w4*)(w + base + cbp_at<PAD>(pb));
    H[1] = *(const cbt_w4*)(w + base + cbp_at<PAD>(pb + 4));
    if (wj == 0) H[0] = H[1] = cbt_w4{0u, 0u, 0u, 0u};
    cbq_far_terms<K, PAD>(L, w, base, hb, wj);
#pragma unroll
    for (int i = 0; i < 4; i++) cbq_near_terms<K>(L[i], H[i + 2], H[i + 1], H[i]);
#pragma unroll
    for (int i = 0; i < 4; i++) *(cbt_w4*)(w + base + cbp_at<PAD>(lb + wj + 4 * i)) = L[i];
  }
}

// hi_fix of one chunk held in registers (hi = the chunk's upper half)
template <int K>
CB_HD void cbr_hi_fix(uint32_t* hi) {
  constexpr int JM = cbt_jmax<K>();
  constexpr int hw = 1 << (K - 5);
  constexpr int nwr = JM >= 5 ? 1 << (JM - 5) : 1;
  uint32_t corr[nwr];
#pragma unroll
  for (int wi = 0; wi < nwr; wi++) {
    uint32_t c = 0;
    cbt_hi_terms<K>(c, hi, hw, wi);  // reads words >= nwr only
    corr[wi] = c;
  }
#pragma unroll
  for (int wi = 0; wi < nwr; wi++) hi[wi] ^= corr[wi];
}

// level K (5..9) on a 1024-bit block W[32] in registers
template <int K, bool INV>
CB_HD void cbr_level(uint32_t* W) {
  constexpr int hw = 1 << (K - 5);
#pragma unroll
  for (int c = 0; c < 32; c += 2 * hw) {
    uint32_t* lo = W + c;
    uint32_t* hi = lo + hw;
    if constexpr (!INV) cbr_hi_fix<K>(hi);
#pragma unroll
    for (int wj = 0; wj < hw; wj++) {
      uint32_t v = lo[wj];
      cbt_lo_terms<K>(v, hi, wj);
      lo[wj] = v;
    }
    if constexpr (INV) cbr_hi_fix<K>(hi);
  }
}

// levels 9..0 (forward) or 0..9 (inverse) of one 1024-bit block; with
// INWORD = false the in-word levels 4..0 are left to the caller (they are the
// first inverse and the last forward step, so they can ride on the copy
// between global memory and LDS)
template <bool INV, bool INWORD = true>
CB_HD void cbr_block(uint32_t* W) {
  if constexpr (!INV) {
    cbr_level<9, false>(W);
    cbr_level<8, false>(W);
    cbr_level<7, false>(W);
    cbr_level<6, false>(W);
    cbr_level<5, false>(W);
    if constexpr (INWORD) {
#pragma unroll
      for (int i = 0; i < 32; i++) W[i] = cbt_word_all(W[i], false);
    }
  } else {
    if constexpr (INWORD) {
#pragma unroll
      for (int i = 0; i < 32; i++) {
        W[i] = cbt_word_all(W[i], true);
#ifdef __HIP_DEVICE_COMPILE__
        // one word at a time: interleaving all 32 in-word transforms would
        // hold more registers than the whole block form needs otherwise
        __builtin_amdgcn_sched_barrier(0);
#endif
      }
    }
    cbr_level<5, true>(W);
    cbr_level<6, true>(W);
    cbr_level<7, true>(W);
    cbr_level<8, true>(W);
    cbr_level<9, true>(W);
  }
}

template <bool PAD, bool INV, bool INWORD = true>
CB_HD void cbp_blocks(uint32_t* w, int64_t nw, int64_t tid, int64_t nth) {
  const int64_t nblk = nw >> 5;
  for (int64_t b = tid; b < nblk; b += nth) {
    uint32_t W[32];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const cbt_w4 v = *cbp_quad<PAD>(w, 32 * b + 4 * q);
      W[4 * q] = v.x;
      W[4 * q + 1] = v.y;
      W[4 * q + 2] = v.z;
      W[4 * q + 3] = v.w;
    }
    cbr_block<INV, INWORD>(W);
#pragma unroll
    for (int q = 0; q < 8; q++)
      *cbp_quad<PAD>(w, 32 * b + 4 * q) = cbt_w4{W[4 * q], W[4 * q + 1], W[4 * q + 2], W[4 * q + 3]};
  }
}

// The dispatcher stays out of line.  Its pointer argument is proved an LDS
// address (so that the accesses are LDS instructions, not flat ones) only
// while every caller passes the workgroup's LDS base itself; a string at an
// offset within it is therefore given as w plus a word offset `base`.
template <bool PAD, typename Sync>
CB_HD void cbp_level(uint32_t* w, int base, int64_t nw, int k, bool inverse, int64_t tid, int64_t nth, Sync sync) {
#define CBP_CASE(K)                                                  \
  case K:                                                            \
    if (!inverse) {                                                  \
      cbp_hi_fix4<K, PAD>(w, base, (int)nw, (int)tid, (int)nth);     \
      sync();                                                        \
      cbq_lo_update<K, PAD>(w, base, (int)nw, (int)tid, (int)nth);   \
    } else {                                                         \
      cbq_lo_update<K, PAD>(w, base, (int)nw, (int)tid, (int)nth);   \
      sync();                                                        \
      cbp_hi_fix4<K, PAD>(w, base, (int)nw, (int)tid, (int)nth);     \
    }                                                                \
    break;
  switch (k) {
    CBP_CASE(10)
    CBP_CASE(11)
    CBP_CASE(12)
    CBP_CASE(13)
    CBP_CASE(14)
    CBP_CASE(15)
    CBP_CASE(16)
    CBP_CASE(17)
    CBP_CASE(18)
    CBP_CASE(19)
    default:
      break;
  }
#undef CBP_CASE
  sync();
}

// whole conversion of a 2^mp-coefficient string, 10 <= mp <= 20, stored with
// or without the padding at w + base (same results as cbt_convert; with
// INWORD = false the caller applies cbt_word_all to every word before an
// inverse or after a forward conversion).  The direction is a template
// argument: one kernel holding both block forms would need more registers
// than either.
template <bool PAD, bool INV, bool INWORD = true, typename Sync>
CB_HD void cbp_convert(uint32_t* w, int base, int mp, int64_t tid, int64_t nth, Sync sync) {
  const int64_t nw = ((int64_t)1 << mp) >> 5;
  if constexpr (!INV) {
    for (int k = mp - 1; k >= 10; k--) cbp_level<PAD>(w, base, nw, k, false, tid, nth, sync);
    cbp_blocks<PAD, false, INWORD>(w + base, nw, tid, nth);
    sync();
  } else {
    cbp_blocks<PAD, true, INWORD>(w + base, nw, tid, nth);
    sync();
    for (int k = 10; k <= mp - 1; k++) cbp_level<PAD>(w, base, nw, k, true, tid, nth, sync);
  }
}
