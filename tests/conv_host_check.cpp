// Host build of the bit-row basis conversion (cuot_amd/csrc/conv_bits.h and
// conv_fast.h) checked against a direct element-level restatement of the
// division / multiplication by the subspace polynomials, plus round trips.
// Built and run by tests/test_conv_host.py.
#include <cstdio>
#include <random>
#include <vector>

#include "../cuot_amd/csrc/conv_fast.h"

static void lower(int k, std::vector<int64_t>& e) {
  e.clear();
  for (int j = 0; j < k; j++)
    if ((j & ~k) == 0) e.push_back(1ll << j);
}

// level k on chunks of 2^(k+1) bits: s_k(x) = x^half + sum_{e in T_k} x^e
static void ref_conv(std::vector<int>& X, int mp, bool inv) {
  const int64_t Mp = 1ll << mp;
  std::vector<int64_t> e;
  auto hi_fix = [&](int k) {
    lower(k, e);
    if (e.empty()) return;
    const int64_t half = 1ll << k, emax = e.back();
    for (int64_t g = 0; g < Mp / (2 * half); g++) {
      const int64_t b = g * 2 * half + half;
      std::vector<int> nw(emax);
      for (int64_t i = 0; i < emax; i++) {
        int v = X[b + i];
        for (auto ee : e)
          if (ee > i) v ^= X[b + i + half - ee];
        nw[i] = v;
      }
      for (int64_t i = 0; i < emax; i++) X[b + i] = nw[i];
    }
  };
  auto lo_up = [&](int k) {
    lower(k, e);
    if (e.empty()) return;
    const int64_t half = 1ll << k;
    for (int64_t g = 0; g < Mp / (2 * half); g++) {
      const int64_t lo = g * 2 * half, hi = lo + half;
      std::vector<int> nw(half);
      for (int64_t j = 0; j < half; j++) {
        int v = X[lo + j];
        for (auto ee : e)
          if (ee <= j) v ^= X[hi + j - ee];
        nw[j] = v;
      }
      for (int64_t j = 0; j < half; j++) X[lo + j] = nw[j];
    }
  };
  if (!inv) {
    for (int k = mp - 1; k >= 0; k--) {
      hi_fix(k);
      lo_up(k);
    }
  } else {
    for (int k = 0; k < mp; k++) {
      lo_up(k);
      hi_fix(k);
    }
  }
}

int main() {
  std::mt19937_64 rng(1);
  auto nosync = []() {};
  for (int mp = 5; mp <= 13; mp++)
    for (int inv = 0; inv < 2; inv++)
      for (int trial = 0; trial < 2; trial++) {
        const int64_t n = 1ll << mp;
        std::vector<int> X(n);
        std::vector<uint32_t> W(n / 32, 0);
        for (int64_t i = 0; i < n; i++) {
          X[i] = rng() & 1;
          if (X[i]) W[i >> 5] |= 1u << (i & 31);
        }
        std::vector<uint32_t> W2 = W;
        ref_conv(X, mp, inv);
        cb_convert(W.data(), mp, inv, 0, 1, nosync);
        cbt_convert(W2.data(), mp, inv, 0, 1, nosync);
        for (int64_t i = 0; i < n; i++) {
          const uint32_t b = (W[i >> 5] >> (i & 31)) & 1, b2 = (W2[i >> 5] >> (i & 31)) & 1;
          if (b != (uint32_t)X[i] || b2 != (uint32_t)X[i]) {
            printf("MISMATCH mp=%d inv=%d bit=%lld\n", mp, inv, (long long)i);
            return 1;
          }
        }
      }
  for (int mp = 5; mp <= 20; mp++) {
    const int64_t n = 1ll << mp;
    std::vector<uint32_t> W(n / 32), W0;
    for (auto& x : W) x = (uint32_t)rng();
    W0 = W;
    std::vector<uint32_t> A = W;
    cb_convert(A.data(), mp, false, 0, 1, nosync);
    cbt_convert(W.data(), mp, false, 0, 1, nosync);
    if (A != W) {
      printf("specialised forward differs mp=%d\n", mp);
      return 1;
    }
    cbt_convert(W.data(), mp, true, 0, 1, nosync);
    if (W != W0) {
      printf("round trip fails mp=%d\n", mp);
      return 1;
    }
  }
  // blocked form (levels >= 10 on the string, 9..0 per 1024-bit block), with
  // and without the padded storage
  for (int mp = 10; mp <= 20; mp++)
    for (int inv = 0; inv < 2; inv++) {
      const int64_t n = 1ll << mp, nw = n / 32;
      std::vector<uint32_t> W(nw);
      for (auto& x : W) x = (uint32_t)rng();
      std::vector<uint32_t> A = W, B = W;
      cb_convert(A.data(), mp, inv, 0, 1, nosync);
      if (inv)
        cbp_convert<false, true>(B.data(), 0, mp, 0, 1, nosync);
      else
        cbp_convert<false, false>(B.data(), 0, mp, 0, 1, nosync);
      std::vector<uint32_t> P(cbp_words(nw, true) + 4, 0xdeadbeefu);
      for (int64_t i = 0; i < nw; i++) P[cbp_at<true>(i)] = W[i];
      if (inv)
        cbp_convert<true, true>(P.data() + 4, -4, mp, 0, 1, nosync);
      else
        cbp_convert<true, false>(P.data() + 4, -4, mp, 0, 1, nosync);
      // in-word levels outside the blocked form
      std::vector<uint32_t> C = W;
      if (inv) {
        for (auto& x : C) x = cbt_word_all(x, true);
        cbp_convert<false, true, false>(C.data(), 0, mp, 0, 1, nosync);
      } else {
        cbp_convert<false, false, false>(C.data(), 0, mp, 0, 1, nosync);
        for (auto& x : C) x = cbt_word_all(x, false);
      }
      bool ok = A == B && A == C;
      for (int64_t i = 0; i < nw; i++) ok = ok && P[cbp_at<true>(i)] == A[i];
      for (int64_t i = 0; i < nw; i += 32)
        for (int pd = 32; pd < 36; pd++) ok = ok && P[i + (i >> 5) * 4 + pd] == 0xdeadbeefu;
      if (!ok) {
        printf("blocked form differs mp=%d inv=%d\n", mp, inv);
        return 1;
      }
    }
  printf("conversion ok\n");
  return 0;
}
