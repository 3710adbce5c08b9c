#include <cstdio>
#include <random>

// Host build of the GF(2^32) arithmetic used by the QC encoder
// (cuot_amd/csrc/gf32.h, the tower GF(2^16)[theta]):
//   * the scalar product is a field multiplication: commutative, associative,
//     distributive, x^(2^32 - 1) = 1 for x != 0 (order 2^32 - 1 group), theta
//     of degree 32, the subfield {x1 = 0} fixed by x -> x^(2^16);
//   * the chunk embedding: unpack(pack(u) * pack(w)) equals the carry-less
//     product u * w for 16-bit chunk polynomials (textbook clmul);
//   * the Cantor basis: beta_i^2 + beta_i = beta_{i-1}, beta_0..15 in the subfield;
//   * the bitsliced products (subfield constants, general constants) and the
//     bitsliced pack/unpack equal the scalar ones on every one of 32 rows.
#include "../cuot_amd/csrc/gf32.h"

static uint32_t gpow(uint32_t x, uint64_t e) {
  uint32_t r = 1;
  while (e) {
    if (e & 1) r = gf32_mul(r, x);
    x = gf32_mul(x, x);
    e >>= 1;
  }
  return r;
}

static uint64_t clmul(uint32_t a, uint32_t b) {
  uint64_t r = 0;
  for (int i = 0; i < 32; i++)
    if ((b >> i) & 1) r ^= (uint64_t)a << i;
  return r;
}

static uint32_t row_of(const uint32_t* planes, int n, int row) {
  uint32_t x = 0;
  for (int j = 0; j < n; j++) x |= ((planes[j] >> row) & 1u) << j;
  return x;
}

int main() {
  std::mt19937 r(3);
  // field axioms
  for (int t = 0; t < 3000; t++) {
    uint32_t a = r(), b = r(), c = r();
    if (gf32_mul(a, b) != gf32_mul(b, a) || gf32_mul(a, gf32_mul(b, c)) != gf32_mul(gf32_mul(a, b), c) ||
        gf32_mul(a, b ^ c) != (gf32_mul(a, b) ^ gf32_mul(a, c))) {
      printf("FAIL field axioms\n");
      return 1;
    }
    if (a && gpow(a, 0xFFFFFFFFull) != 1u) {
      printf("FAIL x^(2^32-1) != 1 for %08x\n", a);
      return 1;
    }
    const uint32_t s = c & 0xFFFFu;
    if (gpow(s, 1ull << 16) != s) {
      printf("FAIL subfield\n");
      return 1;
    }
  }
  if (gpow(0x10000u, 1ull << 16) == 0x10000u || gpow(0x10000u, 1ull << 32) != 0x10000u) {
    printf("FAIL theta degree\n");
    return 1;
  }
  // chunk embedding
  for (int t = 0; t < 3000; t++) {
    const uint32_t u = r() & 0xFFFFu, w = r() & 0xFFFFu;
    const uint32_t p = gf32_mul(gf32_pack_chunk(u), gf32_pack_chunk(w));
    uint32_t in[32], out[31];
    for (int j = 0; j < 32; j++) in[j] = ((p >> j) & 1u) ? 0xFFFFFFFFu : 0u;
    gf32_unpack_planes(in, out);
    uint64_t back = 0;
    for (int i = 0; i < 31; i++) back |= (uint64_t)(out[i] & 1u) << i;
    if (back != clmul(u, w)) {
      printf("FAIL chunk embedding u=%04x w=%04x\n", u, w);
      return 1;
    }
  }
  // Cantor basis
  uint32_t beta[32];
  if (gf32_cantor_basis(beta) || beta[0] != 1u) {
    printf("FAIL cantor\n");
    return 1;
  }
  for (int i = 1; i < 32; i++)
    if ((gf32_mul(beta[i], beta[i]) ^ beta[i]) != beta[i - 1] || (i < 16 && (beta[i] >> 16))) {
      printf("FAIL cantor %d\n", i);
      return 1;
    }
  // bitsliced products and maps
  for (int trial = 0; trial < 2000; trial++) {
    uint32_t v[32], a1[32], a2[32], a3[32], a4[32];
    for (int j = 0; j < 32; j++) v[j] = r();
    const uint32_t c = trial < 32 ? (1u << trial) : r();
    const uint32_t cs = c & 0xFFFFu;
    gf32_mulc<true>(v, c, a1);
    gf32_mulc<false>(v, c, a2);
    gf32_mulc_sub<true>(v, cs, a3);
    gf32_mulc_sub<false>(v, cs, a4);
    uint32_t pk[32], up[31], h0[16], h1[16], m16[16];
    gf32_pack_planes(v, pk);
    gf32_pack_planes_half<0>(v, h0);
    gf32_pack_planes_half<16>(v, h1);
    gf16_mulc(v, cs, m16);
    for (int j = 0; j < 16; j++)
      if (h0[j] != pk[j] || h1[j] != pk[16 + j] || m16[j] != a3[j]) {
        printf("FAIL half forms\n");
        return 1;
      }
    gf32_unpack_planes(v, up);
    for (int row = 0; row < 32; row++) {
      const uint32_t x = row_of(v, 32, row);
      const uint32_t y = gf32_mul(c, x), ys = gf32_mul(cs, x);
      if (row_of(a1, 32, row) != y || row_of(a2, 32, row) != y || row_of(a3, 32, row) != ys ||
          row_of(a4, 32, row) != ys) {
        printf("FAIL bitsliced product trial %d row %d c=%08x\n", trial, row, c);
        return 1;
      }
      if (row_of(pk, 32, row) != gf32_pack_chunk(x & 0xFFFFu)) {
        printf("FAIL bitsliced pack\n");
        return 1;
      }
      uint32_t ref = 0;
      for (int i = 0; i < 31; i++) ref |= (uint32_t)(__builtin_popcount(kGfChunk.unpack[i] & x) & 1) << i;
      if (row_of(up, 31, row) != ref) {
        printf("FAIL bitsliced unpack\n");
        return 1;
      }
    }
  }
  printf("gf32 tower ok\n");
}
