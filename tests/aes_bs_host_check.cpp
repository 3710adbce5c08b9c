// Host build of the bitsliced AES-128 (cuot_amd/csrc/aes_bs.h) as a small
// shared library for tests/test_aes_bs_host.py: bs_sbox_host gives the
// circuit's S-box table, bs_ecb_host encrypts n blocks in groups of 32 (the
// key schedule is FIPS-197's, with the S-box taken from the circuit).
#include <string.h>

#include "../cuot_amd/csrc/aes_bs.h"

extern "C" void bs_sbox_host(uint8_t out[256]) {
  // 8 groups of 32 inputs: bit k of slice i = bit i of input 32 g + k
  for (int g = 0; g < 8; g++) {
    uint32_t x[8], y[8];
    for (int i = 0; i < 8; i++) {
      x[i] = 0;
      for (int k = 0; k < 32; k++) x[i] |= (uint32_t)(((32 * g + k) >> i) & 1) << k;
    }
    AES_BS_SBOX(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], y[0], y[1], y[2], y[3], y[4], y[5], y[6], y[7]);
    for (int k = 0; k < 32; k++) {
      uint8_t v = 0;
      for (int i = 0; i < 8; i++) v |= (uint8_t)(((y[i] >> k) & 1) << i);
      out[32 * g + k] = v ^ 0x63;
    }
  }
}

static void expand_key(const uint32_t key[4], uint32_t rk[44]) {
  uint8_t sb[256];
  bs_sbox_host(sb);
  static const uint8_t rcon[10] = {0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40, 0x80, 0x1b, 0x36};
  memcpy(rk, key, 16);
  for (int w = 4; w < 44; w++) {
    uint32_t t = rk[w - 1];
    if (w % 4 == 0) {
      t = (t >> 8) | (t << 24);  // RotWord on little-endian words
      t = (uint32_t)sb[t & 0xff] | ((uint32_t)sb[(t >> 8) & 0xff] << 8) | ((uint32_t)sb[(t >> 16) & 0xff] << 16) |
          ((uint32_t)sb[t >> 24] << 24);
      t ^= rcon[w / 4 - 1];
    }
    rk[w] = rk[w - 4] ^ t;
  }
}

extern "C" void bs_ecb_host(const uint32_t key[4], const uint32_t* in, uint32_t* out, int64_t n) {
  uint32_t rk[44], fk[44];
  expand_key(key, rk);
  bs_fold_sbox_constant(rk, fk);
  for (int64_t b = 0; b < n; b += 32) {
    uint32_t S[128];
    for (int w = 0; w < 4; w++)
      for (int k = 0; k < 32; k++) S[32 * w + k] = b + k < n ? in[4 * (b + k) + w] : 0u;
    for (int w = 0; w < 4; w++) bs_transpose32(S + 32 * w);
    bs_encrypt(S, fk);
    for (int w = 0; w < 4; w++) bs_transpose32(S + 32 * w);
    for (int w = 0; w < 4; w++)
      for (int k = 0; k < 32; k++)
        if (b + k < n) out[4 * (b + k) + w] = S[32 * w + k];
  }
}
