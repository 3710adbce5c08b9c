// aes.hip -- AES-128 tables (computed from the FIPS-197 definition), host key
// schedule, and the ECB / counter-mode kernels.
#include <string.h>

#include "aes_dev.h"
#include "kernels.h"

// ---------------------------------------------------------------------------
// Host: S-box from the definition, key schedule, T0.
// ---------------------------------------------------------------------------
static uint8_t gf8_mul(uint8_t a, uint8_t b) {
  uint8_t r = 0;
  while (b) {
    if (b & 1) r ^= a;
    a = (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1b : 0));
    b >>= 1;
  }
  return r;
}

static void sbox_host(uint8_t S[256]) {
  for (int x = 0; x < 256; x++) {
    uint8_t inv = 0;
    if (x) {
      uint8_t p = (uint8_t)x, r = 1;
      for (int e = 254; e; e >>= 1) {
        if (e & 1) r = gf8_mul(r, p);
        p = gf8_mul(p, p);
      }
      inv = r;
    }
    uint8_t s = inv;
    for (int k = 1; k <= 4; k++) s ^= (uint8_t)((inv << k) | (inv >> (8 - k)));
    S[x] = (uint8_t)(s ^ 0x63);
  }
}

void cuot_aes_T0_host(uint32_t T0[256]) {
  uint8_t S[256];
  sbox_host(S);
  for (int x = 0; x < 256; x++) {
    uint8_t s = S[x];
    T0[x] = (uint32_t)gf8_mul(s, 2) | ((uint32_t)s << 8) | ((uint32_t)s << 16) | ((uint32_t)gf8_mul(s, 3) << 24);
  }
}

void cuot_aes_expand_host(const uint8_t key[16], uint32_t rk[44]) {
  uint8_t S[256];
  sbox_host(S);
  uint8_t w[176];
  memcpy(w, key, 16);
  uint8_t rcon = 1;
  for (int i = 4; i < 44; i++) {
    uint8_t t[4];
    memcpy(t, w + 4 * (i - 1), 4);
    if (i % 4 == 0) {
      uint8_t t0 = t[0];
      t[0] = (uint8_t)(S[t[1]] ^ rcon);
      t[1] = S[t[2]];
      t[2] = S[t[3]];
      t[3] = S[t0];
      rcon = gf8_mul(rcon, 2);
    }
    for (int j = 0; j < 4; j++) w[4 * i + j] = (uint8_t)(w[4 * (i - 4) + j] ^ t[j]);
  }
  memcpy(rk, w, 176);  // little-endian words = bytes in order
}

// ---------------------------------------------------------------------------
// Device table registry (one copy per device, uploaded by cuot_init).
// ---------------------------------------------------------------------------
static uint32_t* g_T0_dev[64];

hipError_t cuot_upload_tables(int device) {
  if (device < 0 || device >= 64) return hipErrorInvalidDevice;
  if (g_T0_dev[device]) return hipSuccess;
  uint32_t T0[256];
  cuot_aes_T0_host(T0);
  uint32_t* d = nullptr;
  hipError_t e = hipMalloc(&d, sizeof(T0));
  if (e != hipSuccess) return e;
  e = hipMemcpy(d, T0, sizeof(T0), hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  g_T0_dev[device] = d;
  return hipSuccess;
}

const uint32_t* cuot_device_T0() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  return g_T0_dev[dev];
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
struct EcbArgs {
  const uint32_t* T0;
  uint32_t rk[44];
  const blk* in;
  blk* out;
  int64_t n;
  uint64_t stream;  // prg: counter high word
  uint64_t start;   // prg: counter base
};

template <bool PRG>
__global__ __launch_bounds__(256) void aes_stream_kernel(EcbArgs a) {
  __shared__ uint32_t tab[CUOT_AES_LDS_WORDS];
  aes_stage_table(tab, a.T0, threadIdx.x, 256);
  __syncthreads();
  const AesLds aes = aes_lds(tab);
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n; i += stride) {
    blk x;
    if (PRG) {
      const uint64_t c = a.start + (uint64_t)i;
      x = make_uint4((uint32_t)c, (uint32_t)(c >> 32), (uint32_t)a.stream, (uint32_t)(a.stream >> 32));
    } else {
      x = a.in[i];
    }
    a.out[i] = aes_enc(aes, a.rk, x);
  }
}

static unsigned stream_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return (unsigned)g;
}

hipError_t cuot_launch_aes_ecb(const uint32_t rk[44], const blk* in, blk* out, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (cuot_knobs().aes_impl == 2) return cuot_launch_aes_ecb_bs(rk, in, out, n, st);
  EcbArgs a;
  a.T0 = cuot_device_T0();
  memcpy(a.rk, rk, sizeof(a.rk));
  a.in = in;
  a.out = out;
  a.n = n;
  a.stream = a.start = 0;
  cuot_rec_begin(st, "aes_stream", 32.0 * (double)n);
  hipLaunchKernelGGL(aes_stream_kernel<false>, dim3(stream_grid(n)), dim3(256), 0, st, a);
  cuot_rec_end(st);
  return hipGetLastError();
}

hipError_t cuot_launch_prg(const uint32_t rk[44], uint64_t stream, uint64_t start, int64_t n, blk* out,
                           hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (cuot_knobs().aes_impl == 2) return cuot_launch_prg_bs(rk, stream, start, n, out, st);
  EcbArgs a;
  a.T0 = cuot_device_T0();
  memcpy(a.rk, rk, sizeof(a.rk));
  a.in = nullptr;
  a.out = out;
  a.n = n;
  a.stream = stream;
  a.start = start;
  cuot_rec_begin(st, "aes_stream", 16.0 * (double)n);
  hipLaunchKernelGGL(aes_stream_kernel<true>, dim3(stream_grid(n)), dim3(256), 0, st, a);
  cuot_rec_end(st);
  return hipGetLastError();
}

void cuot_aes_encrypt_host(const uint32_t rk[44], const uint8_t in[16], uint8_t out[16]) {
  uint8_t S[256];
  sbox_host(S);
  const uint8_t* k = (const uint8_t*)rk;
  uint8_t s[16], u[16];
  for (int i = 0; i < 16; i++) s[i] = in[i] ^ k[i];
  for (int r = 1; r <= 10; r++) {
    for (int c = 0; c < 4; c++)
      for (int row = 0; row < 4; row++) u[row + 4 * c] = S[s[row + 4 * ((c + row) & 3)]];
    if (r < 10) {
      for (int c = 0; c < 4; c++) {
        const uint8_t a0 = u[4 * c], a1 = u[4 * c + 1], a2 = u[4 * c + 2], a3 = u[4 * c + 3];
        s[4 * c + 0] = gf8_mul(a0, 2) ^ gf8_mul(a1, 3) ^ a2 ^ a3;
        s[4 * c + 1] = a0 ^ gf8_mul(a1, 2) ^ gf8_mul(a2, 3) ^ a3;
        s[4 * c + 2] = a0 ^ a1 ^ gf8_mul(a2, 2) ^ gf8_mul(a3, 3);
        s[4 * c + 3] = gf8_mul(a0, 3) ^ a1 ^ a2 ^ gf8_mul(a3, 2);
      }
    } else {
      memcpy(s, u, 16);
    }
    for (int i = 0; i < 16; i++) s[i] ^= k[16 * r + i];
  }
  memcpy(out, s, 16);
}
