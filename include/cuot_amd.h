/*
 * cuot_amd.h -- C ABI of the MI355X (gfx950) silent-OT correlation engine.
 *
 * Plain C: opaque context, plain pointers and sizes, int status codes, no
 * exceptions across the ABI.  Device pointers are HIP device addresses
 * (suffix _d); host pointers have suffix _h.  `stream` is a hipStream_t
 * passed as void* (NULL = the context's default stream).  Stage entry points
 * are asynchronous on `stream`; the drop-in entry points that replace the
 * reference's blocking calls block until their results are on the host.
 *
 * Block = 16 bytes in memory order (cuot_block.lo = bytes 0..7, little endian),
 * i.e. the reference's `blk` (gpu/gpu_define.h:12-18) and emp-tool's `block`.
 * Bit vectors are packed LSB-first: bit i = byte i/8, bit i%8.
 */
#ifndef CUOT_AMD_H
#define CUOT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint64_t lo, hi;
} cuot_block;

typedef struct cuot_ctx cuot_ctx;

enum {
  CUOT_OK = 0,
  CUOT_ERR_ARG = 1,     /* invalid argument (sizes, pointers, parameters) */
  CUOT_ERR_HIP = 2,     /* HIP runtime error (message in cuot_last_error) */
  CUOT_ERR_NOMEM = 3,   /* device allocation or budget exhausted */
  CUOT_ERR_STATE = 4,   /* call out of order / object in wrong state */
  CUOT_ERR_BUSY = 5     /* admission refused (serving) */
};

enum { CUOT_MODE_SOT = 0, CUOT_MODE_FERRET = 1 };
enum { CUOT_ROLE_SENDER = 0, CUOT_ROLE_RECVER = 1 };

/* ---- context ------------------------------------------------------------- */
/* Binds a device, uploads the AES tables.  Replaces cuda_setdev
 * (ferret/emp-ot/emp-ot/ferret/dev_layer.h:9). */
int cuot_init(int device, cuot_ctx **out);
int cuot_destroy(cuot_ctx *ctx);
const char *cuot_last_error(void);
int cuot_device_sync(cuot_ctx *ctx);

/* ---- AES-128 (FIPS-197) ---------------------------------------------------
 * Replaces Aes::encrypt (gpu/aes_expand.cu:32-36) with the standard key
 * schedule (the reference's gpu/aes_expand.cu:59 Rcon offset is not kept) and
 * the key passed per call instead of global __constant__ slots
 * (gpu/aes_expand.cu:15-30). */
int cuot_aes_ecb(cuot_ctx *ctx, const cuot_block *key_h, const cuot_block *in_d, cuot_block *out_d,
                 int64_t n, void *stream);
/* out[i] = AES_seed({lo = start + i, hi = stream_id}): counter-mode stream used
 * for all seed derivation (replaces make_block + Aes::encrypt,
 * gpu/gpu_ops.cu:102-109). */
int cuot_prg(cuot_ctx *ctx, const cuot_block *seed_h, uint64_t stream_id, uint64_t start, int64_t n,
             cuot_block *out_d, void *stream);

/* ---- GGM PPRF expansion ----------------------------------------------------
 * t trees of depth h (1 <= h <= 16), child_L = AES_k0(s)^s, child_R =
 * AES_k1(s)^s.  Leaves are tree-major in natural order; global leaf index
 * g = tree*2^h + i is written only if g < n_out (truncation).  Replaces
 * Aes::expand + Mat::sum per level (gpu/aes_expand.cu:38-46,
 * gpu/gpu_matrix.cu:155-169) as driven by silent/sender.cu:166-198,
 * silent/recver.cu:333-372 and dev_layer.cu:40-164. */
typedef struct {
  int64_t t;
  int32_t h;
  int32_t mode; /* CUOT_MODE_FERRET clears leaf LSBs as spcot_sender.h:118-121 */
  int64_t n_out;
  cuot_block k0, k1;
} cuot_ggm_params;

/* sums_d: [t][h][2] = xor of even / odd nodes at depth d+1. */
int cuot_ggm_sender(cuot_ctx *ctx, const cuot_ggm_params *p, const cuot_block *roots_d, cuot_block *leaves_d,
                    cuot_block *sums_d, void *stream);
/* choice_d: [t][h] bytes (1 = receiver chose the odd side at that level; the
 * punctured path follows the other side).  csum_d: [t][h] chosen sums.
 * fin_d: [t]: SOT = final-row value (m_c[h] ^ mc[h-1]); Ferret = secret sum. */
int cuot_ggm_recver(cuot_ctx *ctx, const cuot_ggm_params *p, const uint8_t *choice_d, const cuot_block *csum_d,
                    const cuot_block *fin_d, cuot_block *leaves_d, void *stream);

#ifdef __cplusplus
}
#endif
#endif
