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
 * gpu/gpu_matrix.cu:155-169) as driven by silent/sender.cu:76-108,
 * silent/recver.cu:124-163 and dev_layer.cu:40-164. */
typedef struct {
  int64_t t;
  int32_t h;
  int32_t mode; /* CUOT_MODE_FERRET clears leaf LSBs as spcot_sender.h:52-55 */
  int64_t n_out;
  cuot_block k0, k1;
} cuot_ggm_params;

/* sums_d: [t][h][2] = xor of even / odd nodes at depth d+1. */
int cuot_ggm_sender(cuot_ctx *ctx, const cuot_ggm_params *p, const cuot_block *roots_d, cuot_block *leaves_d,
                    cuot_block *sums_d, void *stream);
/* The same sender outputs from the bitsliced AES variant, breadth-first with
 * one launch per level through a caller work buffer of
 * cuot_ggm_sender_bs_work_bytes(p) bytes (two levels of t * 2^(h-1) blocks);
 * the LDS T-table form above is the one the library's pipelines use.  Like
 * cuot_ggm_sender it replaces the per-level Aes::expand + Mat::sum of
 * gpu/aes_expand.cu:38-46, gpu/gpu_matrix.cu:155-169. */
size_t cuot_ggm_sender_bs_work_bytes(const cuot_ggm_params *p);
int cuot_ggm_sender_bs(cuot_ctx *ctx, const cuot_ggm_params *p, const cuot_block *roots_d, cuot_block *leaves_d,
                       cuot_block *sums_d, void *work_d, size_t work_bytes, void *stream);
/* choice_d: [t][h] bytes (1 = receiver chose the odd side at that level; the
 * punctured path follows the other side).  csum_d: [t][h] chosen sums.
 * fin_d: [t]: SOT = final-row value (m_c[h] ^ mc[h-1]); Ferret = secret sum. */
int cuot_ggm_recver(cuot_ctx *ctx, const cuot_ggm_params *p, const uint8_t *choice_d, const cuot_block *csum_d,
                    const cuot_block *fin_d, cuot_block *leaves_d, void *stream);
/* The same receiver outputs from the bitsliced AES variant (breadth-first,
 * one expansion and one sibling-fix launch per level) through a caller work
 * buffer of cuot_ggm_recver_bs_work_bytes(p) bytes; for the cipher comparison,
 * the pipelines use cuot_ggm_recver.  Replaces the receiver level loop of
 * silent/recver.cu:124-163 and dev_layer.cu:62-143 like cuot_ggm_recver. */
size_t cuot_ggm_recver_bs_work_bytes(const cuot_ggm_params *p);
int cuot_ggm_recver_bs(cuot_ctx *ctx, const cuot_ggm_params *p, const uint8_t *choice_d, const cuot_block *csum_d,
                       const cuot_block *fin_d, cuot_block *leaves_d, void *work_d, size_t work_bytes, void *stream);

/* ---- Ferret 10-local primal LPN -------------------------------------------
 * out[i] = noise[i] ^ XOR_{j<10} kk[idx(i,j)], indices from AES under `seed`
 * with emp-ot's counter and reduction rules (lpn_f2.h:46-79) over `ntasks`
 * equal tasks (LpnF2::task per task, lpn_f2.h:72-79; the task split is the commented-out
 * LpnF2::bench, lpn_f2.h:100-115; emp-ot's single-thread run is ntasks = 1).
 * out may alias noise.  Replaces cuda_primal_lpn's device work
 * (dev_layer.cu:146-186) without materialising the index array. */
typedef struct {
  int64_t n;
  int64_t k;
  int32_t d;      /* must be 10 (LpnF2<IO, 10>, ferret_cot.h) */
  int32_t ntasks; /* >= 1 */
  cuot_block seed;
} cuot_lpn_params;

int cuot_lpn(cuot_ctx *ctx, const cuot_lpn_params *p, const cuot_block *noise_d, const cuot_block *kk_d,
             cuot_block *out_d, void *stream);
/* Both roles in one pass (sender: noise_s/kk_s, receiver: noise_r/kk_r). */
int cuot_lpn_dual(cuot_ctx *ctx, const cuot_lpn_params *p, const cuot_block *noise_s_d,
                  const cuot_block *kk_s_d, cuot_block *out_s_d, const cuot_block *noise_r_d,
                  const cuot_block *kk_r_d, cuot_block *out_r_d, void *stream);

/* ---- quasi-cyclic encoder (exact GF(2)) ------------------------------------
 * y_i = x_i ^ XOR_{j: a_j = 1} x_{n + ((i - j) mod n)}, i < n, for all 128
 * bit-rows of the blocks at once: the workload's 2n -> n circulant code.
 * Replaces QuasiCyclic::encode_dense's float32 cuFFT path
 * (silent/quasi_cyclic.cu:118-128) and its per-8-row batching by an exact
 * additive FFT over GF(2^32), or, for small codes (17 <= n <= 2^16, at or
 * below the knob qc_mfma_max), by the dense 0/1 int8 product on the matrix
 * cores with int32 sums taken mod 2.  a_bits_d: n bits.  ahat_d:
 * cuot_qc_ahat_len(n) words computed once per code by cuot_qc_prepare
 * (QuasiCyclic::QuasiCyclic, silent/quasi_cyclic.cu:51-106): the transform,
 * followed for small codes by a's bits, so both encoders run on it.
 * scratch_d: cuot_qc_scratch_bytes(n) bytes.  y_d may alias x_d. */
int64_t cuot_qc_ahat_len(int64_t n);
size_t cuot_qc_scratch_bytes(int64_t n);
int cuot_qc_prepare(cuot_ctx *ctx, int64_t n, const uint8_t *a_bits_d, uint32_t *ahat_d, void *scratch_d,
                    void *stream);
int cuot_qc_encode(cuot_ctx *ctx, int64_t n, const uint32_t *ahat_d, const cuot_block *x_d, cuot_block *y_d,
                   void *scratch_d, void *stream);
/* Receiver's choice bits: the same code on the weight-w noise vector e with
 * ones at pos_d[0..w) (positions in [0, 2n); others ignored).  Replaces
 * QuasiCyclic::encode_sparse (silent/quasi_cyclic.cu:130-161).  a_bits_d must
 * be readable for ((n + 31) / 32 + 1) 32-bit words (pad with zeros).
 * out_bits_d: (n + 31) / 32 words. */
int cuot_qc_sparse(cuot_ctx *ctx, int64_t n, const uint8_t *a_bits_d, const int64_t *pos_d, int64_t w,
                   uint8_t *out_bits_d, void *stream);

/* ---- 128-row bit transpose --------------------------------------------------
 * rows_d[k][m/8] bit m%8 = bit k of block m (bit k = byte k/8, bit k%8), the
 * order of Mat::bit_transpose (gpu/gpu_matrix.cu:95-121, gpu/gpu_ops.cu:31-70).
 * N must be a multiple of 32. */
int cuot_bit_transpose(cuot_ctx *ctx, const cuot_block *in_d, int64_t N, uint8_t *rows_d, void *stream);
int cuot_bit_untranspose(cuot_ctx *ctx, const uint8_t *rows_d, int64_t N, cuot_block *out_d, void *stream);

/* ---- whole instances (both roles in-process) --------------------------------
 * Everything derives from `seed` by the counter-mode PRG (stream ids:
 * 0 Delta (LSB forced 1), 1 roots, 2 punctured indices, 3 circulant a,
 * 4/5 base-OT messages, 6 Ferret LPN seed, 7/8 Ferret bootstrap K and
 * choice bits).  The workspace is caller-owned device memory of at least
 * *_workspace_bytes(cfg) bytes (0 = invalid config).
 *
 * SOT: n COTs from 2n = t*2^h GGM leaves (last tree truncated), QC 2n -> n.
 * Mirrors SOTSender/SOTRecver::{base_ot, seed_expand, dual_lpn}
 * (silent/roles.h:53-85) with base OTs simulated from the seed.
 * Out: v_d (sender, n blocks), w_d (receiver, n blocks), b_d (receiver choice
 * bits, (n+31)/32 words), delta_d (1 block, may be NULL):
 * w_i = v_i ^ b_i * Delta for every i. */
typedef struct {
  int64_t n;
  int64_t t;
  int32_t h;
  int32_t reserved;
  cuot_block k0, k1; /* GGM PRG keys (SilentConfig leftKey/rightKey) */
  cuot_block seed;
} cuot_sot_config;

enum {
  CUOT_SOT_STAGE_SETUP = 1,      /* seed -> Delta, roots, alpha, a, base-OT messages */
  CUOT_SOT_STAGE_QC_PREPARE = 2, /* A^ = transform of a */
  CUOT_SOT_STAGE_GGM_SENDER = 4, /* sender GGM + masked rows + receiver's view */
  CUOT_SOT_STAGE_GGM_RECVER = 8, /* receiver GGM */
  CUOT_SOT_STAGE_QC_SENDER = 16,
  CUOT_SOT_STAGE_QC_RECVER = 32,
  CUOT_SOT_STAGE_SPARSE = 64,
  CUOT_SOT_STAGE_ALL = 127
};

size_t cuot_sot_workspace_bytes(const cuot_sot_config *cfg);
/* Runs the selected stages in order (for per-stage timing; later stages read
 * what earlier ones left in the workspace). */
int cuot_sot_run_stages(cuot_ctx *ctx, const cuot_sot_config *cfg, void *ws_d, size_t ws_bytes, cuot_block *v_d,
                        cuot_block *w_d, uint8_t *b_d, cuot_block *delta_d, uint32_t stages, void *stream);
int cuot_sot_run(cuot_ctx *ctx, const cuot_sot_config *cfg, void *ws_d, size_t ws_bytes, cuot_block *v_d,
                 cuot_block *w_d, uint8_t *b_d, cuot_block *delta_d, void *stream);

/* Ferret: MPCOT over t trees of 2^h leaves truncated to n, then the 10-local
 * LPN with k bootstrap COTs (FerretCOT::extend, ferret_cot.hpp:79-91, with
 * emp-ot's TwoKeyPRP keys).  Out: out_s_d, out_r_d (n blocks each);
 * out_r = out_s ^ lsb(out_r) * Delta. */
typedef struct {
  int64_t n;
  int64_t k;
  int64_t t;
  int32_t h;
  int32_t ntasks;
  cuot_block seed;
  int64_t n_pre; /* sessions only: blocks of emp-ot's ot_pre_data (PrimalLPNParameter::n_pre,
                    constants.h:8-22) carried by the state blob and pre-OT files; 0 = M (see below);
                    otherwise >= M */
} cuot_ferret_config;

enum {
  CUOT_FERRET_STAGE_SETUP = 1,
  CUOT_FERRET_STAGE_GGM_SENDER = 2,
  CUOT_FERRET_STAGE_GGM_RECVER = 4,
  CUOT_FERRET_STAGE_LPN = 8,
  CUOT_FERRET_STAGE_ALL = 15
};

size_t cuot_ferret_workspace_bytes(const cuot_ferret_config *cfg);
int cuot_ferret_run_stages(cuot_ctx *ctx, const cuot_ferret_config *cfg, void *ws_d, size_t ws_bytes,
                           cuot_block *out_s_d, cuot_block *out_r_d, cuot_block *delta_d, uint32_t stages,
                           void *stream);
int cuot_ferret_run(cuot_ctx *ctx, const cuot_ferret_config *cfg, void *ws_d, size_t ws_bytes,
                    cuot_block *out_s_d, cuot_block *out_r_d, cuot_block *delta_d, void *stream);

/* cuot_ferret_run without a workspace argument and without any allocation in
 * the call: the workspace is a library-held device array of
 * CUOT_FERRET_NOALLOC_WORKSPACE bytes (32 MiB; enough for n = 10,168,320,
 * k = 452,160, t = 1,270, h = 13, whose workspace is 28.5 MiB), one per GPU,
 * held until the process ends.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Configs whose
 * cuot_ferret_workspace_bytes exceeds the limit return CUOT_ERR_ARG.  Calls
 * that overlap on one GPU share the array and race: issue them on one stream
 * or serialise them.  Same outputs as cuot_ferret_run. */
#define CUOT_FERRET_NOALLOC_WORKSPACE (32u << 20)
int cuot_ferret_run_noalloc(cuot_ctx *ctx, const cuot_ferret_config *cfg, cuot_block *out_s_d,
                            cuot_block *out_r_d, cuot_block *delta_d, void *stream);

/* ---- drop-in: emp-ot Ferret dev_layer ----------------------------------------
 * Same argument meaning and blocking, host-in/host-out behaviour as
 * ferret/emp-ot/emp-ot/ferret/dev_layer.h:9-18, with Mat* replaced by arrays
 * of device pointers (one per GPU; this build drives ngpu == 1) and emp-ot's
 * semantics (TwoKeyPRP keys, AES(s)^s children, LSB-cleared leaves, secret
 * sum) in place of dev_layer.cu's deviations.  Return CUOT_OK or an error. */
int cuot_setdev(int gpu); /* cuda_setdev, dev_layer.h:9 */

/* cuda_mpcot_sender, dev_layer.h:11-12.  expanded_d[0]: t*2^depth blocks;
 * on entry its first t blocks hold the tree roots (the reference's level loop
 * reads its level-0 input from there, dev_layer.cu:35-38), on return the
 * leaves.  buffer_d/sep_d: caller scratch of t*2^depth blocks
 * (mpcot_reg.h:55-61).  lSum_h/rSum_h[tree*depth + level], secret_sum_h[tree]
 * are written on the host. */
int cuot_mpcot_sender(cuot_ctx *ctx, cuot_block *const *expanded_d, cuot_block *const *buffer_d,
                      cuot_block *const *sep_d, cuot_block *lSum_h, cuot_block *rSum_h, cuot_block *secret_sum_h,
                      const int *t, int depth, const cuot_block *delta_h, int ngpu);
/* cuda_mpcot_recver, dev_layer.h:14-15.  cSum_h[tree*depth+level]: chosen
 * sums; choices_h[tree*depth+level]: bool bytes (mpcot_reg.h:146-157). */
int cuot_mpcot_recver(cuot_ctx *ctx, cuot_block *const *expanded_d, cuot_block *const *buffer_d,
                      cuot_block *const *sep_d, const cuot_block *cSum_h, const cuot_block *secret_sum_h,
                      const int *t, int depth, const uint8_t *choices_h, int ngpu);
/* cuda_primal_lpn, dev_layer.h:17-18.  key: emp AES_KEY.rd_key of PRP(seed)
 * (first 16 bytes = seed); expanded_d[0]: noise in, LPN output out (n[0]
 * blocks); nn_h: host copy of the output; kk_d[0]: device scratch of k blocks
 * receiving kk_h; pubMats_d is not used (indices are never materialised). */
int cuot_primal_lpn(cuot_ctx *ctx, int role, cuot_block *const *pubMats_d, int64_t d, const int64_t *n, int k,
                    const uint32_t *key, cuot_block *const *expanded_d, cuot_block *nn_h, cuot_block *const *kk_d,
                    const cuot_block *kk_h, int ngpu);

/* ---- drop-in: SOT roles (silent/roles.h:15-85) -------------------------------
 * SilentConfig fields; logOT = log2 of the number of GGM leaves (2n), depth =
 * logOT - log2(nTree).  Base OTs are simulated from `seed` (both roles must be
 * created with the same seed).  Call order (silent/main.cu:22-72):
 * sender base_ot, receiver base_ot, sender seed_expand, receiver get_punc_key,
 * receiver seed_expand, dual_lpn on both.  Outputs stay device-resident. */
typedef struct {
  int32_t id;
  int32_t logOT;
  uint64_t nTree;
  uint32_t leftKey[4];
  uint32_t rightKey[4];
  int32_t gpuPerParty;
  int32_t reserved;
} cuot_silent_config;

typedef struct cuot_sot_role cuot_sot_role;

int cuot_sot_sender_new(cuot_ctx *ctx, const cuot_silent_config *cfg, const cuot_block *seed_h,
                        cuot_sot_role **out); /* SOTSender::SOTSender, sender.cu:12-46 */
int cuot_sot_recver_new(cuot_ctx *ctx, const cuot_silent_config *cfg, const cuot_block *seed_h,
                        cuot_sot_role **out); /* SOTRecver::SOTRecver, recver.cu:11-47 */
int cuot_sot_role_free(cuot_sot_role *r);
int cuot_sot_base_ot(cuot_sot_role *r);       /* base_ot, sender.cu:62-74 / recver.cu:64-76 */
int cuot_sot_seed_expand(cuot_sot_role *r);   /* seed_expand, sender.cu:76-108 / recver.cu:124-163 */
int cuot_sot_get_punc_key(cuot_sot_role *recv, const cuot_sot_role *send); /* recver.cu:117-122 */
int cuot_sot_dual_lpn(cuot_sot_role *r);      /* dual_lpn, sender.cu:110-119 / recver.cu:165-174 */
/* fullVector / puncVector (n blocks), choiceVector (receiver), delta (sender) */
int cuot_sot_outputs(const cuot_sot_role *r, cuot_block **vec_d, uint8_t **choice_bits_d, cuot_block **delta_d,
                     int64_t *n);

/* ---- batches of instances ------------------------------------------------------
 * No counterpart in the reference (which runs one instance per process): many
 * SOT and Ferret instances of different sizes run as one launch sequence, the
 * GGM trees, LPN rows and QC transforms of all instances sharing launches.
 * Every instance's outputs are byte-identical to cuot_sot_run / cuot_ferret_run
 * of the same instance.  Life cycle: new (plan, host only) -> bind (workspace
 * and outputs; uploads the launch tables, synchronous) -> run (any number of
 * times, on a stream; CUOT_BATCH_GRAPH replays a HIP graph captured on the
 * first such call) -> free. */
typedef struct {
  int32_t kind;   /* CUOT_MODE_SOT or CUOT_MODE_FERRET */
  int32_t h;      /* tree depth */
  int64_t n;      /* COTs */
  int64_t t;      /* trees */
  int64_t k;      /* Ferret bootstrap COTs (0 for SOT) */
  int32_t ntasks; /* Ferret LPN task partition (emp-ot threads) */
  int32_t client; /* serving only: 0, or a resident Ferret client in [1, ferret_clients] whose
                     iterating session serves this Ferret request (n COTs; h, t, k, seed unused) */
  cuot_block k0, k1; /* SOT GGM keys (ignored for Ferret: TwoKeyPRP keys) */
  cuot_block seed;
} cuot_instance;

enum {
  CUOT_BATCH_STAGE_SETUP = 1,      /* seeds -> Delta, roots, alpha, a, base-OT messages, bootstrap */
  CUOT_BATCH_STAGE_QC_PREPARE = 2, /* A^ of every SOT instance */
  CUOT_BATCH_STAGE_GGM_SENDER = 4, /* sender GGM of all instances + messages */
  CUOT_BATCH_STAGE_GGM_RECVER = 8, /* receiver GGM of all instances */
  CUOT_BATCH_STAGE_QC = 16,        /* QC encode of both SOT parties */
  CUOT_BATCH_STAGE_SPARSE = 32,    /* SOT choice bits */
  CUOT_BATCH_STAGE_LPN = 64,       /* Ferret LPN of both parties */
  CUOT_BATCH_STAGE_ALL = 127,
  CUOT_BATCH_GRAPH = 1 << 16,      /* run flag: replay a captured HIP graph */
  CUOT_BATCH_OVERLAP = 1 << 17     /* run flag (all stages): A^ preparation, sparse encodes and Ferret LPN
                                      on the context's auxiliary stream, forked and joined around the run;
                                      calls on one context come from one host thread at a time */
};

typedef struct cuot_batch cuot_batch;
/* knobs: qc_chunk_bytes caps the shared QC scratch of one encode launch
 * sequence (0 = default). */
int cuot_batch_new(cuot_ctx *ctx, const cuot_instance *inst, int count, size_t qc_chunk_bytes, cuot_batch **out);
size_t cuot_batch_workspace_bytes(const cuot_batch *b);
/* out0[i]: SOT v / Ferret sender (n_i blocks); out1[i]: SOT w / Ferret
 * receiver; bits[i]: SOT choice bits ((n_i+31)/32 words; NULL for Ferret);
 * delta_d: count blocks (may be NULL).  Pointer arrays are host arrays of
 * device pointers. */
int cuot_batch_bind(cuot_batch *b, void *ws_d, size_t ws_bytes, cuot_block *const *out0, cuot_block *const *out1,
                    uint8_t *const *bits, cuot_block *delta_d);
int cuot_batch_run(cuot_batch *b, uint32_t stages, void *stream);
int cuot_batch_free(cuot_batch *b);

/* ---- serving --------------------------------------------------------------------
 * No counterpart in the reference.  A request loop that admits correlation
 * requests (cuot_instance records) against an HBM budget and runs them with
 * continuous batching on several streams: whenever a stream and arena memory
 * are free, the queued requests that fit (arrival order) become one batch.
 * Results are byte-identical to the plain run of each request alone.  With
 * plain = 1 the loop is the FIFO comparator: one stream, one request at a time
 * through cuot_sot_run / cuot_ferret_run. */
typedef struct {
  size_t hbm_budget;     /* bytes of the device arena (outputs + workspaces) */
  int32_t max_batch;     /* requests per batch (0 = 64) */
  int32_t streams;       /* concurrent batches (0 = 2) */
  size_t qc_chunk_bytes; /* cuot_batch_new knob (0 = default) */
  int32_t use_graph;     /* replay batches through captured graphs */
  int32_t plain;         /* FIFO plain comparator */
  int64_t max_batch_cots; /* cap on the COTs of one batch (0 = 2^24); bounds batch duration */
  void *arena_d;          /* optional caller-owned device arena of hbm_budget bytes (NULL: the server
                             allocates and frees its own), e.g. one arena reused by several servers */
  int32_t ferret_clients; /* resident Ferret clients (0 = none): client c holds an iterating Ferret session
                             (Delta, bootstrap COTs and its last extend's outputs stay in HBM between
                             requests; cuot_ferret_session_*) with the parameters of ferret_session and
                             seed ferret_session.seed with lo ^= c.  A Ferret request of client c takes
                             its n COTs as FerretCOT::rcot does (ferret_cot.hpp:171-209): the unused
                             outputs of the client's last extend first, then new extends; requests of
                             one client run in arrival order.  The sessions' device memory counts
                             against hbm_budget. */
  int32_t reserved2;
  cuot_ferret_config ferret_session;
} cuot_server_config;

typedef struct {
  int32_t done;
  int32_t status;
  double t_submit, t_start, t_done; /* host steady-clock seconds */
  cuot_block *out0, *out1;          /* device, valid until release */
  uint8_t *bits;                    /* SOT choice bits (NULL for Ferret) */
  cuot_block *delta;                /* device, one block */
} cuot_server_result;

typedef struct cuot_server cuot_server;
int cuot_server_new(cuot_ctx *ctx, const cuot_server_config *cfg, cuot_server **out);
int cuot_server_submit(cuot_server *s, const cuot_instance *inst, int64_t *ticket);
/* wait != 0 blocks until the request is done.  cuot_server_free fails the
 * requests still queued (status CUOT_ERR_STATE) and wakes their waiters; it
 * must not be called while another thread is inside a cuot_server_* call on
 * the same server. */
int cuot_server_poll(cuot_server *s, int64_t ticket, int wait, cuot_server_result *res);
int cuot_server_release(cuot_server *s, int64_t ticket);
int cuot_server_stats(cuot_server *s, int64_t *batches, size_t *arena_used);
int cuot_server_free(cuot_server *s);
/* Real-time replay of an arrival schedule through a fresh server; latency_s[i]
 * = completion - scheduled arrival of request i.  relation_failures != NULL:
 * every result is checked on the device against its COT relation (SOT
 * w = v ^ b*Delta, Ferret r = s ^ lsb(r)*Delta) before its release, and the
 * number of failing indices over all results is returned there. */
int cuot_serve_replay(cuot_ctx *ctx, const cuot_server_config *cfg, const cuot_instance *reqs,
                      const double *arrival_s, int count, double *latency_s, double *wall_s, int64_t *batches,
                      int64_t *relation_failures);

/* ---- autotuner cost model ---------------------------------------------------------
 * No counterpart in the reference.  A small network (F inputs -> H tanh ->
 * H tanh -> 1) that predicts log kernel time from instance sizes and execution
 * knobs; trained on measured timings with these kernels (mean squared error,
 * Adam), used only to pick knobs, never on the data path.  Parameters are one
 * float32 vector: W1[H][F], b1[H], W2[H][H], b2[H], w3[H], b3. */
int cuot_mlp_param_count(int F, int H); /* -1 if F > 32 or H > 64 */
size_t cuot_mlp_workspace_bytes(int F, int H, int64_t N);
int cuot_mlp_forward(cuot_ctx *ctx, const float *params_d, int F, int H, const float *x_d, int64_t N, float *y_d,
                     void *stream);
/* grad_d[P] = d loss / d params, loss_d[1] = mean (y - target)^2 (deterministic
 * summation order) */
int cuot_mlp_grad(cuot_ctx *ctx, const float *params_d, int F, int H, const float *x_d, const float *target_d,
                  int64_t N, void *ws_d, size_t ws_bytes, float *grad_d, float *loss_d, void *stream);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8); step counts from 1 */
int cuot_mlp_adam(cuot_ctx *ctx, float *params_d, const float *grad_d, float *m_d, float *v_d, int P, int step,
                  float lr, void *stream);

/* ---- execution knobs ---------------------------------------------------------------
 * No counterpart in the reference.  Process-wide execution parameters picked
 * by the autotuner (cuot_amd.tune); 0 selects the built-in rule.  Knobs never
 * change outputs.  Set them while no work is being launched. */
typedef struct {
  int32_t qc_radix; /* FFT layers per upper pass, 2..8 (tiles of one GF(2^16) half hold 256 elements, full ones 128) */
  int32_t qc_cols;  /* FFT columns per tile: 1, 2, 4 or 8 */
  int32_t ggm_dfs;  /* GGM per-lane depth-first levels (valid: h-10 <= L <= min(h-8, 6); else the rule) */
  int32_t lpn_wgs;  /* LPN grid cap */
  int32_t qc_mfma_max; /* QC encodes with 17 <= n <= this run as a 0/1 int8 matrix-core product
                          (at most 2^16; -1: never, 0: the built-in crossover) */
  int32_t aes_impl;    /* cuot_aes_ecb / cuot_prg cipher: 0 rule (T-table), 1 LDS T-table, 2 bitsliced */
} cuot_knob_values;
int cuot_set_knobs(const cuot_knob_values *k);
int cuot_get_knobs(cuot_knob_values *k);
/* the largest n the QC encoder currently runs on the matrix cores (knob
 * qc_mfma_max, or the built-in crossover measured on MI355X when it is 0) */
int64_t cuot_qc_mfma_crossover(void);

/* ---- per-kernel timing ----------------------------------------------------------------
 * No counterpart in the reference.  Between cuot_timing_begin and
 * cuot_timing_end, every instrumented kernel launched from the calling thread
 * is bracketed by HIP events on its launch stream; end synchronizes the
 * device and returns, per launch, the kernel name, its duration and its
 * algorithmic HBM bytes (for roofline reporting). */
typedef struct {
  char kernel[32];
  float ms;
  double bytes;
} cuot_launch_time;
int cuot_timing_begin(cuot_ctx *ctx);
int cuot_timing_end(cuot_ctx *ctx, cuot_launch_time *out, int max_out, int *count);

/* ---- iterating Ferret session ---------------------------------------------------
 * Replaces FerretCOT's extend loop (ferret_cot.hpp:79-103 extend / extend_f2k,
 * 171-209 rcot, preot.h:38-54 OTPre) with both parties on one GPU and the
 * state in HBM: Delta and a bootstrap of M = k + h*t + 128 COTs per party.
 * Each extend turns the first h*t bootstrap COTs into the MPCOT's base OTs
 * through a correlation-robust hash, expands the trees, runs the LPN with
 * kk = bootstrap + 128, writes n COTs per party (the first ot_limit = n - M
 * are the caller's) and keeps the last M as the next bootstrap. */
typedef struct cuot_ferret_session cuot_ferret_session;
int cuot_ferret_session_new(cuot_ctx *ctx, const cuot_ferret_config *cfg, cuot_ferret_session **out);
int cuot_ferret_session_info(const cuot_ferret_session *s, int64_t *M, int64_t *ot_limit, int64_t *round);
int cuot_ferret_session_delta(const cuot_ferret_session *s, cuot_block *delta_d, void *stream);
/* out_s_d, out_r_d: n blocks each (device); out_r = out_s ^ lsb(out_r) * Delta */
int cuot_ferret_session_extend(cuot_ferret_session *s, cuot_block *out_s_d, cuot_block *out_r_d, void *stream);
/* one extend in two calls, for per-stage timing as emp-ot logs it
 * (ferret_cot.hpp:84-92: SeedExp = pre-OT + MPCOT, LPN = LPN + bootstrap
 * refresh): stages 1 then 2 on the same outputs (3 = both = _extend);
 * stage 4 (between 1 and 2) runs the consistency check alone on the current
 * outputs; out of order -> CUOT_ERR_STATE */
int cuot_ferret_session_extend_stages(cuot_ferret_session *s, cuot_block *out_s_d, cuot_block *out_r_d, int stages,
                                      void *stream);
int cuot_ferret_session_free(cuot_ferret_session *s);
/* Malicious mode (MpcotReg::set_malicious, mpcot_reg.h:69-71): every MPCOT is
 * followed by the f2k consistency check of both parties (spcot_sender.h:88-98,
 * spcot_recver.h:97-107, mpcot_reg.h:177-220) with emp-tool's arithmetic
 * (SHA-256 seeds, chi_j = seed^(j+1) and inner products in GF(2^128) mod
 * x^128 + x^7 + x^2 + x + 1, packing of the first 128 bootstrap COTs).
 * Leaves not stored (global index >= n) are left out, and so is the
 * punctured leaf's chi when that leaf is not stored. */
int cuot_ferret_session_set_malicious(cuot_ferret_session *s, int on);
/* the check of the current extend (waits for the device): ok = 1 when
 * ALICE's and BOB's digests agree; x' (128 bytes, BOB's message), digests (32
 * bytes each); CUOT_ERR_STATE when no check ran for the latest extend (the
 * MPCOT stage of an extend discards the previous extend's result) */
int cuot_ferret_session_check(const cuot_ferret_session *s, int *ok, uint8_t *xprime, uint8_t *dig_a,
                              uint8_t *dig_b);

/* ---- online derandomisation of COTs (both parties, asynchronous) -------------------
 * data_s: the sender's COT blocks K, data_r: the receiver's M = K ^ c*Delta;
 * one byte per bit (bool arrays, as the reference's interfaces).
 * cuot_cot_online replaces FerretCOT::online_sender / online_recver
 * (ferret_cot.hpp:277-295, the second half of send_cot / recv_cot): with the
 * receiver's chosen bits b, bo = b ^ LSB(M) (the receiver's message) and
 * K ^= bo*Delta in place, so that M = K ^ b*Delta.
 * cuot_cot_to_ot replaces COT::send / COT::recv (cot.h:16-56) on such COTs:
 * pad[2i] = H_i(K_i) ^ m0_i, pad[2i+1] = H_i(K_i ^ Delta) ^ m1_i (the sender's
 * message), out_r[i] = pad[2i + r_i] ^ H_i(M_i) = m_{r_i, i}.
 * cuot_cot_to_rot replaces COT::send_rot / recv_rot (cot.h:58-91):
 * data0 = H_i(K_i), data1 = H_i(K_i ^ Delta), out_r = H_i(M_i).
 * H_i = emp-tool's MITCCRH (mitccrh.h, emp-tool @44b1dde): AES under the key
 * s ^ makeBlock(key_start + i, 0), XOR the input; s is the sender's random
 * block (COT::send draws it), key_start the keys earlier calls consumed
 * (0 for a fresh hash; emp-tool advances it by 8 per started group of 8). */
int cuot_cot_online(cuot_ctx *ctx, cuot_block *data_s_d, const cuot_block *data_r_d, const uint8_t *b_d,
                    const cuot_block *delta_d, uint8_t *bo_d, int64_t n, void *stream);
int cuot_cot_to_ot(cuot_ctx *ctx, const cuot_block *data_s_d, const cuot_block *data_r_d,
                   const cuot_block *delta_d, cuot_block s, uint64_t key_start, const uint8_t *r_d,
                   const cuot_block *m0_d, const cuot_block *m1_d, cuot_block *pad_d, cuot_block *out_r_d,
                   int64_t n, void *stream);
int cuot_cot_to_rot(cuot_ctx *ctx, const cuot_block *data_s_d, const cuot_block *data_r_d,
                    const cuot_block *delta_d, cuot_block s, uint64_t key_start, cuot_block *data0_d,
                    cuot_block *data1_d, cuot_block *out_r_d, int64_t n, void *stream);

/* ---- session state blob and pre-OT files (emp-ot byte layouts) --------------------
 * party: 1 = ALICE (sender, holds Delta), 2 = BOB (receiver).  The bootstrap of
 * M = k + h*t + 128 COTs is the head of emp-ot's ot_pre_data of P = cfg->n_pre
 * blocks (P = M when cfg->n_pre is 0); blocks [M, P) are written as zeros and
 * ignored on reading (emp-ot's extend reads only the first M).
 * State blob (FerretCOT::assemble_state / disassemble_state / state_size,
 * ferret_cot.hpp:309-347): int64 party, n, t, k; 16-byte Delta (zero for BOB);
 * P blocks, 48 + 16 P bytes.  The session stays usable after assembling.
 * Loading checks the party and (n, t, k) like disassemble_state (mismatch ->
 * CUOT_ERR_ARG) and resumes at `round` (the roots of round r come from the
 * seed's stream 1 at r*t, a counter the emp-ot blob does not carry).
 * Pre-OT file (write_pre_data128_to_file / read_pre_data128_from_file,
 * ferret_cot.hpp:215-247): int64 party; ALICE only: 16-byte Delta; int64 n,
 * t, k; P blocks.  Reading checks party and parameters ("wrong party",
 * "wrong parameters") and removes both files afterwards, as emp-ot does. */
int64_t cuot_ferret_state_size(const cuot_ferret_config *cfg); /* -1 if invalid */
int cuot_ferret_session_assemble_state(const cuot_ferret_session *s, int party, void *data, int64_t size);
int cuot_ferret_session_disassemble_state(cuot_ctx *ctx, const cuot_ferret_config *cfg, int64_t round,
                                          const void *alice, const void *bob, int64_t size,
                                          cuot_ferret_session **out);
int cuot_ferret_session_write_pre_file(const cuot_ferret_session *s, int party, const char *path);
int cuot_ferret_session_read_pre_files(cuot_ctx *ctx, const cuot_ferret_config *cfg, int64_t round,
                                       const char *alice_path, const char *bob_path, cuot_ferret_session **out);

#ifdef __cplusplus
}
#endif
#endif
