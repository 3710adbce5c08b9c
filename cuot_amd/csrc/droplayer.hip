// droplayer.hip -- drop-in replacements for the reference's GPU entry points,
// keeping their argument meaning and blocking host-in/host-out behaviour:
//   emp-ot Ferret dev_layer (ferret/emp-ot/emp-ot/ferret/dev_layer.h:9-18):
//     cuda_setdev, cuda_mpcot_sender, cuda_mpcot_recver, cuda_primal_lpn
//   SOT roles (silent/roles.h:15-85): SOTSender / SOTRecver with
//     base_ot, seed_expand, get_punc_key, dual_lpn.
// The Ferret calls follow emp-ot's semantics (spcot_sender.h, spcot_recver.h,
// lpn_f2.h) rather than the reference GPU code's deviations (uninitialised
// roots, missing feed-forward, unfilled secret sum; dev_layer.cu:20-186).
#include <string.h>

#include <vector>

#include "ctx.h"
#include "engine.h"
#include "kernels.h"

// ---------------------------------------------------------------------------
// Ferret dev_layer
// ---------------------------------------------------------------------------
extern "C" int cuot_setdev(int gpu) {
  CUOT_HIP_TRY(hipSetDevice(gpu));
  return CUOT_OK;
}

static int check_single_gpu(int ngpu, const int* t) {
  CUOT_CHECK_ARG(ngpu == 1, "this build drives one GPU per process (ngpu must be 1)");
  CUOT_CHECK_ARG(t && t[0] >= 1, "t[0] must be >= 1");
  return CUOT_OK;
}

static void ferret_ggm_args(GgmKArgs& a, int64_t t, int depth) {
  memset(&a, 0, sizeof(a));
  a.T0 = cuot_device_T0();
  const cuot_block k0 = {0, 0}, k1 = {1, 0};  // TwoKeyPRP(zero_block, makeBlock(0, 1))
  cuot_aes_expand_host((const uint8_t*)&k0, a.rk);
  cuot_aes_expand_host((const uint8_t*)&k1, a.rk + 44);
  a.t = t;
  a.h = depth;
  a.mode = CUOT_MODE_FERRET;
  a.n_out = t << depth;
}

extern "C" int cuot_mpcot_sender(cuot_ctx* ctx, cuot_block* const* expanded_d, cuot_block* const* buffer_d,
                                 cuot_block* const* sep_d, cuot_block* lSum_h, cuot_block* rSum_h,
                                 cuot_block* secret_sum_h, const int* t, int depth, const cuot_block* delta_h,
                                 int ngpu) {
  CUOT_CHECK_ARG(ctx, "ctx is NULL");
  int rc = check_single_gpu(ngpu, t);
  if (rc) return rc;
  CUOT_CHECK_ARG(depth >= 1 && depth <= 16, "depth must be in [1, 16]");
  CUOT_CHECK_ARG(expanded_d && buffer_d && sep_d && expanded_d[0] && buffer_d[0] && sep_d[0], "device buffer NULL");
  CUOT_CHECK_ARG(lSum_h && rSum_h && secret_sum_h && delta_h, "host buffer NULL");
  const int64_t T = t[0];
  CUOT_CHECK_ARG((T << depth) >= 4 * T * depth + T + 1, "buffer too small for scratch");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  // scratch carved from buffer (t * 2^depth blocks, mpcot_reg.h:55-61):
  blk* roots = (blk*)buffer_d[0];
  blk* sums = roots + T;             // [t][depth][2]
  blk* lsum = sums + 2 * T * depth;  // [t][depth]
  blk* rsum = lsum + T * depth;
  blk* ssum = (blk*)sep_d[0];
  blk* delta = ssum + T;
  // roots: the first t blocks of expanded (where the reference's level loop
  // reads its level-0 input, dev_layer.cu:35-38); copied out before the
  // leaves overwrite them
  CUOT_HIP_TRY(hipMemcpyAsync(roots, expanded_d[0], T * 16, hipMemcpyDeviceToDevice, st));
  CUOT_HIP_TRY(hipMemcpyAsync(delta, delta_h, 16, hipMemcpyHostToDevice, st));
  GgmKArgs a;
  ferret_ggm_args(a, T, depth);
  a.roots = roots;
  a.leaves = (blk*)expanded_d[0];
  a.sums = sums;
  CUOT_HIP_TRY(cuot_launch_ggm(a, false, st));
  CUOT_HIP_TRY(cuot_launch_ferret_sums(sums, delta, nullptr, T, depth, ssum, nullptr, lsum, rsum, st));
  CUOT_HIP_TRY(hipMemcpyAsync(lSum_h, lsum, T * depth * 16, hipMemcpyDeviceToHost, st));
  CUOT_HIP_TRY(hipMemcpyAsync(rSum_h, rsum, T * depth * 16, hipMemcpyDeviceToHost, st));
  CUOT_HIP_TRY(hipMemcpyAsync(secret_sum_h, ssum, T * 16, hipMemcpyDeviceToHost, st));
  CUOT_HIP_TRY(hipStreamSynchronize(st));
  return CUOT_OK;
}

extern "C" int cuot_mpcot_recver(cuot_ctx* ctx, cuot_block* const* expanded_d, cuot_block* const* buffer_d,
                                 cuot_block* const* sep_d, const cuot_block* cSum_h, const cuot_block* secret_sum_h,
                                 const int* t, int depth, const uint8_t* choices_h, int ngpu) {
  CUOT_CHECK_ARG(ctx, "ctx is NULL");
  int rc = check_single_gpu(ngpu, t);
  if (rc) return rc;
  CUOT_CHECK_ARG(depth >= 1 && depth <= 16, "depth must be in [1, 16]");
  CUOT_CHECK_ARG(expanded_d && buffer_d && sep_d && expanded_d[0] && buffer_d[0] && sep_d[0], "device buffer NULL");
  CUOT_CHECK_ARG(cSum_h && secret_sum_h && choices_h, "host buffer NULL");
  const int64_t T = t[0];
  CUOT_CHECK_ARG((T << depth) >= 2 * T * depth + T + 1, "buffer too small for scratch");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  blk* csum = (blk*)buffer_d[0];
  blk* fin = csum + T * depth;
  uint8_t* choice = (uint8_t*)sep_d[0];
  CUOT_HIP_TRY(hipMemcpyAsync(csum, cSum_h, T * depth * 16, hipMemcpyHostToDevice, st));
  CUOT_HIP_TRY(hipMemcpyAsync(fin, secret_sum_h, T * 16, hipMemcpyHostToDevice, st));
  CUOT_HIP_TRY(hipMemcpyAsync(choice, choices_h, T * depth, hipMemcpyHostToDevice, st));
  GgmKArgs a;
  ferret_ggm_args(a, T, depth);
  a.leaves = (blk*)expanded_d[0];
  a.choice = choice;
  a.csum = csum;
  a.fin = fin;
  CUOT_HIP_TRY(cuot_launch_ggm(a, true, st));
  CUOT_HIP_TRY(hipStreamSynchronize(st));
  return CUOT_OK;
}

extern "C" int cuot_primal_lpn(cuot_ctx* ctx, int role, cuot_block* const* pubMats_d, int64_t d, const int64_t* n,
                               int k, const uint32_t* key, cuot_block* const* expanded_d, cuot_block* nn_h,
                               cuot_block* const* kk_d, const cuot_block* kk_h, int ngpu) {
  (void)role;
  (void)pubMats_d;  // the index matrix is never materialised
  CUOT_CHECK_ARG(ctx, "ctx is NULL");
  CUOT_CHECK_ARG(ngpu == 1, "this build drives one GPU per process (ngpu must be 1)");
  CUOT_CHECK_ARG(d == 10, "only d = 10 is supported (LpnF2<IO, 10>)");
  CUOT_CHECK_ARG(n && n[0] >= 1, "n[0] must be >= 1");
  CUOT_CHECK_ARG(k >= 1, "k must be >= 1");
  CUOT_CHECK_ARG(key && expanded_d && expanded_d[0] && kk_d && kk_d[0] && kk_h && nn_h, "buffer NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  LpnLaunch l;
  memset(&l, 0, sizeof(l));
  // key = AES_KEY.rd_key of PRP(seed): its first 16 bytes are the seed itself
  cuot_aes_expand_host((const uint8_t*)key, l.rk);
  l.n = n[0];
  l.k = k;
  l.ntasks = 1;
  l.nroles = 1;
  CUOT_HIP_TRY(hipMemcpyAsync(kk_d[0], kk_h, (size_t)k * 16, hipMemcpyHostToDevice, st));
  l.noise[0] = (const blk*)expanded_d[0];
  l.kk[0] = (const blk*)kk_d[0];
  l.out[0] = (blk*)expanded_d[0];
  CUOT_HIP_TRY(cuot_launch_lpn(l, st));
  CUOT_HIP_TRY(hipMemcpyAsync(nn_h, expanded_d[0], (size_t)n[0] * 16, hipMemcpyDeviceToHost, st));
  CUOT_HIP_TRY(hipStreamSynchronize(st));
  return CUOT_OK;
}

// ---------------------------------------------------------------------------
// SOT roles
// ---------------------------------------------------------------------------
struct cuot_sot_role {
  cuot_ctx* ctx = nullptr;
  int role = 0;
  SotPlan plan;
  void* ws = nullptr;
  blk* out = nullptr;     // sender fullVector / receiver puncVector (n blocks)
  uint32_t* bits = nullptr;  // receiver choiceVector ((n+31)/32 words)
  int stage = 0;          // 0 new, 1 base_ot, 2 punc key, 3 expanded, 4 lpn
};

static int role_new(cuot_ctx* ctx, const cuot_silent_config* cfg, const cuot_block* seed_h, int role,
                    cuot_sot_role** out) {
  CUOT_CHECK_ARG(ctx && cfg && seed_h && out, "NULL argument");
  CUOT_CHECK_ARG(cfg->logOT >= 2 && cfg->logOT <= 34, "logOT out of range");
  CUOT_CHECK_ARG(cfg->nTree >= 1 && (cfg->nTree & (cfg->nTree - 1)) == 0, "nTree must be a power of two");
  const int lt = __builtin_ctzll(cfg->nTree);
  CUOT_CHECK_ARG(cfg->logOT - lt >= 1 && cfg->logOT - lt <= 16, "tree depth logOT - log2(nTree) must be in [1, 16]");
  CUOT_CHECK_ARG(cfg->gpuPerParty == 1, "one GPU per party in this build");
  cuot_sot_role* r = new cuot_sot_role();
  r->ctx = ctx;
  r->role = role;
  r->plan.cfg.n = (int64_t)1 << (cfg->logOT - 1);
  r->plan.cfg.t = (int64_t)cfg->nTree;
  r->plan.cfg.h = cfg->logOT - lt;
  r->plan.cfg.reserved = 0;
  memcpy(&r->plan.cfg.k0, cfg->leftKey, 16);
  memcpy(&r->plan.cfg.k1, cfg->rightKey, 16);
  r->plan.cfg.seed = *seed_h;
  int rc = r->plan.validate();
  if (rc) {
    delete r;
    return rc;
  }
  if (hipSetDevice(ctx->device) != hipSuccess) {
    delete r;
    cuot_set_last_error("hipSetDevice failed", __FILE__, __LINE__);
    return CUOT_ERR_HIP;
  }
  const size_t bytes = r->plan.carve(nullptr);
  const int64_t n = r->plan.cfg.n;
  if (hipMalloc(&r->ws, bytes) != hipSuccess || hipMalloc((void**)&r->out, (size_t)n * 16) != hipSuccess ||
      hipMalloc((void**)&r->bits, (size_t)((n + 31) / 32) * 4) != hipSuccess) {
    hipFree(r->ws);
    hipFree(r->out);
    hipFree(r->bits);
    delete r;
    cuot_set_last_error("device allocation failed", __FILE__, __LINE__);
    return CUOT_ERR_NOMEM;
  }
  r->plan.carve(r->ws);
  *out = r;
  return CUOT_OK;
}

extern "C" int cuot_sot_sender_new(cuot_ctx* ctx, const cuot_silent_config* cfg, const cuot_block* seed_h,
                                   cuot_sot_role** out) {
  return role_new(ctx, cfg, seed_h, CUOT_ROLE_SENDER, out);
}

extern "C" int cuot_sot_recver_new(cuot_ctx* ctx, const cuot_silent_config* cfg, const cuot_block* seed_h,
                                   cuot_sot_role** out) {
  return role_new(ctx, cfg, seed_h, CUOT_ROLE_RECVER, out);
}

extern "C" int cuot_sot_role_free(cuot_sot_role* r) {
  if (!r) return CUOT_OK;
  hipSetDevice(r->ctx->device);
  hipFree(r->ws);
  hipFree(r->out);
  hipFree(r->bits);
  delete r;
  return CUOT_OK;
}

// base_ot: the base OTs are simulated from the shared seed (both roles derive
// the same messages and the receiver's choices), then the code's A^ is
// prepared (QuasiCyclic constructor, silent/quasi_cyclic.cu:51-106).
extern "C" int cuot_sot_base_ot(cuot_sot_role* r) {
  CUOT_CHECK_ARG(r, "role NULL");
  CUOT_HIP_TRY(hipSetDevice(r->ctx->device));
  hipStream_t st = r->ctx->stream;
  CUOT_HIP_TRY(r->plan.run_setup(st));
  CUOT_HIP_TRY(r->plan.run_prepare(st));
  CUOT_HIP_TRY(hipStreamSynchronize(st));
  r->stage = 1;
  return CUOT_OK;
}

extern "C" int cuot_sot_seed_expand(cuot_sot_role* r) {
  CUOT_CHECK_ARG(r, "role NULL");
  CUOT_HIP_TRY(hipSetDevice(r->ctx->device));
  hipStream_t st = r->ctx->stream;
  if (r->role == CUOT_ROLE_SENDER) {
    if (r->stage < 1) {
      cuot_set_last_error("seed_expand before base_ot", __FILE__, __LINE__);
      return CUOT_ERR_STATE;
    }
    CUOT_HIP_TRY(r->plan.run_expand_sender(st));
  } else {
    if (r->stage < 2) {
      cuot_set_last_error("receiver seed_expand before get_punc_key", __FILE__, __LINE__);
      return CUOT_ERR_STATE;
    }
    CUOT_HIP_TRY(r->plan.run_expand_recver(st));
  }
  CUOT_HIP_TRY(hipStreamSynchronize(st));
  r->stage = 3;
  return CUOT_OK;
}

// Receiver obtains the sender's masked level-sum rows (SOTRecver::get_punc_key,
// silent/recver.cu:117-122) and forms its view: chosen sums and final row.
extern "C" int cuot_sot_get_punc_key(cuot_sot_role* recv, const cuot_sot_role* send) {
  CUOT_CHECK_ARG(recv && send, "role NULL");
  CUOT_CHECK_ARG(recv->role == CUOT_ROLE_RECVER && send->role == CUOT_ROLE_SENDER, "wrong roles");
  if (send->stage < 3 || recv->stage < 1) {
    cuot_set_last_error("get_punc_key needs sender seed_expand and receiver base_ot", __FILE__, __LINE__);
    return CUOT_ERR_STATE;
  }
  const SotPlan& s = send->plan;
  SotPlan& p = recv->plan;
  CUOT_CHECK_ARG(s.cfg.t == p.cfg.t && s.cfg.h == p.cfg.h && s.cfg.n == p.cfg.n, "configs differ");
  CUOT_HIP_TRY(hipSetDevice(recv->ctx->device));
  hipStream_t st = recv->ctx->stream;
  const size_t rows = (size_t)(p.cfg.h + 1) * p.cfg.t * 16;
  CUOT_HIP_TRY(hipMemcpyAsync(p.M0, s.M0, rows, hipMemcpyDeviceToDevice, st));
  CUOT_HIP_TRY(hipMemcpyAsync(p.M1, s.M1, rows, hipMemcpyDeviceToDevice, st));
  CUOT_HIP_TRY(cuot_launch_sot_recv_prep(p.choice, p.M0, p.M1, p.m0b, p.m1b, p.cfg.t, p.cfg.h, p.csum, p.fin, st));
  CUOT_HIP_TRY(hipStreamSynchronize(st));
  recv->stage = 2;
  return CUOT_OK;
}

extern "C" int cuot_sot_dual_lpn(cuot_sot_role* r) {
  CUOT_CHECK_ARG(r, "role NULL");
  if (r->stage < 3) {
    cuot_set_last_error("dual_lpn before seed_expand", __FILE__, __LINE__);
    return CUOT_ERR_STATE;
  }
  CUOT_HIP_TRY(hipSetDevice(r->ctx->device));
  hipStream_t st = r->ctx->stream;
  SotPlan& p = r->plan;
  if (r->role == CUOT_ROLE_SENDER) {
    CUOT_HIP_TRY(cuot_launch_qc_encode(p.ahat, p.cfg.n, p.leaves_s, r->out, p.qc_scratch, st));
  } else {
    CUOT_HIP_TRY(cuot_launch_qc_encode(p.ahat, p.cfg.n, p.leaves_r, r->out, p.qc_scratch, st));
    CUOT_HIP_TRY(cuot_launch_sparse(p.a_words, p.cfg.n, p.pos, (int)p.cfg.t, r->bits, st));
  }
  CUOT_HIP_TRY(hipStreamSynchronize(st));
  r->stage = 4;
  return CUOT_OK;
}

// Results: sender fullVector / receiver puncVector (n blocks), the receiver's
// choiceVector (bits), Delta (sender), all device pointers owned by the role.
extern "C" int cuot_sot_outputs(const cuot_sot_role* r, cuot_block** vec_d, uint8_t** choice_bits_d,
                                cuot_block** delta_d, int64_t* n) {
  CUOT_CHECK_ARG(r, "role NULL");
  if (vec_d) *vec_d = (cuot_block*)r->out;
  if (choice_bits_d) *choice_bits_d = r->role == CUOT_ROLE_RECVER ? (uint8_t*)r->bits : nullptr;
  if (delta_d) *delta_d = r->role == CUOT_ROLE_SENDER ? (cuot_block*)r->plan.delta : nullptr;
  if (n) *n = r->plan.cfg.n;
  return CUOT_OK;
}
