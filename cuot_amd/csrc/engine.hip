// engine.hip -- whole SOT and Ferret instances on one GPU, both roles
// in-process, each stage one launch sequence on the caller's stream.
//
// SOT (silent/main.cu:22-72 call order, silent/sender.cu, silent/recver.cu):
//   setup (seed -> Delta, roots, alpha, a, base-OT messages) -> A^ prepare ->
//   sender GGM -> masked base-OT rows -> receiver's view -> receiver GGM ->
//   QC(sender leaves) -> QC(receiver leaves) -> sparse encode of the noise.
//   Result: v (sender), w and b (receiver) with w_i = v_i ^ b_i * Delta.
// Ferret (ferret_cot.hpp:79-91 extend, mpcot_reg.h, lpn_f2.h):
//   setup -> sender MPCOT -> secret sum + receiver's chosen sums -> receiver
//   MPCOT -> bootstrap COTs (K, M = K ^ c*Delta) -> dual-role LPN.
#include <string.h>

#include <mutex>

#include "ctx.h"
#include "engine.h"
#include "kernels.h"

namespace {

struct Carve {
  char* base;
  size_t off = 0;
  template <typename T>
  T* take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    T* p = base ? (T*)(base + off) : nullptr;
    off += count * sizeof(T);
    return p;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// SOT
// ---------------------------------------------------------------------------
size_t SotPlan::carve(void* ws) {
  Carve c{(char*)ws};
  const int64_t t = cfg.t, h = cfg.h, n = cfg.n;
  delta = c.take<blk>(1);
  roots = c.take<blk>(t);
  alpha = c.take<blk>(t);
  choice = c.take<uint8_t>(t * h);
  pos = c.take<int64_t>(t);
  a_words_total = ((n + 127) / 128) * 4 + 4;  // PRG blocks as words + padding
  a_words = c.take<uint32_t>(a_words_total);
  m0b = c.take<blk>(t * h);
  m1b = c.take<blk>(t * h);
  sums = c.take<blk>(2 * t * h);
  M0 = c.take<blk>((h + 1) * t);
  M1 = c.take<blk>((h + 1) * t);
  csum = c.take<blk>(t * h);
  fin = c.take<blk>(t);
  ahat = c.take<uint32_t>(cuot_qc_log_len(n) <= 30 ? (size_t)cuot_qc_ahat_words(n) : 0);
  leaves_s = c.take<blk>(2 * n);
  leaves_r = c.take<blk>(2 * n);
  qc_scratch = shared_qc ? nullptr : c.take<char>(cuot_qc_scratch_bytes_for(n));
  qc_scratch2 = shared_qc ? nullptr : c.take<char>(cuot_qc_scratch_bytes_for(n));
  return c.off + 256;
}

int SotPlan::validate() const {
  CUOT_CHECK_ARG(cfg.n >= 1, "n < 1");
  CUOT_CHECK_ARG(cfg.h >= 1 && cfg.h <= 16, "h must be in [1, 16]");
  CUOT_CHECK_ARG(cfg.t >= 1 && cfg.t < (1ll << 31), "t out of range");
  CUOT_CHECK_ARG((cfg.t << cfg.h) >= 2 * cfg.n, "t * 2^h must cover the 2n leaves");
  CUOT_CHECK_ARG(cuot_qc_log_len(cfg.n) <= 30, "n too large");
  return CUOT_OK;
}

hipError_t SotPlan::run_setup(hipStream_t st) {
  const int64_t t = cfg.t, h = cfg.h, n = cfg.n;
  uint32_t rk[44];
  cuot_aes_expand_host((const uint8_t*)&cfg.seed, rk);
  hipError_t e;
  if ((e = cuot_launch_prg(rk, 0, 0, 1, delta, st))) return e;
  if ((e = cuot_launch_prg(rk, 1, 0, t, roots, st))) return e;
  if ((e = cuot_launch_prg(rk, 2, 0, t, alpha, st))) return e;
  if ((e = cuot_launch_prg(rk, 3, 0, (n + 127) / 128, (blk*)a_words, st))) return e;
  if ((e = cuot_launch_prg(rk, 4, 0, h * t, m0b, st))) return e;
  if ((e = cuot_launch_prg(rk, 5, 0, h * t, m1b, st))) return e;
  return cuot_launch_setup_fix(delta, alpha, t, (int)h, choice, pos, a_words, n, a_words_total, st);
}

hipError_t SotPlan::run_prepare(hipStream_t st) {
  return cuot_launch_qc_prepare((const uint8_t*)a_words, cfg.n, ahat, qc_scratch, st);
}

void SotPlan::ggm_args(GgmKArgs& a) const {
  memset(&a, 0, sizeof(a));
  a.T0 = cuot_device_T0();
  cuot_aes_expand_host((const uint8_t*)&cfg.k0, a.rk);
  cuot_aes_expand_host((const uint8_t*)&cfg.k1, a.rk + 44);
  a.t = cfg.t;
  a.h = cfg.h;
  a.mode = CUOT_MODE_SOT;
  a.n_out = 2 * cfg.n;
}

hipError_t SotPlan::run_expand_sender(hipStream_t st) {
  GgmKArgs a;
  ggm_args(a);
  a.roots = roots;
  a.leaves = leaves_s;
  a.sums = sums;
  hipError_t e;
  if ((e = cuot_launch_ggm(a, false, st))) return e;
  return run_sender_msgs(st);
}

hipError_t SotPlan::run_sender_msgs(hipStream_t st) {
  hipError_t e;
  if ((e = cuot_launch_sot_msgs(sums, m0b, m1b, delta, cfg.t, cfg.h, M0, M1, st))) return e;
  return cuot_launch_sot_recv_prep(choice, M0, M1, m0b, m1b, cfg.t, cfg.h, csum, fin, st);
}

hipError_t SotPlan::run_expand_recver(hipStream_t st) {
  GgmKArgs a;
  ggm_args(a);
  a.leaves = leaves_r;
  a.choice = choice;
  a.csum = csum;
  a.fin = fin;
  return cuot_launch_ggm(a, true, st);
}

hipError_t SotPlan::run_expand(hipStream_t st) {
  hipError_t e = run_expand_sender(st);
  return e ? e : run_expand_recver(st);
}

hipError_t SotPlan::run_lpn(blk* v, blk* w, uint32_t* b, hipStream_t st) {
  hipError_t e;
  if ((e = cuot_launch_qc_encode(ahat, cfg.n, leaves_s, v, qc_scratch, st))) return e;
  if ((e = cuot_launch_qc_encode(ahat, cfg.n, leaves_r, w, qc_scratch, st))) return e;
  return cuot_launch_sparse(a_words, cfg.n, pos, (int)cfg.t, b, st);
}

// ---------------------------------------------------------------------------
// Ferret
// ---------------------------------------------------------------------------
size_t FerretPlan::carve(void* ws) {
  Carve c{(char*)ws};
  const int64_t t = cfg.t, h = cfg.h, k = cfg.k;
  delta = c.take<blk>(1);
  roots = c.take<blk>(t);
  alpha = c.take<blk>(t);
  choice = c.take<uint8_t>(t * h);
  sums = c.take<blk>(2 * t * h);
  csum = c.take<blk>(t * h);
  ssum = c.take<blk>(t);
  lpn_seed = c.take<blk>(1);
  K = c.take<blk>(k);
  Mr = c.take<blk>(k);
  cbits = c.take<blk>((k + 127) / 128);
  kkp = c.take<blk>(2 * k);
  return c.off + 256;
}

int FerretPlan::validate() const {
  CUOT_CHECK_ARG(cfg.n >= 1, "n < 1");
  CUOT_CHECK_ARG(cfg.h >= 1 && cfg.h <= 16, "h must be in [1, 16]");
  CUOT_CHECK_ARG(cfg.t >= 1 && cfg.t < (1ll << 31), "t out of range");
  CUOT_CHECK_ARG((cfg.t << cfg.h) >= cfg.n, "t * 2^h must cover n");
  CUOT_CHECK_ARG(cfg.k >= 1 && cfg.k < (1ll << 31), "k out of range");
  CUOT_CHECK_ARG(cfg.k < cfg.n, "k must be below n (the LPN compresses k bootstrap COTs into n)");
  CUOT_CHECK_ARG(cfg.ntasks >= 1, "ntasks < 1");
  return CUOT_OK;
}

hipError_t FerretPlan::run_setup(hipStream_t st) {
  uint32_t rk[44];
  cuot_aes_expand_host((const uint8_t*)&cfg.seed, rk);
  hipError_t e;
  if ((e = cuot_launch_prg(rk, 0, 0, 1, delta, st))) return e;
  if ((e = cuot_launch_prg(rk, 1, 0, cfg.t, roots, st))) return e;
  if ((e = cuot_launch_prg(rk, 2, 0, cfg.t, alpha, st))) return e;
  if ((e = cuot_launch_prg(rk, 6, 0, 1, lpn_seed, st))) return e;
  if ((e = cuot_launch_prg(rk, 7, 0, cfg.k, K, st))) return e;
  if ((e = cuot_launch_prg(rk, 8, 0, (cfg.k + 127) / 128, cbits, st))) return e;
  if ((e = cuot_launch_setup_fix(delta, alpha, cfg.t, cfg.h, choice, nullptr, nullptr, 0, 0, st))) return e;
  return cuot_launch_ferret_boot(K, Mr, (const uint8_t*)cbits, delta, cfg.k, st);
}

void FerretPlan::ggm_args(GgmKArgs& a) const {
  memset(&a, 0, sizeof(a));
  a.T0 = cuot_device_T0();
  const cuot_block k0 = {0, 0}, k1 = {1, 0};  // TwoKeyPRP(zero_block, makeBlock(0, 1))
  cuot_aes_expand_host((const uint8_t*)&k0, a.rk);
  cuot_aes_expand_host((const uint8_t*)&k1, a.rk + 44);
  a.t = cfg.t;
  a.h = cfg.h;
  a.mode = CUOT_MODE_FERRET;
  a.n_out = cfg.n;
}

hipError_t FerretPlan::run_mpcot_sender(blk* out_s, hipStream_t st) {
  GgmKArgs a;
  ggm_args(a);
  a.roots = roots;
  a.leaves = out_s;
  a.sums = sums;
  hipError_t e;
  if ((e = cuot_launch_ggm(a, false, st))) return e;
  return run_sender_sums(st);
}

hipError_t FerretPlan::run_sender_sums(hipStream_t st) {
  return cuot_launch_ferret_sums(sums, delta, choice, cfg.t, cfg.h, ssum, csum, nullptr, nullptr, st);
}

hipError_t FerretPlan::run_mpcot_recver(blk* out_r, hipStream_t st) {
  GgmKArgs a;
  ggm_args(a);
  a.leaves = out_r;
  a.choice = choice;
  a.csum = csum;
  a.fin = ssum;
  return cuot_launch_ggm(a, true, st);
}

hipError_t FerretPlan::run_mpcot(blk* out_s, blk* out_r, hipStream_t st) {
  hipError_t e = run_mpcot_sender(out_s, st);
  return e ? e : run_mpcot_recver(out_r, st);
}

// The LPN key is the seed block; the host needs it for the AES schedule, so
// the seed is derived on the host side from the instance seed (one AES).
void FerretPlan::lpn_launch(blk* out_s, blk* out_r, LpnLaunch& l) const {
  uint32_t rk_seed[44];
  cuot_aes_expand_host((const uint8_t*)&cfg.seed, rk_seed);
  // LPN seed = AES_seed({0, 6}) computed on the host (identical to stream 6)
  uint8_t ctr[16] = {0};
  ctr[8] = 6;
  uint8_t seed6[16];
  cuot_aes_encrypt_host(rk_seed, ctr, seed6);
  memset(&l, 0, sizeof(l));
  cuot_aes_expand_host(seed6, l.rk);
  l.n = cfg.n;
  l.k = cfg.k;
  l.ntasks = cfg.ntasks;
  l.nroles = 2;
  l.noise[0] = out_s;
  l.kk[0] = K;
  l.out[0] = out_s;
  l.noise[1] = out_r;
  l.kk[1] = Mr;
  l.out[1] = out_r;
  l.kkp = kkp;
}

hipError_t FerretPlan::fill_kkp(hipStream_t st) { return cuot_launch_interleave(K, Mr, cfg.k, kkp, st); }

hipError_t FerretPlan::run_lpn(blk* out_s, blk* out_r, hipStream_t st) {
  hipError_t e = fill_kkp(st);
  if (e) return e;
  LpnLaunch l;
  lpn_launch(out_s, out_r, l);
  return cuot_launch_lpn(l, st);
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" size_t cuot_sot_workspace_bytes(const cuot_sot_config* cfg) {
  if (!cfg) return 0;
  SotPlan p;
  p.cfg = *cfg;
  if (p.validate() != CUOT_OK) return 0;
  return p.carve(nullptr);
}

extern "C" int cuot_sot_run_stages(cuot_ctx* ctx, const cuot_sot_config* cfg, void* ws_d, size_t ws_bytes,
                                   cuot_block* v_d, cuot_block* w_d, uint8_t* b_d, cuot_block* delta_d,
                                   uint32_t stages, void* stream) {
  CUOT_CHECK_ARG(ctx && cfg, "ctx/cfg NULL");
  SotPlan p;
  p.cfg = *cfg;
  int rc = p.validate();
  if (rc) return rc;
  CUOT_CHECK_ARG(ws_d && v_d && w_d && b_d, "buffer is NULL");
  CUOT_CHECK_ARG(ws_bytes >= p.carve(nullptr), "workspace too small (cuot_sot_workspace_bytes)");
  p.carve(ws_d);
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = ctx->st(stream);
  if (stages & CUOT_SOT_STAGE_SETUP) CUOT_HIP_TRY(p.run_setup(st));
  if (stages & CUOT_SOT_STAGE_QC_PREPARE) CUOT_HIP_TRY(p.run_prepare(st));
  if (stages & CUOT_SOT_STAGE_GGM_SENDER) CUOT_HIP_TRY(p.run_expand_sender(st));
  if (stages & CUOT_SOT_STAGE_GGM_RECVER) CUOT_HIP_TRY(p.run_expand_recver(st));
  if (stages & CUOT_SOT_STAGE_QC_SENDER)
    CUOT_HIP_TRY(cuot_launch_qc_encode(p.ahat, p.cfg.n, p.leaves_s, (blk*)v_d, p.qc_scratch, st));
  if (stages & CUOT_SOT_STAGE_QC_RECVER)
    CUOT_HIP_TRY(cuot_launch_qc_encode(p.ahat, p.cfg.n, p.leaves_r, (blk*)w_d, p.qc_scratch, st));
  if (stages & CUOT_SOT_STAGE_SPARSE)
    CUOT_HIP_TRY(cuot_launch_sparse(p.a_words, p.cfg.n, p.pos, (int)p.cfg.t, (uint32_t*)b_d, st));
  if (delta_d && (stages & CUOT_SOT_STAGE_SETUP))
    CUOT_HIP_TRY(hipMemcpyAsync(delta_d, p.delta, 16, hipMemcpyDeviceToDevice, st));
  return CUOT_OK;
}

// Whole instance with the independent stages overlapped on two streams (the
// caller's and the context's second stream, joined before returning):
//   caller: setup -> sender GGM + masked rows -> [A^] -> sender QC -> [join]
//   second: [setup] -> A^ prepare -> sparse -> [rows] -> receiver GGM -> receiver QC
// The two encodes use separate scratch.  Same bytes as the staged run.
extern "C" int cuot_sot_run(cuot_ctx* ctx, const cuot_sot_config* cfg, void* ws_d, size_t ws_bytes, cuot_block* v_d,
                            cuot_block* w_d, uint8_t* b_d, cuot_block* delta_d, void* stream) {
  CUOT_CHECK_ARG(ctx && cfg, "ctx/cfg NULL");
  SotPlan p;
  p.cfg = *cfg;
  int rc = p.validate();
  if (rc) return rc;
  CUOT_CHECK_ARG(ws_d && v_d && w_d && b_d, "buffer is NULL");
  CUOT_CHECK_ARG(ws_bytes >= p.carve(nullptr), "workspace too small (cuot_sot_workspace_bytes)");
  p.carve(ws_d);
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  CUOT_HIP_TRY(ctx->ensure_aux());
  CUOT_HIP_TRY(cuot_qc_configure());
  hipStream_t st = ctx->st(stream), ax = ctx->aux;
  hipEvent_t e_setup = ctx->ev[0], e_prep = ctx->ev[1], e_rows = ctx->ev[2], e_done = ctx->ev[3];
  CUOT_HIP_TRY(p.run_setup(st));
  if (delta_d) CUOT_HIP_TRY(hipMemcpyAsync(delta_d, p.delta, 16, hipMemcpyDeviceToDevice, st));
  CUOT_HIP_TRY(hipEventRecord(e_setup, st));
  CUOT_HIP_TRY(hipStreamWaitEvent(ax, e_setup, 0));
  CUOT_HIP_TRY(p.run_prepare(ax));
  CUOT_HIP_TRY(hipEventRecord(e_prep, ax));
  CUOT_HIP_TRY(cuot_launch_sparse(p.a_words, p.cfg.n, p.pos, (int)p.cfg.t, (uint32_t*)b_d, ax));
  CUOT_HIP_TRY(p.run_expand_sender(st));
  CUOT_HIP_TRY(hipEventRecord(e_rows, st));
  CUOT_HIP_TRY(hipStreamWaitEvent(ax, e_rows, 0));
  CUOT_HIP_TRY(p.run_expand_recver(ax));
  CUOT_HIP_TRY(cuot_launch_qc_encode(p.ahat, p.cfg.n, p.leaves_r, (blk*)w_d, p.qc_scratch2, ax));
  CUOT_HIP_TRY(hipEventRecord(e_done, ax));
  CUOT_HIP_TRY(hipStreamWaitEvent(st, e_prep, 0));
  CUOT_HIP_TRY(cuot_launch_qc_encode(p.ahat, p.cfg.n, p.leaves_s, (blk*)v_d, p.qc_scratch, st));
  CUOT_HIP_TRY(hipStreamWaitEvent(st, e_done, 0));
  return CUOT_OK;
}

extern "C" size_t cuot_ferret_workspace_bytes(const cuot_ferret_config* cfg) {
  if (!cfg) return 0;
  FerretPlan p;
  p.cfg = *cfg;
  if (p.validate() != CUOT_OK) return 0;
  return p.carve(nullptr);
}

extern "C" int cuot_ferret_run_stages(cuot_ctx* ctx, const cuot_ferret_config* cfg, void* ws_d, size_t ws_bytes,
                                      cuot_block* out_s_d, cuot_block* out_r_d, cuot_block* delta_d,
                                      uint32_t stages, void* stream) {
  CUOT_CHECK_ARG(ctx && cfg, "ctx/cfg NULL");
  FerretPlan p;
  p.cfg = *cfg;
  int rc = p.validate();
  if (rc) return rc;
  CUOT_CHECK_ARG(ws_d && out_s_d && out_r_d, "buffer is NULL");
  CUOT_CHECK_ARG(ws_bytes >= p.carve(nullptr), "workspace too small (cuot_ferret_workspace_bytes)");
  p.carve(ws_d);
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = ctx->st(stream);
  if (stages & CUOT_FERRET_STAGE_SETUP) CUOT_HIP_TRY(p.run_setup(st));
  if (stages & CUOT_FERRET_STAGE_GGM_SENDER) CUOT_HIP_TRY(p.run_mpcot_sender((blk*)out_s_d, st));
  if (stages & CUOT_FERRET_STAGE_GGM_RECVER) CUOT_HIP_TRY(p.run_mpcot_recver((blk*)out_r_d, st));
  if (stages & CUOT_FERRET_STAGE_LPN) CUOT_HIP_TRY(p.run_lpn((blk*)out_s_d, (blk*)out_r_d, st));
  if (delta_d && (stages & CUOT_FERRET_STAGE_SETUP))
    CUOT_HIP_TRY(hipMemcpyAsync(delta_d, p.delta, 16, hipMemcpyDeviceToDevice, st));
  return CUOT_OK;
}

extern "C" int cuot_ferret_run(cuot_ctx* ctx, const cuot_ferret_config* cfg, void* ws_d, size_t ws_bytes,
                               cuot_block* out_s_d, cuot_block* out_r_d, cuot_block* delta_d, void* stream) {
  return cuot_ferret_run_stages(ctx, cfg, ws_d, ws_bytes, out_s_d, out_r_d, delta_d, CUOT_FERRET_STAGE_ALL, stream);
}

// ---------------------------------------------------------------------------
// Ferret with library-held scratch (no allocation in the call)
// ---------------------------------------------------------------------------
// The workspace of cuot_ferret_run lives in this module-scope device array:
// every GPU the library runs on holds its own copy until the process ends.
// Host code never uses the array's name as an address: cuot_noalloc_setup
// (called by cuot_init with the device current) takes each device's address
// once with hipGetSymbolAddress and every later use goes through that
// pointer.  All calls on one GPU share the array, so calls that overlap on
// one GPU race; every run writes each workspace word before it reads it
// (seed PRG streams, bootstrap, level sums, interleaved kk), so nothing has to
// start at zero.
static __device__ __align__(256) unsigned char scratch[CUOT_FERRET_NOALLOC_WORKSPACE];
static void* g_scratch_dev[64];
static std::mutex g_scratch_mu;

int cuot_noalloc_setup(int device) {
  CUOT_CHECK_ARG(device >= 0 && device < 64, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return CUOT_OK;
  void* p = nullptr;
  CUOT_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  CUOT_CHECK_ARG(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
  g_scratch_dev[device] = p;
  return CUOT_OK;
}

extern "C" int cuot_ferret_run_noalloc(cuot_ctx* ctx, const cuot_ferret_config* cfg, cuot_block* out_s_d,
                                       cuot_block* out_r_d, cuot_block* delta_d, void* stream) {
  CUOT_CHECK_ARG(ctx && cfg, "ctx/cfg NULL");
  FerretPlan p;
  p.cfg = *cfg;
  const int rc = p.validate();
  if (rc) return rc;
  CUOT_CHECK_ARG(p.carve(nullptr) <= (size_t)CUOT_FERRET_NOALLOC_WORKSPACE,
                 "workspace of this config (cuot_ferret_workspace_bytes) exceeds the library scratch of "
                 "CUOT_FERRET_NOALLOC_WORKSPACE bytes");
  CUOT_CHECK_ARG(ctx->device >= 0 && ctx->device < 64, "device index out of range");
  void* ws = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ws = g_scratch_dev[ctx->device];
  }
  CUOT_CHECK_ARG(ws != nullptr, "scratch address not set up on this device (cuot_init)");
  return cuot_ferret_run_stages(ctx, cfg, ws, (size_t)CUOT_FERRET_NOALLOC_WORKSPACE, out_s_d, out_r_d, delta_d,
                                CUOT_FERRET_STAGE_ALL, stream);
}
