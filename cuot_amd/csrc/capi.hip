// capi.hip -- extern "C" entry points: context, errors, AES and GGM stages.
#include <stdio.h>
#include <string.h>

#include <string>

#include "ctx.h"
#include "kernels.h"

static thread_local std::string g_last_error;

void cuot_set_last_error(const char* msg, const char* file, int line) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%s (%s:%d)", msg ? msg : "?", file, line);
  g_last_error = buf;
}

hipError_t cuot_upload_tables(int device);
int cuot_noalloc_setup(int device);  // engine.hip: per-device address of the library scratch

extern "C" const char* cuot_last_error(void) { return g_last_error.c_str(); }

extern "C" int cuot_init(int device, cuot_ctx** out) {
  CUOT_CHECK_ARG(out != nullptr, "out is NULL");
  int n = 0;
  CUOT_HIP_TRY(hipGetDeviceCount(&n));
  CUOT_CHECK_ARG(device >= 0 && device < n, "device index out of range");
  CUOT_HIP_TRY(hipSetDevice(device));
  CUOT_HIP_TRY(cuot_upload_tables(device));
  {
    const int rc = cuot_noalloc_setup(device);
    if (rc) return rc;
  }
  cuot_ctx* c = new cuot_ctx();
  c->device = device;
  c->stream = nullptr;  // legacy default stream unless the caller passes one
  if (c->ensure_streams() != hipSuccess) {
    c->release_all();
    delete c;
    cuot_set_last_error("stream creation failed", __FILE__, __LINE__);
    return CUOT_ERR_HIP;
  }
  *out = c;
  return CUOT_OK;
}

extern "C" int cuot_destroy(cuot_ctx* ctx) {
  if (!ctx) return CUOT_OK;
  hipSetDevice(ctx->device);
  ctx->release_all();
  delete ctx;
  return CUOT_OK;
}

extern "C" int cuot_device_sync(cuot_ctx* ctx) {
  CUOT_CHECK_ARG(ctx, "ctx is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  CUOT_HIP_TRY(hipDeviceSynchronize());
  return CUOT_OK;
}

extern "C" int cuot_aes_ecb(cuot_ctx* ctx, const cuot_block* key_h, const cuot_block* in_d, cuot_block* out_d,
                            int64_t n, void* stream) {
  CUOT_CHECK_ARG(ctx && key_h, "ctx/key is NULL");
  CUOT_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return CUOT_OK;
  CUOT_CHECK_ARG(in_d && out_d, "buffer is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  uint32_t rk[44];
  cuot_aes_expand_host((const uint8_t*)key_h, rk);
  CUOT_HIP_TRY(cuot_launch_aes_ecb(rk, (const blk*)in_d, (blk*)out_d, n, ctx->st(stream)));
  return CUOT_OK;
}

extern "C" int cuot_prg(cuot_ctx* ctx, const cuot_block* seed_h, uint64_t stream_id, uint64_t start, int64_t n,
                        cuot_block* out_d, void* stream) {
  CUOT_CHECK_ARG(ctx && seed_h, "ctx/seed is NULL");
  CUOT_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return CUOT_OK;
  CUOT_CHECK_ARG(out_d, "out is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  uint32_t rk[44];
  cuot_aes_expand_host((const uint8_t*)seed_h, rk);
  CUOT_HIP_TRY(cuot_launch_prg(rk, stream_id, start, n, (blk*)out_d, ctx->st(stream)));
  return CUOT_OK;
}

static int ggm_check(const cuot_ggm_params* p) {
  CUOT_CHECK_ARG(p != nullptr, "params NULL");
  CUOT_CHECK_ARG(p->t >= 1 && p->t <= (1ll << 31) - 1, "t out of range");
  CUOT_CHECK_ARG(p->h >= 1 && p->h <= 16, "h must be in [1, 16]");
  CUOT_CHECK_ARG(p->mode == CUOT_MODE_SOT || p->mode == CUOT_MODE_FERRET, "bad mode");
  CUOT_CHECK_ARG(p->n_out >= 0, "n_out < 0");
  return CUOT_OK;
}

static void ggm_fill(GgmKArgs& a, const cuot_ggm_params* p) {
  memset(&a, 0, sizeof(a));
  a.T0 = cuot_device_T0();
  cuot_aes_expand_host((const uint8_t*)&p->k0, a.rk);
  cuot_aes_expand_host((const uint8_t*)&p->k1, a.rk + 44);
  a.t = p->t;
  a.h = p->h;
  a.mode = p->mode;
  int64_t total = p->t << p->h;
  a.n_out = p->n_out < total ? p->n_out : total;
}

extern "C" int cuot_ggm_sender(cuot_ctx* ctx, const cuot_ggm_params* p, const cuot_block* roots_d,
                               cuot_block* leaves_d, cuot_block* sums_d, void* stream) {
  CUOT_CHECK_ARG(ctx, "ctx is NULL");
  int rc = ggm_check(p);
  if (rc) return rc;
  CUOT_CHECK_ARG(roots_d && sums_d && (leaves_d || p->n_out == 0), "buffer is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  GgmKArgs a;
  ggm_fill(a, p);
  a.roots = (const blk*)roots_d;
  a.leaves = (blk*)leaves_d;
  a.sums = (blk*)sums_d;
  CUOT_HIP_TRY(cuot_launch_ggm(a, false, ctx->st(stream)));
  return CUOT_OK;
}

extern "C" size_t cuot_ggm_sender_bs_work_bytes(const cuot_ggm_params* p) {
  if (ggm_check(p)) return 0;
  return cuot_ggm_bs_work_bytes(p->t, p->h);
}

extern "C" int cuot_ggm_sender_bs(cuot_ctx* ctx, const cuot_ggm_params* p, const cuot_block* roots_d,
                                  cuot_block* leaves_d, cuot_block* sums_d, void* work_d, size_t work_bytes,
                                  void* stream) {
  CUOT_CHECK_ARG(ctx, "ctx is NULL");
  int rc = ggm_check(p);
  if (rc) return rc;
  CUOT_CHECK_ARG(roots_d && sums_d && work_d && (leaves_d || p->n_out == 0), "buffer is NULL");
  CUOT_CHECK_ARG(work_bytes >= cuot_ggm_bs_work_bytes(p->t, p->h), "work buffer too small");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  GgmKArgs a;
  ggm_fill(a, p);
  CUOT_HIP_TRY(cuot_launch_ggm_sender_bs(a.rk, a.t, a.h, a.mode == CUOT_MODE_FERRET, a.n_out,
                                         (const blk*)roots_d, (blk*)leaves_d, (blk*)sums_d, work_d,
                                         ctx->st(stream)));
  return CUOT_OK;
}

extern "C" size_t cuot_ggm_recver_bs_work_bytes(const cuot_ggm_params* p) {
  if (ggm_check(p)) return 0;
  return cuot_ggm_recver_bs_work_bytes(p->t, p->h);
}

extern "C" int cuot_ggm_recver_bs(cuot_ctx* ctx, const cuot_ggm_params* p, const uint8_t* choice_d,
                                  const cuot_block* csum_d, const cuot_block* fin_d, cuot_block* leaves_d,
                                  void* work_d, size_t work_bytes, void* stream) {
  CUOT_CHECK_ARG(ctx, "ctx is NULL");
  int rc = ggm_check(p);
  if (rc) return rc;
  CUOT_CHECK_ARG(choice_d && csum_d && fin_d && work_d && (leaves_d || p->n_out == 0), "buffer is NULL");
  CUOT_CHECK_ARG(work_bytes >= cuot_ggm_recver_bs_work_bytes(p->t, p->h), "work buffer too small");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  GgmKArgs a;
  ggm_fill(a, p);
  CUOT_HIP_TRY(cuot_launch_ggm_recver_bs(a.rk, a.t, a.h, a.mode == CUOT_MODE_FERRET, a.n_out, choice_d,
                                         (const blk*)csum_d, (const blk*)fin_d, (blk*)leaves_d, work_d,
                                         ctx->st(stream)));
  return CUOT_OK;
}

extern "C" int cuot_ggm_recver(cuot_ctx* ctx, const cuot_ggm_params* p, const uint8_t* choice_d,
                               const cuot_block* csum_d, const cuot_block* fin_d, cuot_block* leaves_d,
                               void* stream) {
  CUOT_CHECK_ARG(ctx, "ctx is NULL");
  int rc = ggm_check(p);
  if (rc) return rc;
  CUOT_CHECK_ARG(choice_d && csum_d && fin_d && (leaves_d || p->n_out == 0), "buffer is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  GgmKArgs a;
  ggm_fill(a, p);
  a.choice = choice_d;
  a.csum = (const blk*)csum_d;
  a.fin = (const blk*)fin_d;
  a.leaves = (blk*)leaves_d;
  CUOT_HIP_TRY(cuot_launch_ggm(a, true, ctx->st(stream)));
  return CUOT_OK;
}

static int lpn_prepare(cuot_ctx* ctx, const cuot_lpn_params* p, LpnLaunch& l) {
  CUOT_CHECK_ARG(ctx && p, "ctx/params NULL");
  CUOT_CHECK_ARG(p->d == 10, "only d = 10 is supported (LpnF2<IO, 10>)");
  CUOT_CHECK_ARG(p->n >= 0, "n < 0");
  CUOT_CHECK_ARG(p->k >= 1 && p->k < (1ll << 31), "k out of range");
  CUOT_CHECK_ARG(p->ntasks >= 1, "ntasks < 1");
  memset(&l, 0, sizeof(l));
  cuot_aes_expand_host((const uint8_t*)&p->seed, l.rk);
  l.n = p->n;
  l.k = p->k;
  l.ntasks = p->ntasks;
  return CUOT_OK;
}

extern "C" int cuot_lpn(cuot_ctx* ctx, const cuot_lpn_params* p, const cuot_block* noise_d, const cuot_block* kk_d,
                        cuot_block* out_d, void* stream) {
  LpnLaunch l;
  int rc = lpn_prepare(ctx, p, l);
  if (rc) return rc;
  if (p->n == 0) return CUOT_OK;
  CUOT_CHECK_ARG(noise_d && kk_d && out_d, "buffer is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  l.nroles = 1;
  l.noise[0] = (const blk*)noise_d;
  l.kk[0] = (const blk*)kk_d;
  l.out[0] = (blk*)out_d;
  CUOT_HIP_TRY(cuot_launch_lpn(l, ctx->st(stream)));
  return CUOT_OK;
}

extern "C" int cuot_lpn_dual(cuot_ctx* ctx, const cuot_lpn_params* p, const cuot_block* noise_s_d,
                             const cuot_block* kk_s_d, cuot_block* out_s_d, const cuot_block* noise_r_d,
                             const cuot_block* kk_r_d, cuot_block* out_r_d, void* stream) {
  LpnLaunch l;
  int rc = lpn_prepare(ctx, p, l);
  if (rc) return rc;
  if (p->n == 0) return CUOT_OK;
  CUOT_CHECK_ARG(noise_s_d && kk_s_d && out_s_d && noise_r_d && kk_r_d && out_r_d, "buffer is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  l.nroles = 2;
  l.noise[0] = (const blk*)noise_s_d;
  l.kk[0] = (const blk*)kk_s_d;
  l.out[0] = (blk*)out_s_d;
  l.noise[1] = (const blk*)noise_r_d;
  l.kk[1] = (const blk*)kk_r_d;
  l.out[1] = (blk*)out_r_d;
  CUOT_HIP_TRY(cuot_launch_lpn(l, ctx->st(stream)));
  return CUOT_OK;
}

extern "C" int64_t cuot_qc_ahat_len(int64_t n) {
  if (n < 1 || cuot_qc_log_len(n) > 30) return 0;
  return cuot_qc_ahat_words(n);
}

extern "C" size_t cuot_qc_scratch_bytes(int64_t n) {
  if (n < 1) return 0;
  return cuot_qc_scratch_bytes_for(n);
}

static int qc_check(int64_t n) {
  CUOT_CHECK_ARG(n >= 1 && n <= (1ll << 34), "n out of range");
  CUOT_CHECK_ARG(cuot_qc_log_len(n) <= 30, "n too large for the GF(2^32) transform");
  return CUOT_OK;
}

extern "C" int cuot_qc_prepare(cuot_ctx* ctx, int64_t n, const uint8_t* a_bits_d, uint32_t* ahat_d, void* scratch_d,
                               void* stream) {
  CUOT_CHECK_ARG(ctx, "ctx is NULL");
  int rc = qc_check(n);
  if (rc) return rc;
  CUOT_CHECK_ARG(a_bits_d && ahat_d && scratch_d, "buffer is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  CUOT_HIP_TRY(cuot_launch_qc_prepare(a_bits_d, n, ahat_d, scratch_d, ctx->st(stream)));
  return CUOT_OK;
}

extern "C" int cuot_qc_encode(cuot_ctx* ctx, int64_t n, const uint32_t* ahat_d, const cuot_block* x_d,
                              cuot_block* y_d, void* scratch_d, void* stream) {
  CUOT_CHECK_ARG(ctx, "ctx is NULL");
  int rc = qc_check(n);
  if (rc) return rc;
  CUOT_CHECK_ARG(ahat_d && x_d && y_d && scratch_d, "buffer is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  CUOT_HIP_TRY(cuot_launch_qc_encode(ahat_d, n, x_d, y_d, scratch_d, ctx->st(stream)));
  return CUOT_OK;
}

extern "C" int cuot_qc_sparse(cuot_ctx* ctx, int64_t n, const uint8_t* a_bits_d, const int64_t* pos_d, int64_t w,
                              uint8_t* out_bits_d, void* stream) {
  CUOT_CHECK_ARG(ctx, "ctx is NULL");
  CUOT_CHECK_ARG(n >= 1, "n < 1");
  CUOT_CHECK_ARG(w >= 0 && w < (1ll << 31), "w out of range");
  CUOT_CHECK_ARG(a_bits_d && out_bits_d && (pos_d || w == 0), "buffer is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  CUOT_HIP_TRY(cuot_launch_sparse((const uint32_t*)a_bits_d, n, pos_d, (int)w, (uint32_t*)out_bits_d,
                                  ctx->st(stream)));
  return CUOT_OK;
}

extern "C" int cuot_bit_transpose(cuot_ctx* ctx, const cuot_block* in_d, int64_t N, uint8_t* rows_d, void* stream) {
  CUOT_CHECK_ARG(ctx, "ctx is NULL");
  CUOT_CHECK_ARG(N >= 0 && N % 32 == 0, "N must be a non-negative multiple of 32");
  if (N == 0) return CUOT_OK;
  CUOT_CHECK_ARG(in_d && rows_d, "buffer is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  CUOT_HIP_TRY(cuot_launch_transpose(in_d, N, rows_d, false, ctx->st(stream)));
  return CUOT_OK;
}

extern "C" int cuot_bit_untranspose(cuot_ctx* ctx, const uint8_t* rows_d, int64_t N, cuot_block* out_d,
                                    void* stream) {
  CUOT_CHECK_ARG(ctx, "ctx is NULL");
  CUOT_CHECK_ARG(N >= 0 && N % 32 == 0, "N must be a non-negative multiple of 32");
  if (N == 0) return CUOT_OK;
  CUOT_CHECK_ARG(rows_d && out_d, "buffer is NULL");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  CUOT_HIP_TRY(cuot_launch_transpose(rows_d, N, out_d, true, ctx->st(stream)));
  return CUOT_OK;
}

// ---- execution knobs ---------------------------------------------------------
namespace {
CuotKnobs g_knobs = {0, 0, 0, 0, 0, 0};
}

const CuotKnobs& cuot_knobs() { return g_knobs; }

extern "C" int cuot_set_knobs(const cuot_knob_values* k) {
  CUOT_CHECK_ARG(k, "NULL knobs");
  CUOT_CHECK_ARG(k->qc_radix == 0 || (k->qc_radix >= 2 && k->qc_radix <= 8), "qc_radix must be 0 or 2..8");
  CUOT_CHECK_ARG(k->qc_cols == 0 || k->qc_cols == 1 || k->qc_cols == 2 || k->qc_cols == 4 || k->qc_cols == 8,
                 "qc_cols must be 0, 1, 2, 4 or 8");
  CUOT_CHECK_ARG(k->ggm_dfs >= 0 && k->ggm_dfs <= 6, "ggm_dfs must be in [0, 6]");
  CUOT_CHECK_ARG(k->lpn_wgs >= 0, "lpn_wgs must be >= 0");
  CUOT_CHECK_ARG(k->qc_mfma_max >= -1, "qc_mfma_max must be >= -1");
  CUOT_CHECK_ARG(k->aes_impl >= 0 && k->aes_impl <= 2, "aes_impl must be 0, 1 or 2");
  g_knobs.qc_radix = k->qc_radix;
  g_knobs.qc_cols = k->qc_cols;
  g_knobs.ggm_dfs = k->ggm_dfs;
  g_knobs.lpn_wgs = k->lpn_wgs;
  g_knobs.qc_mfma_max = k->qc_mfma_max;
  g_knobs.aes_impl = k->aes_impl;
  return CUOT_OK;
}

extern "C" int64_t cuot_qc_mfma_crossover(void) { return cuot_qc_mfma_limit(); }

extern "C" int cuot_get_knobs(cuot_knob_values* k) {
  CUOT_CHECK_ARG(k, "NULL knobs");
  k->qc_radix = g_knobs.qc_radix;
  k->qc_cols = g_knobs.qc_cols;
  k->ggm_dfs = g_knobs.ggm_dfs;
  k->lpn_wgs = g_knobs.lpn_wgs;
  k->qc_mfma_max = g_knobs.qc_mfma_max;
  k->aes_impl = g_knobs.aes_impl;
  return CUOT_OK;
}

// ---- per-launch timing ----------------------------------------------------------
namespace {
thread_local LaunchRecorder* t_rec = nullptr;
}

LaunchRecorder* cuot_recorder() { return t_rec; }

void cuot_rec_begin(hipStream_t st, const char* name, double bytes) {
  if (!t_rec) return;
  LaunchRecorder::Item it;
  it.name = name;
  it.bytes = bytes;
  it.a = it.b = nullptr;
  if (hipEventCreate(&it.a) != hipSuccess || hipEventCreate(&it.b) != hipSuccess) return;
  hipEventRecord(it.a, st);
  t_rec->items.push_back(it);
}

void cuot_rec_end(hipStream_t st) {
  if (!t_rec || t_rec->items.empty()) return;
  hipEventRecord(t_rec->items.back().b, st);
}

extern "C" int cuot_timing_begin(cuot_ctx* ctx) {
  CUOT_CHECK_ARG(ctx, "NULL ctx");
  if (t_rec) {
    cuot_set_last_error("timing already active on this thread", __FILE__, __LINE__);
    return CUOT_ERR_STATE;
  }
  t_rec = new LaunchRecorder();
  return CUOT_OK;
}

extern "C" int cuot_timing_end(cuot_ctx* ctx, cuot_launch_time* out, int max_out, int* count) {
  CUOT_CHECK_ARG(ctx && count && (out || max_out == 0), "NULL argument");
  if (!t_rec) {
    cuot_set_last_error("timing not active on this thread", __FILE__, __LINE__);
    return CUOT_ERR_STATE;
  }
  LaunchRecorder* rec = t_rec;
  t_rec = nullptr;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  int c = 0;
  for (auto& it : rec->items) {
    float ms = 0;
    if (e == hipSuccess && c < max_out && hipEventElapsedTime(&ms, it.a, it.b) == hipSuccess) {
      memset(out[c].kernel, 0, sizeof(out[c].kernel));
      strncpy(out[c].kernel, it.name, sizeof(out[c].kernel) - 1);
      out[c].ms = ms;
      out[c].bytes = it.bytes;
      c++;
    }
    hipEventDestroy(it.a);
    hipEventDestroy(it.b);
  }
  delete rec;
  *count = c;
  CUOT_HIP_TRY(e);
  return CUOT_OK;
}
