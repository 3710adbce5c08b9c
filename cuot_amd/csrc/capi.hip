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

extern "C" const char* cuot_last_error(void) { return g_last_error.c_str(); }

extern "C" int cuot_init(int device, cuot_ctx** out) {
  CUOT_CHECK_ARG(out != nullptr, "out is NULL");
  int n = 0;
  CUOT_HIP_TRY(hipGetDeviceCount(&n));
  CUOT_CHECK_ARG(device >= 0 && device < n, "device index out of range");
  CUOT_HIP_TRY(hipSetDevice(device));
  CUOT_HIP_TRY(cuot_upload_tables(device));
  cuot_ctx* c = new cuot_ctx();
  c->device = device;
  c->stream = nullptr;  // legacy default stream unless the caller passes one
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
