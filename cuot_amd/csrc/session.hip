// session.hip -- an iterating, device-resident Ferret session (emp-ot's
// FerretCOT extend loop, ferret_cot.hpp:79-103, 171-209), both roles on one GPU.
//
// The session keeps Delta and a bootstrap of M = k + h*t + 128 COTs per party
// in HBM.  One extend:
//   1. pre-OT (preot.h:38-54): the first h*t bootstrap COTs become random OTs
//      through the circular correlation-robust hash H: sender pads
//      P0 = H(K), P1 = H(K ^ Delta); receiver bits = LSB(M), Pr = H(M);
//   2. MPCOT: sender GGM trees (roots from the seed, stream 1, advanced by t
//      per extend), level sums sent as m0 ^ P0 / m1 ^ P1; the receiver's
//      choice per level is its bit, its chosen sum pad_b ^ Pr (its own view);
//      the secret sum closes the punctured leaf (spcot_recver.h:53-64);
//   3. LPN with kk = bootstrap + 128 (ferret_cot.hpp:89);
//   4. the last M outputs become the next bootstrap (ferret_cot.hpp:100).
// H(x) = AES0(s(x)) ^ s(x) with s(x) = (hi ^ lo, hi) as (hi, lo) halves and
// AES0 = AES-128 under the all-zero key.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "aes_dev.h"
#include "ctx.h"
#include "engine.h"
#include "kernels.h"

namespace {

struct CcrhArgs {
  const uint32_t* T0;
  uint32_t rk[44];
  const blk* boot;
  int64_t npre;
  const blk* delta;
  int recv;
  blk* out0;     // P0 (sender) or Pr (receiver)
  blk* out1;     // P1 (sender)
  uint8_t* bits; // receiver choice bits
};

CUOT_DEV blk ccrh(const AesLds& aes, const uint32_t* rk, blk x) {
  // s(x): hi' = hi ^ lo, lo' = hi (hi = words 2,3)
  blk s;
  s.x = x.z;
  s.y = x.w;
  s.z = x.z ^ x.x;
  s.w = x.w ^ x.y;
  return bxor(aes_enc(aes, rk, s), s);
}

__global__ __launch_bounds__(256) void ccrh_pre_kernel(CcrhArgs a) {
  __shared__ uint32_t tab[CUOT_AES_LDS_WORDS];
  aes_stage_table(tab, a.T0, threadIdx.x, 256);
  __syncthreads();
  const AesLds aes = aes_lds(tab);
  const blk D = *a.delta;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.npre; i += (int64_t)gridDim.x * blockDim.x) {
    const blk v = a.boot[i];
    if (!a.recv) {
      a.out0[i] = ccrh(aes, a.rk, v);
      a.out1[i] = ccrh(aes, a.rk, bxor(v, D));
    } else {
      a.out0[i] = ccrh(aes, a.rk, v);
      a.bits[i] = (uint8_t)(v.x & 1u);
    }
  }
}

// receiver's chosen level sums from the sender's pads (sums [tree][level][L,R])
__global__ void ot_unmask_kernel(const blk* __restrict__ sums, const blk* __restrict__ P0, const blk* __restrict__ P1,
                                 const blk* __restrict__ Pr, const uint8_t* __restrict__ bits, int64_t npre,
                                 blk* __restrict__ csum) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npre; i += (int64_t)gridDim.x * blockDim.x) {
    const blk pad0 = bxor(sums[2 * i], P0[i]);
    const blk pad1 = bxor(sums[2 * i + 1], P1[i]);
    csum[i] = bxor(bits[i] ? pad1 : pad0, Pr[i]);
  }
}

unsigned grid_of(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (unsigned)g;
}

}  // namespace

struct cuot_ferret_session {
  cuot_ctx* ctx = nullptr;
  FerretPlan plan;  // parameters, GGM/LPN argument builders
  int64_t npre = 0, M = 0, ot_limit = 0, round = 0;
  int64_t n_pre_file = 0;  // blocks of ot_pre_data in the state blob / pre-OT files (>= M)
  bool seeded = false;  // the MPCOT half of the current extend has run
  bool malicious = false;  // run the consistency check after each MPCOT
  bool checked = false;    // a consistency check has been enqueued for the current extend
  blk* vwc = nullptr;      // per-tree V, W, chi_alpha (3 t blocks)
  uint8_t* chk = nullptr;  // x' (128), digests (2 x 32), ok (int32)
  char* mem = nullptr;
  blk *delta, *boot_s, *boot_r, *P0, *P1, *Pr, *roots, *sums, *csum, *ssum, *cbits, *kkp;
  uint8_t* bits;
};

namespace {

size_t carve_session(cuot_ferret_session* s, char* base) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    off = (off + 255) & ~(size_t)255;
    char* p = base ? base + off : nullptr;
    off += bytes;
    return p;
  };
  const int64_t t = s->plan.cfg.t, h = s->plan.cfg.h;
  s->delta = (blk*)take(16);
  s->boot_s = (blk*)take((size_t)s->M * 16);
  s->boot_r = (blk*)take((size_t)s->M * 16);
  s->P0 = (blk*)take((size_t)s->npre * 16);
  s->P1 = (blk*)take((size_t)s->npre * 16);
  s->Pr = (blk*)take((size_t)s->npre * 16);
  s->bits = (uint8_t*)take((size_t)s->npre);
  s->roots = (blk*)take((size_t)t * 16);
  s->sums = (blk*)take((size_t)2 * t * h * 16);
  s->csum = (blk*)take((size_t)t * h * 16);
  s->ssum = (blk*)take((size_t)t * 16);
  s->cbits = (blk*)take((size_t)((s->M + 127) / 128) * 16);
  s->kkp = (blk*)take((size_t)2 * s->plan.cfg.k * 16);
  s->vwc = (blk*)take((size_t)3 * t * 16);
  s->chk = (uint8_t*)take(256);
  return off + 256;
}

// the f2k consistency check on the current trees' leaves and the first 128
// bootstrap COTs (mpcot_reg.h:97-98: pre_cot_data = ot_input)
hipError_t consistency(cuot_ferret_session* s, const blk* out_s, const blk* out_r, hipStream_t st) {
  const cuot_ferret_config& cfg = s->plan.cfg;
  s->checked = true;
  return cuot_launch_consist(out_s, out_r, cfg.n, cfg.t, cfg.h, s->ssum, s->bits, s->boot_s, s->boot_r, s->delta,
                             s->vwc, s->chk, st);
}

// the first half of an extend (emp-ot's SeedExp event): pre-OT and MPCOT
hipError_t extend_seed(cuot_ferret_session* s, blk* out_s, blk* out_r, hipStream_t st) {
  const cuot_ferret_config& cfg = s->plan.cfg;
  hipError_t e;
  s->checked = false;  // a check result belongs to the extend whose MPCOT it checked
  // 1. pre-OT through the hash
  CcrhArgs c;
  memset(&c, 0, sizeof(c));
  c.T0 = cuot_device_T0();
  const cuot_block zero = {0, 0};
  cuot_aes_expand_host((const uint8_t*)&zero, c.rk);
  c.npre = s->npre;
  c.delta = s->delta;
  c.boot = s->boot_s;
  c.out0 = s->P0;
  c.out1 = s->P1;
  hipLaunchKernelGGL(ccrh_pre_kernel, dim3(grid_of(s->npre)), dim3(256), 0, st, c);
  c.recv = 1;
  c.boot = s->boot_r;
  c.out0 = s->Pr;
  c.out1 = nullptr;
  c.bits = s->bits;
  hipLaunchKernelGGL(ccrh_pre_kernel, dim3(grid_of(s->npre)), dim3(256), 0, st, c);
  if ((e = hipGetLastError())) return e;
  // 2. sender MPCOT
  uint32_t rk_seed[44];
  cuot_aes_expand_host((const uint8_t*)&cfg.seed, rk_seed);
  if ((e = cuot_launch_prg(rk_seed, 1, (uint64_t)(s->round * cfg.t), cfg.t, s->roots, st))) return e;
  GgmKArgs a;
  s->plan.ggm_args(a);
  a.roots = s->roots;
  a.leaves = out_s;
  a.sums = s->sums;
  if ((e = cuot_launch_ggm(a, false, st))) return e;
  if ((e = cuot_launch_ferret_sums(s->sums, s->delta, nullptr, cfg.t, cfg.h, s->ssum, nullptr, nullptr, nullptr, st)))
    return e;
  hipLaunchKernelGGL(ot_unmask_kernel, dim3(grid_of(s->npre)), dim3(256), 0, st, s->sums, s->P0, s->P1, s->Pr, s->bits,
                     s->npre, s->csum);
  if ((e = hipGetLastError())) return e;
  //    receiver MPCOT from its bits and unmasked sums
  GgmKArgs r;
  s->plan.ggm_args(r);
  r.leaves = out_r;
  r.choice = s->bits;
  r.csum = s->csum;
  r.fin = s->ssum;
  if ((e = cuot_launch_ggm(r, true, st))) return e;
  if (s->malicious) return consistency(s, out_s, out_r, st);
  return hipSuccess;
}

// the second half (emp-ot's LPN event): LPN and the bootstrap refresh
hipError_t extend_lpn(cuot_ferret_session* s, blk* out_s, blk* out_r, hipStream_t st) {
  const cuot_ferret_config& cfg = s->plan.cfg;
  hipError_t e;
  // 3. LPN with kk = bootstrap + 128
  if ((e = cuot_launch_interleave(s->boot_s + 128, s->boot_r + 128, cfg.k, s->kkp, st))) return e;
  LpnLaunch l;
  s->plan.lpn_launch(out_s, out_r, l);
  l.kk[0] = s->boot_s + 128;
  l.kk[1] = s->boot_r + 128;
  l.kkp = s->kkp;
  if ((e = cuot_launch_lpn(l, st))) return e;
  // 4. refresh the bootstrap from the last M outputs
  if ((e = hipMemcpyAsync(s->boot_s, out_s + s->ot_limit, (size_t)s->M * 16, hipMemcpyDeviceToDevice, st))) return e;
  if ((e = hipMemcpyAsync(s->boot_r, out_r + s->ot_limit, (size_t)s->M * 16, hipMemcpyDeviceToDevice, st))) return e;
  s->round++;
  return hipSuccess;
}

}  // namespace

namespace {

// validated parameters and one device allocation carved into the state buffers
int session_alloc(cuot_ctx* ctx, const cuot_ferret_config* cfg, cuot_ferret_session** out) {
  *out = nullptr;
  cuot_ferret_session* s = new cuot_ferret_session();
  s->ctx = ctx;
  s->plan.cfg = *cfg;
  if (s->plan.cfg.ntasks < 1) s->plan.cfg.ntasks = 1;
  int rc = s->plan.validate();
  if (rc) {
    delete s;
    return rc;
  }
  s->npre = (int64_t)cfg->h * cfg->t;
  s->M = cfg->k + s->npre + 128;
  s->ot_limit = cfg->n - s->M;
  s->n_pre_file = cfg->n_pre ? cfg->n_pre : s->M;
  if (s->n_pre_file < s->M) {
    delete s;
    cuot_set_last_error("n_pre must be 0 or at least k + h*t + 128 (LPN parameter not matched)", __FILE__,
                        __LINE__);
    return CUOT_ERR_ARG;
  }
  if (s->ot_limit <= 0) {
    delete s;
    cuot_set_last_error("n must exceed k + h*t + 128 (the bootstrap carried to the next extend)", __FILE__,
                        __LINE__);
    return CUOT_ERR_ARG;
  }
  const size_t bytes = carve_session(s, nullptr);
  if (hipSetDevice(ctx->device) != hipSuccess || hipMalloc((void**)&s->mem, bytes) != hipSuccess) {
    delete s;
    cuot_set_last_error("cannot allocate the session state", __FILE__, __LINE__);
    return CUOT_ERR_NOMEM;
  }
  carve_session(s, s->mem);
  *out = s;
  return CUOT_OK;
}

void session_destroy(cuot_ferret_session* s) {
  if (s->mem) hipFree(s->mem);
  delete s;
}

// ---- state blob and pre-OT file (emp-ot layouts) ------------------------------
// party ids as in emp-ot: ALICE = 1 (sender, holds Delta), BOB = 2 (receiver)
constexpr int64_t kAlice = 1, kBob = 2;

int64_t state_bytes(int64_t P) { return 4 * (int64_t)sizeof(int64_t) + 16 + 16 * P; }

// blocks of ot_pre_data in the blob and files: cfg->n_pre, or M when 0;
// -1 when the parameters are invalid or n_pre < M
int64_t pre_blocks(const cuot_ferret_config* cfg) {
  if (!cfg || cfg->n <= 0 || cfg->k <= 0 || cfg->t <= 0 || cfg->h <= 0 || cfg->n_pre < 0) return -1;
  const int64_t M = cfg->k + (int64_t)cfg->h * cfg->t + 128;
  if (M >= cfg->n) return -1;
  const int64_t P = cfg->n_pre ? cfg->n_pre : M;
  return P >= M ? P : -1;
}

// copies one party's Delta (zero for the receiver) and its n_pre_file blocks
// of ot_pre_data (the M bootstrap COTs, then zeros) to the host after all
// work on the device has finished
int export_party(const cuot_ferret_session* s, int64_t party, blk* delta_h, void* boot_h) {
  CUOT_HIP_TRY(hipSetDevice(s->ctx->device));
  CUOT_HIP_TRY(hipDeviceSynchronize());
  if (party == kAlice) {
    CUOT_HIP_TRY(hipMemcpy(delta_h, s->delta, 16, hipMemcpyDeviceToHost));
  } else {
    memset(delta_h, 0, 16);
  }
  CUOT_HIP_TRY(hipMemcpy(boot_h, party == kAlice ? s->boot_s : s->boot_r, (size_t)s->M * 16, hipMemcpyDeviceToHost));
  memset((char*)boot_h + (size_t)s->M * 16, 0, (size_t)(s->n_pre_file - s->M) * 16);
  return CUOT_OK;
}

// a session at `round` from both parties' Delta and bootstraps (host)
int import_session(cuot_ctx* ctx, const cuot_ferret_config* cfg, int64_t round, const void* delta_h,
                   const void* boot_s_h, const void* boot_r_h, cuot_ferret_session** out) {
  CUOT_CHECK_ARG(round >= 0, "round must be >= 0");
  cuot_ferret_session* s = nullptr;
  int rc = session_alloc(ctx, cfg, &s);
  if (rc) return rc;
  s->round = round;
  hipError_t e = hipMemcpy(s->delta, delta_h, 16, hipMemcpyHostToDevice);
  if (!e) e = hipMemcpy(s->boot_s, boot_s_h, (size_t)s->M * 16, hipMemcpyHostToDevice);
  if (!e) e = hipMemcpy(s->boot_r, boot_r_h, (size_t)s->M * 16, hipMemcpyHostToDevice);
  if (e) {
    session_destroy(s);
    CUOT_HIP_TRY(e);
  }
  *out = s;
  return CUOT_OK;
}

bool read_exact(FILE* f, void* dst, size_t bytes) { return fread(dst, 1, bytes, f) == bytes; }

}  // namespace

extern "C" int cuot_ferret_session_new(cuot_ctx* ctx, const cuot_ferret_config* cfg, cuot_ferret_session** out) {
  CUOT_CHECK_ARG(ctx && cfg && out, "NULL argument");
  cuot_ferret_session* s = nullptr;
  int rc = session_alloc(ctx, cfg, &s);
  if (rc) return rc;
  // Delta and the initial bootstrap from the seed (streams 0, 7, 8)
  uint32_t rk[44];
  cuot_aes_expand_host((const uint8_t*)&cfg->seed, rk);
  hipStream_t st = ctx->stream;
  hipError_t e = cuot_launch_prg(rk, 0, 0, 1, s->delta, st);
  if (!e) e = cuot_launch_setup_fix(s->delta, nullptr, 0, cfg->h, nullptr, nullptr, nullptr, 0, 0, st);
  if (!e) e = cuot_launch_prg(rk, 7, 0, s->M, s->boot_s, st);
  if (!e) e = cuot_launch_prg(rk, 8, 0, (s->M + 127) / 128, s->cbits, st);
  if (!e) e = cuot_launch_ferret_boot(s->boot_s, s->boot_r, (const uint8_t*)s->cbits, s->delta, s->M, st);
  if (!e) e = hipStreamSynchronize(st);
  if (e) {
    session_destroy(s);
    CUOT_HIP_TRY(e);
  }
  *out = s;
  return CUOT_OK;
}

extern "C" int64_t cuot_ferret_state_size(const cuot_ferret_config* cfg) {
  const int64_t P = pre_blocks(cfg);
  return P > 0 ? state_bytes(P) : -1;
}

extern "C" int cuot_ferret_session_assemble_state(const cuot_ferret_session* s, int party, void* data, int64_t size) {
  CUOT_CHECK_ARG(s && data, "NULL argument");
  CUOT_CHECK_ARG(party == kAlice || party == kBob, "party must be 1 (ALICE, sender) or 2 (BOB, receiver)");
  CUOT_CHECK_ARG(size >= state_bytes(s->n_pre_file), "state buffer smaller than cuot_ferret_state_size");
  unsigned char* a = (unsigned char*)data;
  const int64_t head[4] = {(int64_t)party, s->plan.cfg.n, (int64_t)s->plan.cfg.t, s->plan.cfg.k};
  memcpy(a, head, sizeof(head));
  return export_party(s, party, (blk*)(a + sizeof(head)), a + sizeof(head) + 16);
}

extern "C" int cuot_ferret_session_disassemble_state(cuot_ctx* ctx, const cuot_ferret_config* cfg, int64_t round,
                                                     const void* alice, const void* bob, int64_t size,
                                                     cuot_ferret_session** out) {
  CUOT_CHECK_ARG(ctx && cfg && alice && bob && out, "NULL argument");
  *out = nullptr;
  const int64_t need = cuot_ferret_state_size(cfg);
  CUOT_CHECK_ARG(need > 0 && size >= need, "state buffers smaller than cuot_ferret_state_size");
  const unsigned char* sa = (const unsigned char*)alice;
  const unsigned char* sb = (const unsigned char*)bob;
  int64_t ha[4], hb[4];
  memcpy(ha, sa, sizeof(ha));
  memcpy(hb, sb, sizeof(hb));
  const bool match = ha[0] == kAlice && hb[0] == kBob && ha[1] == cfg->n && hb[1] == cfg->n && ha[2] == cfg->t &&
                     hb[2] == cfg->t && ha[3] == cfg->k && hb[3] == cfg->k;
  CUOT_CHECK_ARG(match, "state does not match the parties or the parameters (party, n, t, k)");
  return import_session(ctx, cfg, round, sa + 32, sa + 48, sb + 48, out);
}

extern "C" int cuot_ferret_session_write_pre_file(const cuot_ferret_session* s, int party, const char* path) {
  CUOT_CHECK_ARG(s && path, "NULL argument");
  CUOT_CHECK_ARG(party == kAlice || party == kBob, "party must be 1 (ALICE, sender) or 2 (BOB, receiver)");
  std::vector<unsigned char> boot((size_t)s->n_pre_file * 16);
  blk delta;
  int rc = export_party(s, party, &delta, boot.data());
  if (rc) return rc;
  FILE* f = fopen(path, "wb");
  CUOT_CHECK_ARG(f, "create a directory to store pre-OT data");
  const int64_t p64 = party;
  const int64_t nkt[3] = {s->plan.cfg.n, (int64_t)s->plan.cfg.t, s->plan.cfg.k};
  bool ok = fwrite(&p64, sizeof(p64), 1, f) == 1;
  if (party == kAlice) ok = ok && fwrite(&delta, 16, 1, f) == 1;
  ok = ok && fwrite(nkt, sizeof(nkt), 1, f) == 1;
  ok = ok && fwrite(boot.data(), 1, boot.size(), f) == boot.size();
  ok = (fclose(f) == 0) && ok;
  CUOT_CHECK_ARG(ok, "cannot write the pre-OT file");
  return CUOT_OK;
}

extern "C" int cuot_ferret_session_read_pre_files(cuot_ctx* ctx, const cuot_ferret_config* cfg, int64_t round,
                                                  const char* alice_path, const char* bob_path,
                                                  cuot_ferret_session** out) {
  CUOT_CHECK_ARG(ctx && cfg && alice_path && bob_path && out, "NULL argument");
  *out = nullptr;
  const int64_t need = cuot_ferret_state_size(cfg);
  CUOT_CHECK_ARG(need > 0, "invalid parameters");
  const int64_t P = (need - 48) / 16;  // ot_pre_data blocks in the file; the first M are the bootstrap
  std::vector<unsigned char> boot[2];
  blk delta;
  const char* paths[2] = {alice_path, bob_path};
  for (int r = 0; r < 2; r++) {
    FILE* f = fopen(paths[r], "rb");
    CUOT_CHECK_ARG(f, "cannot open the pre-OT file");
    int64_t party = 0, nkt[3] = {0, 0, 0};
    bool ok = read_exact(f, &party, sizeof(party));
    const bool party_ok = ok && party == (r ? kBob : kAlice);
    if (party_ok && r == 0) ok = read_exact(f, &delta, 16);
    ok = ok && party_ok && read_exact(f, nkt, sizeof(nkt));
    const bool params_ok = ok && nkt[0] == cfg->n && nkt[1] == cfg->t && nkt[2] == cfg->k;
    boot[r].resize((size_t)P * 16);
    ok = ok && params_ok && read_exact(f, boot[r].data(), boot[r].size());
    fclose(f);
    CUOT_CHECK_ARG(party_ok, "wrong party");
    CUOT_CHECK_ARG(params_ok, "wrong parameters");
    CUOT_CHECK_ARG(ok, "pre-OT file too short");
  }
  int rc = import_session(ctx, cfg, round, &delta, boot[0].data(), boot[1].data(), out);
  if (rc) return rc;
  // the files are consumed, as emp-ot's reader removes them
  remove(alice_path);
  remove(bob_path);
  return CUOT_OK;
}

extern "C" int cuot_ferret_session_info(const cuot_ferret_session* s, int64_t* M, int64_t* ot_limit, int64_t* round) {
  CUOT_CHECK_ARG(s, "NULL session");
  if (M) *M = s->M;
  if (ot_limit) *ot_limit = s->ot_limit;
  if (round) *round = s->round;
  return CUOT_OK;
}

extern "C" int cuot_ferret_session_delta(const cuot_ferret_session* s, cuot_block* delta_d, void* stream) {
  CUOT_CHECK_ARG(s && delta_d, "NULL argument");
  CUOT_HIP_TRY(hipSetDevice(s->ctx->device));
  CUOT_HIP_TRY(hipMemcpyAsync(delta_d, s->delta, 16, hipMemcpyDeviceToDevice, s->ctx->st(stream)));
  return CUOT_OK;
}

extern "C" int cuot_ferret_session_extend_stages(cuot_ferret_session* s, cuot_block* out_s_d, cuot_block* out_r_d,
                                                 int stages, void* stream) {
  CUOT_CHECK_ARG(s && out_s_d && out_r_d, "NULL argument");
  CUOT_CHECK_ARG(stages >= 1 && stages <= 4, "stages must be 1 (MPCOT), 2 (LPN), 3 (both) or 4 (check)");
  if (stages == 4) {  // the consistency check alone, between the two halves
    if (!s->seeded) {
      cuot_set_last_error("the consistency check needs the MPCOT stage of the same extend first", __FILE__,
                          __LINE__);
      return CUOT_ERR_STATE;
    }
    CUOT_HIP_TRY(hipSetDevice(s->ctx->device));
    CUOT_HIP_TRY(consistency(s, (const blk*)out_s_d, (const blk*)out_r_d, s->ctx->st(stream)));
    return CUOT_OK;
  }
  if ((stages & 1) && s->seeded) {
    cuot_set_last_error("the MPCOT stage already ran; run the LPN stage next", __FILE__, __LINE__);
    return CUOT_ERR_STATE;
  }
  if (stages == 2 && !s->seeded) {
    cuot_set_last_error("the LPN stage needs the MPCOT stage of the same extend first", __FILE__, __LINE__);
    return CUOT_ERR_STATE;
  }
  CUOT_HIP_TRY(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->st(stream);
  if (stages & 1) {
    CUOT_HIP_TRY(extend_seed(s, (blk*)out_s_d, (blk*)out_r_d, st));
    s->seeded = true;
  }
  if (stages & 2) {
    CUOT_HIP_TRY(extend_lpn(s, (blk*)out_s_d, (blk*)out_r_d, st));
    s->seeded = false;
  }
  return CUOT_OK;
}

extern "C" int cuot_ferret_session_set_malicious(cuot_ferret_session* s, int on) {
  CUOT_CHECK_ARG(s, "NULL session");
  s->malicious = on != 0;
  return CUOT_OK;
}

extern "C" int cuot_ferret_session_check(const cuot_ferret_session* s, int* ok, uint8_t* xprime, uint8_t* dig_a,
                                         uint8_t* dig_b) {
  CUOT_CHECK_ARG(s && ok, "NULL argument");
  if (!s->checked) {
    cuot_set_last_error("no consistency check has run for the current extend", __FILE__, __LINE__);
    return CUOT_ERR_STATE;
  }
  uint8_t h[196];
  CUOT_HIP_TRY(hipSetDevice(s->ctx->device));
  CUOT_HIP_TRY(hipDeviceSynchronize());
  CUOT_HIP_TRY(hipMemcpy(h, s->chk, sizeof(h), hipMemcpyDeviceToHost));
  int32_t v;
  memcpy(&v, h + 192, 4);
  *ok = v;
  if (xprime) memcpy(xprime, h, 128);
  if (dig_a) memcpy(dig_a, h + 128, 32);
  if (dig_b) memcpy(dig_b, h + 160, 32);
  return CUOT_OK;
}

extern "C" int cuot_ferret_session_extend(cuot_ferret_session* s, cuot_block* out_s_d, cuot_block* out_r_d,
                                          void* stream) {
  return cuot_ferret_session_extend_stages(s, out_s_d, out_r_d, 3, stream);
}

extern "C" int cuot_ferret_session_free(cuot_ferret_session* s) {
  if (!s) return CUOT_OK;
  if (s->mem) {
    hipSetDevice(s->ctx->device);
    hipDeviceSynchronize();
    hipFree(s->mem);
  }
  delete s;
  return CUOT_OK;
}
