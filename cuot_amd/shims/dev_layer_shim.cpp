// dev_layer_shim.cpp -- reference-side replacement for the body of emp-ot's
// GPU layer (ferret/emp-ot/emp-ot/ferret/dev_layer.cu), forwarding the four
// calls declared in ferret/emp-ot/emp-ot/ferret/dev_layer.h:9-18 to
// libcuot_amd.so.  It is compiled by the reference's host compiler against the
// reference's own headers (dev_layer.h, gpu/gpu_matrix.h: Mat::data() is the
// device pointer of a Mat) and linked with -lcuot_amd; this library does not
// build it.  MpcotReg (mpcot_reg.h:130-157) and LpnF2::compute (lpn_f2.h:81-87)
// call these functions unchanged.
#include <stdexcept>

#include "cuot_amd.h"
#include "dev_layer.h"

namespace {

cuot_ctx* g_ctx = nullptr;

void check(int rc) {
  if (rc != CUOT_OK) throw std::runtime_error(cuot_last_error());
}

cuot_block* dev(const Mat& m) { return reinterpret_cast<cuot_block*>(m.data()); }

}  // namespace

// dev_layer.h:9; one context per process (the library drives one GPU per
// process, ngpu == 1)
void cuda_setdev(int gpu) {
  check(cuot_setdev(gpu));
  if (!g_ctx) check(cuot_init(gpu, &g_ctx));
}

// dev_layer.h:11-12.  expanded[0] holds the t[0] tree roots in its first
// blocks on entry (MpcotReg fills them; the reference's level loop reads its
// level-0 input from there, dev_layer.cu:35-38) and the leaves on return.
void cuda_mpcot_sender(Mat* expanded, Mat* buffer, Mat* sep, blk* lSum_h, blk* rSum_h, blk* secret_sum, int* t,
                       int depth, blk* delta, int ngpu) {
  cuot_block* e[1] = {dev(expanded[0])};
  cuot_block* b[1] = {dev(buffer[0])};
  cuot_block* s[1] = {dev(sep[0])};
  check(cuot_mpcot_sender(g_ctx, e, b, s, reinterpret_cast<cuot_block*>(lSum_h), reinterpret_cast<cuot_block*>(rSum_h),
                          reinterpret_cast<cuot_block*>(secret_sum), t, depth,
                          reinterpret_cast<const cuot_block*>(delta), ngpu));
}

// dev_layer.h:14-15
void cuda_mpcot_recver(Mat* expanded, Mat* buffer, Mat* sep, blk* cSum_h, blk* secret_sum, int* t, int depth,
                       bool* choices, int ngpu) {
  cuot_block* e[1] = {dev(expanded[0])};
  cuot_block* b[1] = {dev(buffer[0])};
  cuot_block* s[1] = {dev(sep[0])};
  static_assert(sizeof(bool) == 1, "choices are passed as one byte per level");
  check(cuot_mpcot_recver(g_ctx, e, b, s, reinterpret_cast<const cuot_block*>(cSum_h),
                          reinterpret_cast<const cuot_block*>(secret_sum), t, depth,
                          reinterpret_cast<const uint8_t*>(choices), ngpu));
}

// dev_layer.h:17-18.  The index matrix is never materialised, so pubMats is
// not read.
void cuda_primal_lpn(Role role, Mat* pubMats, int64_t d, int64_t* n, int k, uint32_t* key, Mat* expanded, blk* nn,
                     Mat* kk_d, blk* kk, int ngpu) {
  (void)pubMats;
  cuot_block* e[1] = {dev(expanded[0])};
  cuot_block* kd[1] = {dev(kk_d[0])};
  check(cuot_primal_lpn(g_ctx, static_cast<int>(role), nullptr, d, n, k, key, e, reinterpret_cast<cuot_block*>(nn),
                        kd, reinterpret_cast<const cuot_block*>(kk), ngpu));
}
