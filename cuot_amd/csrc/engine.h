// engine.h -- instance plans: workspace carving and stage launch sequences
// for one SOT or Ferret instance (host side).
#pragma once

#include "kernels.h"

struct SotPlan {
  cuot_sot_config cfg;
  blk *delta, *roots, *alpha;
  uint8_t* choice;
  int64_t* pos;
  uint32_t* a_words;
  int64_t a_words_total;
  blk *m0b, *m1b, *sums, *M0, *M1, *csum, *fin;
  uint32_t* ahat;
  blk *leaves_s, *leaves_r;
  char* qc_scratch;
  char* qc_scratch2;       // second encode's scratch (both parties' encodes overlap)
  bool shared_qc = false;  // batch: the QC scratch lives outside the instance

  size_t carve(void* ws);  // returns bytes needed; fills pointers if ws != NULL
  int validate() const;
  hipError_t run_setup(hipStream_t st);
  hipError_t run_prepare(hipStream_t st);
  void ggm_args(GgmKArgs& a) const;
  hipError_t run_expand_sender(hipStream_t st);
  hipError_t run_sender_msgs(hipStream_t st);  // masked rows + receiver's view
  hipError_t run_expand_recver(hipStream_t st);
  hipError_t run_expand(hipStream_t st);
  hipError_t run_lpn(blk* v, blk* w, uint32_t* b, hipStream_t st);
};

struct FerretPlan {
  cuot_ferret_config cfg;
  blk *delta, *roots, *alpha;
  uint8_t* choice;
  blk *sums, *csum, *ssum, *lpn_seed, *K, *Mr, *cbits;
  blk* kkp;  // K and Mr interleaved for the LPN gathers

  size_t carve(void* ws);
  int validate() const;
  hipError_t run_setup(hipStream_t st);
  void ggm_args(GgmKArgs& a) const;
  hipError_t run_mpcot_sender(blk* out_s, hipStream_t st);
  hipError_t run_sender_sums(hipStream_t st);  // secret sum + receiver's chosen sums
  hipError_t run_mpcot_recver(blk* out_r, hipStream_t st);
  hipError_t run_mpcot(blk* out_s, blk* out_r, hipStream_t st);
  void lpn_launch(blk* out_s, blk* out_r, LpnLaunch& l) const;  // gathers from kkp (fill it first)
  hipError_t fill_kkp(hipStream_t st);
  hipError_t run_lpn(blk* out_s, blk* out_r, hipStream_t st);
};

// Batch binding with an explicit stream for the table upload (NULL: synchronous)
// and per-instance Delta destinations; the public cuot_batch_bind wraps it.
int cuot_batch_bind_on(cuot_batch* b, void* ws_d, size_t ws_bytes, cuot_block* const* out0, cuot_block* const* out1,
                       uint8_t* const* bits, cuot_block* const* delta, hipStream_t st);
