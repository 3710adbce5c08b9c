// kernels.h -- internal kernel argument structs and launchers (not ABI).
#pragma once

#include <vector>

#include "common.h"

// GGM expansion of t trees of depth h (one 256-lane workgroup per tree).
struct GgmKArgs {
  const uint32_t* T0;    // AES T-table in global memory
  uint32_t rk[88];       // key-0 then key-1 AES-128 schedules (little-endian words)
  int64_t t;
  int h;
  int mode;              // CUOT_MODE_SOT / CUOT_MODE_FERRET
  int64_t n_out;         // leaves with global index >= n_out are not written
  const blk* roots;      // sender: [t]
  blk* leaves;           // [min(n_out, t<<h)]
  blk* sums;             // sender out: [t][h][2] (xor of left/right children per depth)
  const uint8_t* choice; // receiver: [t][h], 1 = chose the right (odd) side
  const blk* csum;       // receiver: [t][h] chosen level sums
  const blk* fin;        // receiver: [t] SOT final-row value / Ferret secret sum
};
hipError_t cuot_launch_ggm(const GgmKArgs& a, bool receiver, hipStream_t st);
// depth handled per lane below the shared top levels (trees of depth h)
int cuot_ggm_split(int h);
// batch of instances with the same cuot_ggm_split(h) == L and role; tab_d and
// start_d (ninst + 1 prefix of tree counts) in device memory
hipError_t cuot_launch_ggm_batch(const GgmKArgs* tab_d, const int64_t* start_d, int ninst, int64_t total_trees,
                                 int L, bool receiver, hipStream_t st);

// AES-128 ECB / CTR streams.  T0: device AES table (cuot_device_T0()).
const uint32_t* cuot_device_T0();
hipError_t cuot_launch_aes_ecb(const uint32_t rk[44], const blk* in, blk* out, int64_t n, hipStream_t st);
// out[i] = AES_key({lo = start + i, hi = stream})
hipError_t cuot_launch_prg(const uint32_t rk[44], uint64_t stream, uint64_t start, int64_t n, blk* out,
                           hipStream_t st);
// the same two streams with the bitsliced cipher (aes_bs.hip)
hipError_t cuot_launch_aes_ecb_bs(const uint32_t rk[44], const blk* in, blk* out, int64_t n, hipStream_t st);
hipError_t cuot_launch_prg_bs(const uint32_t rk[44], uint64_t stream, uint64_t start, int64_t n, blk* out,
                              hipStream_t st);
// GGM sender (full trees, level sums) breadth-first on the bitsliced cipher;
// rk: k0 then k1 round keys; work: cuot_ggm_bs_work_bytes bytes
size_t cuot_ggm_bs_work_bytes(int64_t t, int h);
hipError_t cuot_launch_ggm_sender_bs(const uint32_t rk[88], int64_t t, int h, int ferret, int64_t n_out,
                                     const blk* roots, blk* leaves, blk* sums, void* work, hipStream_t st);
// receiver from its own view (choice bits, chosen sums, final blocks)
size_t cuot_ggm_recver_bs_work_bytes(int64_t t, int h);
hipError_t cuot_launch_ggm_recver_bs(const uint32_t rk[88], int64_t t, int h, int ferret, int64_t n_out,
                                     const uint8_t* choice, const blk* csum, const blk* fin, blk* leaves, void* work,
                                     hipStream_t st);

// Host helpers (aes_host.cpp).
void cuot_aes_expand_host(const uint8_t key[16], uint32_t rk[44]);
void cuot_aes_T0_host(uint32_t T0[256]);

// Ferret 10-local LPN (lpn.hip).  nroles = 1 or 2 (both roles share indices).
struct LpnLaunch {
  uint32_t rk[44];  // AES-128 schedule of the LPN seed
  int64_t n, k;
  int ntasks;
  int nroles;
  const blk* noise[2];
  const blk* kk[2];
  const blk* kkp;  // optional: kk[0], kk[1] interleaved (2k blocks), one 32-byte gather for both roles
  blk* out[2];
};
hipError_t cuot_launch_lpn(const LpnLaunch& l, hipStream_t st);
// out[2j] = a[j], out[2j+1] = b[j], j < k (proto.hip)
hipError_t cuot_launch_interleave(const blk* a, const blk* b, int64_t k, blk* out, hipStream_t st);
// batch form: per-instance argument records (cuot_lpn_args_bytes each, filled
// by cuot_lpn_args_fill, which returns the instance's WG count or -1) and the
// prefix of WG counts, both in device memory; all instances share nroles
size_t cuot_lpn_args_bytes();
int64_t cuot_lpn_args_fill(const LpnLaunch& l, void* out);
hipError_t cuot_launch_lpn_batch(const void* tabs_d, const int64_t* wg_start_d, int ninst, int64_t total_wgs,
                                 int nroles, hipStream_t st);

// Quasi-cyclic encoder (qc.hip).
int cuot_qc_log_len(int64_t n);
size_t cuot_qc_scratch_bytes_for(int64_t n);
hipError_t cuot_launch_qc_prepare(const uint8_t* a_bits, int64_t n, uint32_t* ahat, void* scratch, hipStream_t st);
hipError_t cuot_launch_qc_encode(const uint32_t* ahat, int64_t n, const void* x, void* y, void* scratch,
                                 hipStream_t st);
// Batches of instances with the same transform length 2^m (one launch per
// step, blockIdx.y = instance).  Tables live in device memory.
struct QcIo {
  const blk* x;  // 2n input blocks
  blk* y;        // n output blocks
  int64_t n;
};
struct QcPrep {
  const uint8_t* a_bits;  // n bits of the circulant's first column
  uint32_t* ahat;         // cuot_qc_ahat_words(n) words
  int64_t n;
};
size_t cuot_qc_batch_scratch_bytes(int m, int B);
bool cuot_qc_batchable(int m);   // B > 1 allowed for transform length 2^m
hipError_t cuot_qc_configure();  // one-time device set-up (outside graph capture)
hipError_t cuot_launch_qc_prepare_batch(const QcPrep* prep_d, int B, int m, hipStream_t st);
// nmax: the largest n of the batch (host side)
hipError_t cuot_launch_qc_encode_batch(const QcIo* io_d, const uint32_t* const* ahats_d, int B, int m, int64_t nmax,
                                       void* scratch, hipStream_t st);
// Small codes on the matrix cores (qc_mfma.hip).  For 17 <= n <= 2^16 the
// prepared words are the 2^m transform words followed by a's n bits, so either
// encoder can run on them; the encode picks by cuot_qc_mfma_selected.
bool cuot_qc_mfma_eligible(int64_t n);
bool cuot_qc_mfma_selected(int64_t n);  // eligible and at or below the crossover knob
int64_t cuot_qc_mfma_limit();           // the crossover in force (knob or built-in)
int64_t cuot_qc_ahat_words(int64_t n);
hipError_t cuot_launch_qc_keep_bits(const QcPrep* pr, QcPrep pr1, int B, int m, int64_t nmax, hipStream_t st);
hipError_t cuot_launch_qc_mfma(const QcIo* io, QcIo io1, const uint32_t* bits1, const uint32_t* const* ahats, int B,
                               int m, int64_t nmax, void* scratch, hipStream_t st);

// Sparse encode and bit transpose (bits.hip).
hipError_t cuot_launch_sparse(const uint32_t* a_words, int64_t n, const int64_t* pos, int w, uint32_t* out,
                              hipStream_t st);
hipError_t cuot_launch_transpose(const void* in, int64_t N, void* out, bool inverse, hipStream_t st);
// planes: blockIdx.y; strides in blocks (block side) and 32-bit words (row side)
hipError_t cuot_launch_transpose_planes(const void* in, int64_t N, void* out, bool inverse, int planes,
                                       int64_t in_ps, int64_t out_ps, hipStream_t st);
// planes grouped by instance: y -> (y / P, y % P); offset = inst * is + p * ps
hipError_t cuot_launch_transpose_planes2(const void* in, int64_t N, void* out, bool inverse, int planes, int P,
                                        int64_t in_is, int64_t in_ps, int64_t out_is, int64_t out_ps,
                                        hipStream_t st);

// Per-launch timing (capi.hip): when a recorder is installed on the calling
// thread, instrumented launch sites bracket their kernel with HIP events on
// the launch stream and note the kernel's algorithmic bytes.
struct LaunchRecorder {
  struct Item {
    const char* name;
    hipEvent_t a, b;
    double bytes;
  };
  std::vector<Item> items;
  hipStream_t st = nullptr;
};
LaunchRecorder* cuot_recorder();
void cuot_rec_begin(hipStream_t st, const char* name, double bytes);
void cuot_rec_end(hipStream_t st);

// Execution knobs (capi.hip): process-wide, picked by the autotuner; 0 means
// the built-in rule.  Knobs never change outputs.
struct CuotKnobs {
  int qc_radix;  // FFT layers per upper pass (2..8)
  int qc_cols;   // FFT columns per tile (1, 2, 4, 8)
  int ggm_dfs;   // per-lane depth-first levels L (h - 10 <= L <= min(h - 8, 6))
  int lpn_wgs;   // LPN grid cap
  int qc_mfma_max;  // matrix-core QC encoder for n at or below (-1: never)
  int aes_impl;     // AES streams: 0 rule / 1 LDS T-table / 2 bitsliced
};
const CuotKnobs& cuot_knobs();

// Protocol glue (proto.hip).
hipError_t cuot_launch_setup_fix(blk* delta, const blk* alpha_blocks, int64_t t, int h, uint8_t* choice,
                                 int64_t* pos, uint32_t* a_words, int64_t n, int64_t a_words_total,
                                 hipStream_t st);
hipError_t cuot_launch_sot_msgs(const blk* sums, const blk* m0b, const blk* m1b, const blk* delta, int64_t t,
                                int h, blk* M0, blk* M1, hipStream_t st);
hipError_t cuot_launch_sot_recv_prep(const uint8_t* choice, const blk* M0, const blk* M1, const blk* m0b,
                                     const blk* m1b, int64_t t, int h, blk* csum, blk* fin, hipStream_t st);
hipError_t cuot_launch_ferret_sums(const blk* sums, const blk* delta, const uint8_t* choice, int64_t t, int h,
                                   blk* ssum, blk* csum, blk* lsum, blk* rsum, hipStream_t st);
hipError_t cuot_launch_ferret_boot(blk* K, blk* Mr, const uint8_t* cbits, const blk* delta, int64_t k,
                                   hipStream_t st);

// Malicious-security consistency check of a Ferret MPCOT (consist.hip):
// per-tree V, W, chi_alpha into vwc (3 t blocks), then both parties' check
// into out (x' 128 bytes, ALICE digest 32, BOB digest 32, ok int32).
hipError_t cuot_launch_consist(const blk* leaves_s, const blk* leaves_r, int64_t n, int64_t t, int h,
                               const blk* ssum, const uint8_t* bits, const blk* pre_s, const blk* pre_r,
                               const blk* delta, blk* vwc, uint8_t* out, hipStream_t st);

// Online derandomisation (derand.hip): chosen-bit COT in place, and the
// MITCCRH hash pass turning COTs into OTs (pad, out_r) or random OTs (rot0,
// rot1 non-NULL; out_r = H(M)).
hipError_t cuot_launch_cot_online(blk* ks, const blk* mr, const blk* delta, const uint8_t* b, uint8_t* bo, int64_t n,
                                  hipStream_t st);
hipError_t cuot_launch_cot_hash(const blk* ks, const blk* mr, const blk* delta, blk s, uint64_t g0, int64_t n,
                                const uint8_t* r, const blk* m0, const blk* m1, blk* pad, blk* out_r, blk* rot0,
                                blk* rot1, hipStream_t st);

// Host single-block AES-128 (aes.hip), rk from cuot_aes_expand_host.
void cuot_aes_encrypt_host(const uint32_t rk[44], const uint8_t in[16], uint8_t out[16]);
