/*
 * oracle.h -- CPU restatement of the silent-OT correlation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (cuot_amd) never links, loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).  Where the reference
 * GPU code deviates from the protocol definition (non-FIPS key schedule,
 * missing feed-forward, float FFT, ...) the oracle follows the protocol
 * definition that the CPU code (emp-ot) and the workload description fix.
 *
 * Pinning status (what checks each function independently of its own text):
 *   or_aes_*         FIPS-197 App. A.1 / C.1 vectors, OpenSSL EVP ECB,
 *                    portable vs AES-NI implementations, and the reference's
 *                    vendored cryptoTools AES (oracle/_ref) at random keys.
 *   or_prg           defined here (seed derivation of the workload); pinned
 *                    through or_aes_* only.
 *   or_ggm_*         brute-force recursive GGM on tiny trees; level-sum
 *                    invariants; sender/receiver puncture invariant.
 *   or_sot_*         w = v xor e*Delta invariant; two independent receiver
 *                    restatements (SOT fill rule vs emp-ot reconstruction).
 *   or_qc_*          O(n^2) brute force vs CLMUL Karatsuba vs special cases
 *                    (a = 1, a = x^s), linearity.
 *   or_ferret_lpn    hand-checked small case; partition invariance for
 *                    rows handled by __compute4 vs __compute1 forms.
 *   or_bit_transpose involution and per-bit brute force.
 *   emp-tool internals (makeBlock order, ParaEnc key assignment) are read
 *   from call sites only: parity unpinned against emp-tool itself, which is
 *   not present in the reference tree (emp-tool @44b1dde, not vendored).
 */
#ifndef CUOT_ORACLE_H
#define CUOT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 128-bit block in memory byte order: bytes 0..7 = lo (little endian),
 * bytes 8..15 = hi.  emp-tool makeBlock(hi, lo) == {lo, hi}
 * (spcot_sender.h:52 clears bit 0 of byte 0 through makeBlock(~0, ~0-1)). */
typedef struct {
  uint64_t lo, hi;
} orb;

/* ---- AES-128 (FIPS-197) ------------------------------------------------ */
void or_aes_expand(const orb *key, uint8_t rk[176]);
/* impl: 0 = portable byte-oriented, 1 = AES-NI (falls back to 0 if absent) */
void or_aes_ecb(const uint8_t rk[176], const orb *in, orb *out, size_t n, int impl);
int or_have_aesni(void);

/* ---- seed derivation (workload definition) ------------------------------ */
/* out[i] = AES_seed( {lo = start + i, hi = stream} ) */
void or_prg(const orb *seed, uint64_t stream, uint64_t start, size_t count, orb *out);

/* ---- GGM tree, two-key PRG (twokeyprp.h:7-63) -------------------------- */
/* One tree of depth h from root.  leaves: 2^h blocks.  lsum/rsum: h blocks,
 * lsum[d] = xor of even nodes at depth d+1, rsum[d] = xor of odd nodes. */
void or_ggm_sender_tree(const uint8_t rk0[176], const uint8_t rk1[176], orb root,
                        int h, orb *leaves, orb *lsum, orb *rsum);

/* Ferret MPCOT, emp-ot semantics, t trees of depth h, leaves truncated to
 * n_out (<= t*2^h).  Sender: spcot_sender.h:49-86; receiver:
 * spcot_recver.h:53-95.  b[t*h + d] = receiver choice bit at level d. */
void or_ferret_mpcot_sender(const orb *roots, int64_t t, int h, orb delta, int64_t n_out,
                            orb *leaves, orb *lsum, orb *rsum, orb *secret_sum, int nthreads);
void or_ferret_mpcot_recver(int64_t t, int h, const uint8_t *b, const orb *csum,
                            const orb *secret_sum, int64_t n_out, orb *leaves, int nthreads);

/* SOT seed expansion (silent/sender.cu:76-108, silent/recver.cu:98-163)
 * with the config's child = AES_k(s) xor s.  m0b/m1b: base-OT messages
 * [h][t]; c[d*t + tr] receiver choice bits; M0/M1 out: [h+1][t]. */
void or_sot_sender_expand(const uint8_t rk0[176], const uint8_t rk1[176], const orb *roots,
                          int64_t t, int h, orb delta, const orb *m0b, const orb *m1b,
                          int64_t n_leaves, orb *leaves, orb *M0, orb *M1, int nthreads);
void or_sot_recver_expand(const uint8_t rk0[176], const uint8_t rk1[176], int64_t t, int h,
                          const uint8_t *c, const orb *M0, const orb *M1, const orb *m0b,
                          const orb *m1b, int64_t n_leaves, orb *leaves, int nthreads);

/* ---- quasi-cyclic encoder (config semantics, not quasi_cyclic.cu's float
 * path): y_i = x_i xor  XOR_{j : a_j = 1} x_{n + ((i - j) mod n)},  i < n.
 * x: 2n blocks, a: n bits packed LSB-first, y: n blocks. */
void or_qc_encode_brute(const orb *x, const uint8_t *a, int64_t n, orb *y);
void or_qc_encode_fast(const orb *x, const uint8_t *a, int64_t n, orb *y, int nthreads);
/* bits version used by the sparse encoder: e: 2n bits, out: n bits */
void or_qc_encode_bits_brute(const uint8_t *e, const uint8_t *a, int64_t n, uint8_t *out);
/* sparse encode (quasi_cyclic.cu:130-161 semantics on the config's code):
 * noise positions pos[0..w) in [0, 2n); out: n bits */
void or_qc_sparse(const int64_t *pos, int64_t w, const uint8_t *a, int64_t n, uint8_t *out);

/* ---- binary polynomial product (test helper of or_qc_encode_fast) ------- */
void or_bpmul(const uint64_t *a, const uint64_t *b, int64_t m, uint64_t *out);

/* ---- 128-row bit transpose (gpu/gpu_ops.cu:31-70 bit order) ------------- */
/* rows[k][m/8] bit (m%8) = bit k of block m (bit k = byte k/8, bit k%8). N % 8 == 0 */
void or_bit_transpose(const orb *in, int64_t N, uint8_t *rows);
void or_bit_untranspose(const uint8_t *rows, int64_t N, orb *out);

/* ---- Ferret 10-local primal LPN (lpn_f2.h:46-79) --------------------- */
/* nn[i] ^= XOR_j kk[idx(i,j)] for i in [0, n); PRP keyed with seed;
 * ntasks equal-width tasks as in LpnF2::bench (lpn_f2.h:100-115, commented out
 * in the reference tree; each task runs LpnF2::task, lpn_f2.h:72-79). */
void or_ferret_lpn(orb *nn, int64_t n, const orb *kk, int64_t k, orb seed, int d, int ntasks,
                   int nthreads);
/* indices only (for tests): idx[i*d + j] */
void or_ferret_lpn_indices(int64_t n, int64_t k, orb seed, int d, int ntasks, uint32_t *idx);

/* ---- whole instances from one seed (workload definition) ----------------
 * Seed derivation, stream ids of or_prg(S, stream, index):
 *   0: Delta = block 0 with bit 0 set (ferret_cot.hpp:27-31 forces LSB 1)
 *   1: GGM roots, one per tree
 *   2: punctured index alpha_t = (low 32 bits of block t) mod 2^h
 *   3: a (QC circulant), bit i = bit i%128 of block i/128, i < n
 *   4, 5: base-OT messages m0b[d][t], m1b[d][t] (index d*t_count + t)
 *   6: Ferret LPN seed (block 0)
 *   7: Ferret bootstrap K_j with bit 0 cleared (base_cot.h:53-58)
 *   8: Ferret bootstrap choice c_j = bit j%128 of block j/128;
 *      receiver's M_j = K_j ^ c_j * Delta (LSB = c_j, base_cot.h:59-70)
 * The receiver's choice at level d is c = 1 - bit (h-1-d) of alpha
 * (silent/recver.cu:78-87, spcot_recver.h:33-41). */
typedef struct {
  int64_t n; /* COTs out; leaves = 2n = t * 2^h (last tree truncated) */
  int64_t t;
  int h;
  orb k0, k1;
  orb seed;
} or_sot_cfg;

/* v, w: n blocks; b: (n+7)/8 bytes; delta: 1 block.  Also returns the raw
 * leaves if leaves_s/leaves_r are non-NULL (2n blocks each). */
void or_sot_run(const or_sot_cfg *cfg, int nthreads, orb *v, orb *w, uint8_t *b, orb *delta, orb *leaves_s,
                orb *leaves_r);

typedef struct {
  int64_t n, k;
  int64_t t;
  int h;
  int ntasks;
  orb seed;
} or_ferret_cfg;

/* out_s, out_r: n blocks; delta: 1 block */
void or_ferret_run(const or_ferret_cfg *cfg, int nthreads, orb *out_s, orb *out_r, orb *delta);


/* ---- iterating Ferret (FerretCOT::extend_f2k, ferret_cot.hpp:79-103) ------
 * State: Delta and a bootstrap of M = k + h*t + 128 COTs per party (sender K
 * with LSB 0, receiver Mr = K ^ c*Delta).  One extend:
 *   pre-OT (preot.h:38-54): P0 = H(K[i]), P1 = H(K[i]^Delta) (sender), bits =
 *     LSB(Mr[i]), Pr = H(Mr[i]) (receiver), i < h*t, H = the circular
 *     correlation-robust hash H(x) = AES0(s(x)) ^ s(x), s(x) = (hi ^ lo, hi)
 *     as (hi, lo) halves, AES0 = AES-128 under the all-zero key;
 *   MPCOT: roots = PRG(seed, stream 1, start r*t); level sums sent as
 *     m0 ^ P0, m1 ^ P1; the receiver's choice bits are bits[tree*h + level]
 *     and its chosen sum pad_b ^ Pr; secret sum as in the single instance;
 *   LPN with kk = bootstrap + 128 (ferret_cot.hpp:89), seed PRG(seed, 6, 0);
 *   the last M outputs become the next bootstrap (ferret_cot.hpp:100).
 * Initial bootstrap: K = PRG(seed, 7) & ~1, c = PRG(seed, 8) bits, M blocks. */
typedef struct {
  or_ferret_cfg cfg;
  int64_t M, ot_limit, round;
  orb delta;
  orb *boot_s, *boot_r; /* M blocks each */
  /* malicious mode: the last extend's consistency check */
  int malicious, check_ok;
  uint8_t xprime[128], dig_a[32], dig_b[32];
} or_ferret_session;

int or_ferret_session_init(or_ferret_session *s, const or_ferret_cfg *cfg);
/* one extend: out_s, out_r receive all n outputs (the first ot_limit are the
 * user's; the rest become the bootstrap) */
void or_ferret_session_extend(or_ferret_session *s, int nthreads, orb *out_s, orb *out_r);
void or_ferret_session_free(or_ferret_session *s);
void or_ccrh(const orb *in, orb *out, size_t n);

/* ---- malicious consistency check (mpcot_reg.h:177-220) ---------------------
 * emp-tool's Hash, gfmul, uni_hash_coeff_gen, vector_inn_prdt_sum_red and
 * GaloisFieldPacking restated (emp-tool @44b1dde is absent).
 * GF(2^128) = GF(2)[x]/(x^128 + x^7 + x^2 + x + 1), bit i = coefficient x^i.
 * Per tree: seed = SHA-256(secret_sum)[0:16] (spcot_sender.h:91-92),
 * chi_j = seed^(j+1), V = sum chi_j leaf_s[j], W = sum chi_j leaf_r[j],
 * chi_alpha = chi at the punctured leaf (spcot_recver.h:104).  Leaves at a
 * global index >= n are not stored and are left out, and so is a chi_alpha
 * whose leaf is not stored (emp-ot's own parameters store every leaf).
 * bits: the receiver's choice bits [tree * h + level]. */
void or_sha256(const uint8_t *msg, size_t len, uint8_t out[32]);
orb or_gf128_mul(orb a, orb b);
void or_ferret_consist_trees(const orb *leaves_s, const orb *leaves_r, int64_t n, int64_t t, int h,
                             const orb *ssum, const uint8_t *bits, orb *V, orb *W, orb *chi_alpha, int nthreads);
/* consistency_check_f2k for both parties: x' = bits of XOR chi_alpha ^
 * LSB(pre_r); digests of XOR V ^ packing(pre_s ^ x' Delta) (ALICE) and
 * XOR W ^ packing(pre_r) (BOB), packing(d) = sum_j d_j x^j; returns 1 when
 * they agree */
/* ---- online derandomisation -----------------------------------------------
 * or_cot_online: FerretCOT::online_sender/online_recver (ferret_cot.hpp:277-295),
 * bo = b ^ LSB(data_r), data_s ^= bo * Delta.  or_cot_to_ot / or_cot_to_rot:
 * COT::send/recv and send_rot/recv_rot (cot.h:16-91) with emp-tool's MITCCRH
 * H_i(x) = AES_{s ^ makeBlock(g0 + i, 0)}(x) ^ x (mitccrh.h, emp-tool
 * @44b1dde restated: parity unpinned for the key derivation). */
void or_cot_online(orb *data_s, const orb *data_r, const uint8_t *b, orb delta, uint8_t *bo, int64_t n);
void or_cot_to_ot(const orb *data_s, const orb *data_r, orb delta, orb s, uint64_t g0, const uint8_t *r,
                  const orb *m0, const orb *m1, orb *pad, orb *out_r, int64_t n, int nthreads);
void or_cot_to_rot(const orb *data_s, const orb *data_r, orb delta, orb s, uint64_t g0, orb *data0, orb *data1,
                   orb *out_r, int64_t n, int nthreads);
int or_ferret_consist_final(const orb *V, const orb *W, const orb *chi_alpha, int64_t t, const orb *pre_s,
                            const orb *pre_r, orb delta, uint8_t xprime[128], uint8_t dig_a[32],
                            uint8_t dig_b[32]);

#ifdef __cplusplus
}
#endif
#endif
