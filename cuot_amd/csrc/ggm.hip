// ggm.hip -- fused breadth-first GGM tree expansion with level sums, for the
// sender (full trees) and the receiver (punctured trees rebuilt from its own
// view: chosen level sums + choice bits), in SOT and Ferret flavours.
//
// One workgroup of 256 lanes owns one tree (the reference launches one kernel
// per level plus a multi-pass reduction per level; gpu/aes_op.cu:176-266,
// gpu/gpu_matrix.cu:155-169).  Per tree:
//   A. top levels 0..top breadth-first in LDS (top = h for h <= 8, else
//      top = h - L with L = min(h - 8, 6), so 2^top is 256..1024 nodes), level
//      sums by a workgroup XOR reduction; the receiver resolves the sibling of
//      its punctured path at every level (silent/recver.cu:98-115,
//      spcot_recver.h:66-95).
//   B. (h > 8) every lane expands 2^top/256 depth-`top` nodes depth-first for
//      L levels, with a fully unrolled path stack in registers; leaves go
//      straight to HBM, per-depth left/right sums are reduced over the
//      workgroup.  The punctured depth-`top` node contributes nothing.
//   C. (receiver, h > 8) the punctured depth-`top` subtree is rebuilt level
//      by level in LDS from the chosen sums and the workgroup sums of phase B.
// Child rule: left = AES_k0(p) ^ p, right = AES_k1(p) ^ p (twokeyprp.h:7-23).
#include "aes_dev.h"
#include "kernels.h"

namespace {

constexpr int WG = 256;
constexpr int LMAX = 6;      // deepest per-lane DFS (register budget)
constexpr int MAXTOP = 10;   // deepest LDS breadth-first level (h <= 16)

CUOT_DEV uint32_t wave_xor(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v ^= (uint32_t)__shfl_xor((int)v, o, 64);
  return v;
}

// XOR-reduce two blocks over the 256-lane workgroup; result in every lane.
// red: LDS scratch of 4 waves x 8 words.  Contains two barriers.
CUOT_DEV void wg_xor2(blk& a, blk& b, uint32_t* red) {
  uint32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = wave_xor(v[i]);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 8; i++) red[wave * 8 + i] = v[i];
  }
  __syncthreads();
  uint32_t r[8];
#pragma unroll
  for (int i = 0; i < 8; i++) r[i] = red[i] ^ red[8 + i] ^ red[16 + i] ^ red[24 + i];
  a = make_uint4(r[0], r[1], r[2], r[3]);
  b = make_uint4(r[4], r[5], r[6], r[7]);
}

CUOT_DEV blk clear_lsb(blk v) {
  v.x &= ~1u;
  return v;
}

struct Lds {
  uint32_t aes[CUOT_AES_LDS_WORDS];
  blk nodes[1 << MAXTOP];
  blk sums[2 * 16];  // [depth][L,R] for depths 1..16 (index depth-1)
  uint32_t red[32];
};

// Breadth-first expansion in LDS of `nlev` levels below nodes[0].
// Receiver mode resolves the sibling on the punctured path at each level
// using chosen sums csum[d0 + r] and external (outside-LDS) side sums ext.
// Returns the punctured index at the last level (receiver) in *punct_out.
// `final_rule`: apply the punctured-leaf rule at the last level.
template <bool RECV>
CUOT_DEV void bfs_levels(const AesLds& aes, const GgmKArgs& a, Lds& s, int64_t tree, int d0, int nlev,
                         const blk* ext, bool final_rule, int* punct_out) {
  const int tid = threadIdx.x;
  int punct = 0;
  for (int r = 0; r < nlev; r++) {
    const int w = 1 << r;
    // each lane expands parents tid, tid+256, ... (at most 2 for 2^10 children)
    blk lc[2], rc[2];
    blk L = bzero(), R = bzero();
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int j = tid + q * WG;
      lc[q] = bzero();
      rc[q] = bzero();
      if (j < w) {
        blk p = s.nodes[j];
        ggm_expand(aes, a.rk, p, lc[q], rc[q]);
        if (RECV && j == punct) {
          lc[q] = bzero();
          rc[q] = bzero();
        }
      }
      L = bxor(L, lc[q]);
      R = bxor(R, rc[q]);
    }
    wg_xor2(L, R, s.red);  // barrier inside: all reads of nodes[] are done
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int j = tid + q * WG;
      if (j < w) {
        s.nodes[2 * j] = lc[q];
        s.nodes[2 * j + 1] = rc[q];
      }
    }
    const int d = d0 + r;  // sum index: depth d+1
    if (!RECV) {
      if (tid == 0) {
        s.sums[2 * d] = L;
        s.sums[2 * d + 1] = R;
      }
    } else {
      if (ext) {
        L = bxor(L, ext[2 * r]);
        R = bxor(R, ext[2 * r + 1]);
      }
      const int c = a.choice[tree * a.h + d] ? 1 : 0;
      const int sib = 2 * punct + c;
      const int np = 2 * punct + (1 - c);
      blk sibv = bxor(a.csum[tree * a.h + d], c ? R : L);
      __syncthreads();
      if (tid == 0) {
        s.nodes[sib] = sibv;
        if (final_rule && r == nlev - 1) {
          blk f = a.fin[tree];
          if (a.mode == CUOT_MODE_SOT) {
            // silent/recver.cu:106-112 final layer: opposite-side sum
            s.nodes[np] = bxor(f, c ? L : R);
          } else {
            // spcot_recver.h:56-63: all other leaves, LSB cleared, xor secret
            s.nodes[np] = bxor(f, clear_lsb(bxor(bxor(L, R), sibv)));
          }
        }
      }
      punct = np;
    }
    __syncthreads();
  }
  *punct_out = punct;
}

// Unrolled depth-first expansion of one node for L levels (phase B).
// Leaf pairs jj = 0 .. 2^(L-1)-1 are produced in order.  Going from pair jj-1
// to jj, the path changes below relative depth r0 = L-1-ctz(jj): the node at
// r0 is the pending right child saved when its parent was expanded, and every
// node below it is a left child expanded on the spot.  All branches depend on
// jj only, so they are uniform across the wave and every array index is
// static after unrolling (no scratch).
// Sums of this subtree are xored into sl/sr[1..L] only when `active`.
template <int L>
CUOT_DEV void dfs_subtree(const AesLds& aes, const uint32_t* rk, blk root, bool active, blk* leaves,
                          int64_t leaf0, int64_t n_out, bool ferret, blk* sl, blk* sr) {
  blk rightc[L > 1 ? L : 1];
  const uint32_t npairs = 1u << (L - 1);
  for (uint32_t jj = 0; jj < npairs; jj++) {
    const int r0 = (jj == 0) ? 0 : (L - 1 - __builtin_ctz(jj));
    blk cur = root;
#pragma unroll
    for (int r = 1; r <= L - 1; r++) {
      if (r == r0) {
        cur = rightc[r];
      } else if (r > r0) {
        blk lc, rc;
        ggm_expand(aes, rk, cur, lc, rc);
        cur = lc;
        rightc[r] = rc;
        sl[r] = bxor(sl[r], bsel(active, lc));
        sr[r] = bxor(sr[r], bsel(active, rc));
      }
    }
    blk lf, rf;
    ggm_expand(aes, rk, cur, lf, rf);
    sl[L] = bxor(sl[L], bsel(active, lf));
    sr[L] = bxor(sr[L], bsel(active, rf));
    if (active) {
      const int64_t g = leaf0 + 2 * (int64_t)jj;
      if (ferret) {
        lf = clear_lsb(lf);
        rf = clear_lsb(rf);
      }
      if (g < n_out) stb(leaves + g, lf);
      if (g + 1 < n_out) stb(leaves + g + 1, rf);
    }
  }
}

template <int L, bool RECV>
CUOT_DEV void ggm_tree(const GgmKArgs& a, Lds& s, int64_t tree) {
  const int tid = threadIdx.x;
  aes_stage_table(s.aes, a.T0, tid, WG);
  if (tid == 0) s.nodes[0] = RECV ? bzero() : a.roots[tree];
  __syncthreads();
  const AesLds aes = aes_lds(s.aes);
  const int h = a.h;
  const bool ferret = a.mode == CUOT_MODE_FERRET;
  const int top = h - L;  // L == 0 iff h <= 8
  int punct = 0;
  bfs_levels<RECV>(aes, a, s, tree, 0, top, nullptr, L == 0, &punct);
  const int64_t tree0 = tree << h;

  if constexpr (L == 0) {
    // whole tree fits the top phase: write leaves from LDS
    if (tid < (1 << h)) {
      blk v = s.nodes[tid];
      if (ferret && !(RECV && tid == punct)) v = clear_lsb(v);
      const int64_t g = tree0 + tid;
      if (g < a.n_out) stb(a.leaves + g, v);
    }
  } else {
    // ---- phase B: per-lane depth-first subtrees --------------------------
    blk tl[L + 1], tr[L + 1];
#pragma unroll
    for (int r = 0; r <= L; r++) {
      tl[r] = bzero();
      tr[r] = bzero();
    }
    const int nf = 1 << (top - 8);  // frontier nodes per lane: 1, 2 or 4
    for (int q = 0; q < nf; q++) {
      const int f = tid + q * WG;
      const bool active = !(RECV && f == punct);
      dfs_subtree<L>(aes, a.rk, s.nodes[f], active, a.leaves, tree0 + ((int64_t)f << L), a.n_out, ferret, tl,
                     tr);
    }
#pragma unroll
    for (int r = 1; r <= L; r++) {
      wg_xor2(tl[r], tr[r], s.red);
      if (tid == 0) {
        s.sums[2 * (top + r - 1)] = tl[r];
        s.sums[2 * (top + r - 1) + 1] = tr[r];
      }
    }
    __syncthreads();
    if constexpr (RECV) {
      // ---- phase C: rebuild the punctured depth-`top` subtree -------------
      const int pt = punct;
      if (tid == 0) s.nodes[0] = bzero();
      __syncthreads();
      int pl = 0;
      bfs_levels<true>(aes, a, s, tree, top, L, &s.sums[2 * top], true, &pl);
      const int64_t base = tree0 + ((int64_t)pt << L);
      if (tid < (1 << L)) {
        blk v = s.nodes[tid];
        if (ferret && tid != pl) v = clear_lsb(v);
        const int64_t g = base + tid;
        if (g < a.n_out) stb(a.leaves + g, v);
      }
    }
  }
  if constexpr (!RECV) {
    __syncthreads();
    if (tid < 2 * h) a.sums[tree * 2 * h + tid] = s.sums[tid];
  }
}

template <int L, bool RECV>
__global__ __launch_bounds__(WG) void ggm_kernel(GgmKArgs a) {
  __shared__ Lds s;
  ggm_tree<L, RECV>(a, s, blockIdx.x);
}

// Batch: one WG per tree of every instance in the table; start[i] is the first
// global tree of instance i (start[ninst] = total).  All instances share L.
template <int L, bool RECV>
__global__ __launch_bounds__(WG) void ggm_batch_kernel(const GgmKArgs* __restrict__ tab,
                                                       const int64_t* __restrict__ start, int ninst) {
  __shared__ Lds s;
  const int64_t g = blockIdx.x;
  int lo = 0, hi = ninst - 1;  // last i with start[i] <= g
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  ggm_tree<L, RECV>(tab[lo], s, g - start[lo]);
}

int ggm_split(int h) {
  if (h <= 8) return 0;
  const int k = cuot_knobs().ggm_dfs;
  const int hi = h - 8 < LMAX ? h - 8 : LMAX;
  if (k >= 1 && k <= hi && h - k <= MAXTOP) return k;
  return hi;
}

template <bool RECV>
hipError_t launch_ggm(const GgmKArgs& a, hipStream_t st) {
  const int L = ggm_split(a.h);
  dim3 grid((unsigned)a.t), block(WG);
  const int64_t nl = (a.t << a.h) < a.n_out ? (a.t << a.h) : a.n_out;
  cuot_rec_begin(st, RECV ? "ggm_kernel<recv>" : "ggm_kernel<send>", 16.0 * (double)nl);
  switch (L) {
#define CUOT_GGM_CASE(X) \
  case X:                \
    hipLaunchKernelGGL((ggm_kernel<X, RECV>), grid, block, 0, st, a); \
    break;
    CUOT_GGM_CASE(0)
    CUOT_GGM_CASE(1)
    CUOT_GGM_CASE(2)
    CUOT_GGM_CASE(3)
    CUOT_GGM_CASE(4)
    CUOT_GGM_CASE(5)
    CUOT_GGM_CASE(6)
#undef CUOT_GGM_CASE
    default:
      cuot_rec_end(st);
      return hipErrorInvalidValue;
  }
  cuot_rec_end(st);
  return hipGetLastError();
}

template <bool RECV>
hipError_t launch_ggm_batch(const GgmKArgs* tab, const int64_t* start, int ninst, int64_t total, int L,
                            hipStream_t st) {
  dim3 grid((unsigned)total), block(WG);
  switch (L) {
#define CUOT_GGM_CASE(X) \
  case X:                \
    hipLaunchKernelGGL((ggm_batch_kernel<X, RECV>), grid, block, 0, st, tab, start, ninst); \
    break;
    CUOT_GGM_CASE(0)
    CUOT_GGM_CASE(1)
    CUOT_GGM_CASE(2)
    CUOT_GGM_CASE(3)
    CUOT_GGM_CASE(4)
    CUOT_GGM_CASE(5)
    CUOT_GGM_CASE(6)
#undef CUOT_GGM_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

int cuot_ggm_split(int h) { return ggm_split(h); }

hipError_t cuot_launch_ggm(const GgmKArgs& a, bool receiver, hipStream_t st) {
  return receiver ? launch_ggm<true>(a, st) : launch_ggm<false>(a, st);
}

hipError_t cuot_launch_ggm_batch(const GgmKArgs* tab_d, const int64_t* start_d, int ninst, int64_t total_trees,
                                 int L, bool receiver, hipStream_t st) {
  if (ninst < 1 || total_trees < 1 || total_trees > 0x7fffffff) return hipErrorInvalidValue;
  return receiver ? launch_ggm_batch<true>(tab_d, start_d, ninst, total_trees, L, st)
                  : launch_ggm_batch<false>(tab_d, start_d, ninst, total_trees, L, st);
}
