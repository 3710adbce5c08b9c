// lpn.hip -- Ferret's 10-local primal LPN as one fused kernel:
//   out[i] = noise[i] ^ XOR_{j<10} kk[idx(i, j)]
// with the indices generated on the fly by AES under the LPN seed exactly as
// emp-ot's LpnF2 does (ferret/emp-ot/emp-ot/ferret/lpn_f2.h:46-79):
//   rows i..i+3 (a "__compute4" group): AES(makeBlock(i, m)) for m < 10 gives
//     40 words; row i+q uses words 10q..10q+9, idx = w & mask, minus k if >= k;
//   the last 1..4 rows of every task ("__compute1"): AES(makeBlock(i, m)) for
//     m < 3, idx = w[j] % k.
// Tasks split [0, n) in equal widths n/ntasks with the last task taking the
// remainder (LpnF2::bench, lpn_f2.h:100-115, commented out in the reference
// tree; each task is LpnF2::task, lpn_f2.h:72-79).  The reference GPU path materialises the index
// array and uses other counters (gpu/gpu_ops.cu:102-109, dev_layer.cu:146-186);
// this kernel keeps the emp-ot definition and never writes indices to HBM.
//
// Both roles share the (public) indices, so one launch can produce the
// sender's and the receiver's outputs from the same AES work.
#include <string.h>

#include "aes_dev.h"
#include "kernels.h"

namespace {

constexpr int D = 10;

// The gathers miss L2 and wait on the last-level cache, so residency
// matters: computing each index AES block just before the first row that
// reads it brought the kernel from 256 VGPRs (one wave per SIMD) to ~157
// (three).

struct LpnKArgs {
  const uint32_t* T0;
  uint32_t rk[44];
  int64_t n, k;
  uint32_t mask;
  int64_t width;   // n / ntasks
  int ntasks;
  int64_t units_per_task;  // units of a full-width task
  int64_t units_total;
  const blk* noise[2];
  const blk* kk[2];
  const blk* kkp;  // both roles' kk interleaved (kkp[2j], kkp[2j+1]) or NULL
  blk* out[2];
  int nroles;
};

// kk gathers: with the interleaved table one 32-byte read serves both roles
template <int NR>
CUOT_DEV void gather_kk(const LpnKArgs& a, uint32_t idx, blk* acc) {
  if (NR == 2 && a.kkp) {
    const blk* p = a.kkp + 2 * (int64_t)idx;
    acc[0] = bxor(acc[0], p[0]);
    acc[NR - 1] = bxor(acc[NR - 1], p[1]);
  } else {
#pragma unroll
    for (int rr = 0; rr < NR; rr++) acc[rr] = bxor(acc[rr], a.kk[rr][idx]);
  }
}

// number of __compute4 groups and __compute1 rows of a task of `rows` rows
CUOT_DEV void task_shape(int64_t rows, int64_t& g4, int64_t& t1) {
  g4 = rows > 4 ? (rows - 4 + 3) / 4 : 0;
  t1 = rows - 4 * g4;
}

template <int NR>
CUOT_DEV void lpn_unit(const AesLds& aes, const LpnKArgs& a, int64_t u) {
  // locate the task and the unit inside it
  int64_t task = u / a.units_per_task;
  if (task >= a.ntasks) task = a.ntasks - 1;
  const int64_t start = task * a.width;
  const int64_t end = (task == a.ntasks - 1) ? a.n : start + a.width;
  const int64_t lu = u - task * a.units_per_task;
  int64_t g4, t1;
  task_shape(end - start, g4, t1);
  if (lu < g4) {
    const int64_t i = start + 4 * lu;
    uint32_t w[4 * D];
    // row q takes words 10q..10q+9, i.e. AES blocks 10q/4 .. (10q+9)/4: each
    // block is computed just before the first row that reads it
    auto block = [&](int m) {
      blk c = make_uint4((uint32_t)m, 0u, (uint32_t)i, (uint32_t)((uint64_t)i >> 32));
      blk e = aes_enc(aes, a.rk, c);
      w[4 * m + 0] = e.x;
      w[4 * m + 1] = e.y;
      w[4 * m + 2] = e.z;
      w[4 * m + 3] = e.w;
    };
    auto row = [&](int q) {
      blk acc[NR];
#pragma unroll
      for (int rr = 0; rr < NR; rr++) acc[rr] = a.noise[rr][i + q];
#pragma unroll
      for (int j = 0; j < D; j++) {
        uint32_t idx = w[q * D + j] & a.mask;
        if (idx >= (uint32_t)a.k) idx -= (uint32_t)a.k;
        gather_kk<NR>(a, idx, acc);
      }
#pragma unroll
      for (int rr = 0; rr < NR; rr++) a.out[rr][i + q] = acc[rr];
    };
    block(0);
    block(1);
    block(2);
    row(0);
    block(3);
    block(4);
    row(1);
    block(5);
    block(6);
    block(7);
    row(2);
    block(8);
    block(9);
    row(3);
  } else if (lu - g4 < t1) {
    const int64_t i = start + 4 * g4 + (lu - g4);
    uint32_t w[12];
#pragma unroll
    for (int m = 0; m < 3; m++) {
      blk c = make_uint4((uint32_t)m, 0u, (uint32_t)i, (uint32_t)((uint64_t)i >> 32));
      blk e = aes_enc(aes, a.rk, c);
      w[4 * m + 0] = e.x;
      w[4 * m + 1] = e.y;
      w[4 * m + 2] = e.z;
      w[4 * m + 3] = e.w;
    }
    blk acc[NR];
#pragma unroll
    for (int rr = 0; rr < NR; rr++) acc[rr] = a.noise[rr][i];
#pragma unroll
    for (int j = 0; j < D; j++) {
      const uint32_t idx = w[j] % (uint32_t)a.k;
      gather_kk<NR>(a, idx, acc);
    }
#pragma unroll
    for (int rr = 0; rr < NR; rr++) a.out[rr][i] = acc[rr];
  }
}

template <int NR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1))) void lpn_kernel(LpnKArgs a) {
  __shared__ uint32_t tab[CUOT_AES_LDS_WORDS];
  aes_stage_table(tab, a.T0, threadIdx.x, 256);
  __syncthreads();
  const AesLds aes = aes_lds(tab);
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x; u < a.units_total; u += stride) lpn_unit<NR>(aes, a, u);
}

// Batch: wg_start[i] = first WG of instance i (prefix of ceil(units / 256));
// one WG covers 256 units of a single instance, so the round keys stay uniform.
template <int NR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1))) void lpn_batch_kernel(const LpnKArgs* __restrict__ tabs,
                                                        const int64_t* __restrict__ wg_start, int ninst) {
  __shared__ uint32_t tab[CUOT_AES_LDS_WORDS];
  const int64_t g = blockIdx.x;
  int lo = 0, hi = ninst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (wg_start[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  const LpnKArgs& a = tabs[lo];
  aes_stage_table(tab, a.T0, threadIdx.x, 256);
  __syncthreads();
  const AesLds aes = aes_lds(tab);
  const int64_t u = (g - wg_start[lo]) * 256 + threadIdx.x;
  if (u < a.units_total) lpn_unit<NR>(aes, a, u);
}

int64_t units_of(int64_t rows) {
  int64_t g4 = rows > 4 ? (rows - 4 + 3) / 4 : 0;
  return g4 + (rows - 4 * g4);
}

}  // namespace

namespace {

bool fill_args(const LpnLaunch& l, LpnKArgs& a) {
  memset(&a, 0, sizeof(a));
  if (l.n <= 0 || l.k < 1 || l.nroles < 1 || l.nroles > 2) return false;
  a.T0 = cuot_device_T0();
  memcpy(a.rk, l.rk, sizeof(a.rk));
  a.n = l.n;
  a.k = l.k;
  uint32_t mask = 1;  // lpn_f2.h:27-31
  while ((int64_t)mask < l.k) mask = (mask << 1) | 1u;
  a.mask = mask;
  a.ntasks = l.ntasks < 1 ? 1 : l.ntasks;
  a.width = l.n / a.ntasks;
  const int64_t last_rows = l.n - (int64_t)(a.ntasks - 1) * a.width;
  a.units_per_task = units_of(a.width);
  if (a.units_per_task < 1) a.units_per_task = 1;
  // full-width tasks then the last one; the kernel maps u -> task by division,
  // so the last task's units start at (ntasks-1)*units_per_task
  a.units_total = (int64_t)(a.ntasks - 1) * a.units_per_task + units_of(last_rows);
  a.nroles = l.nroles;
  for (int r = 0; r < l.nroles; r++) {
    a.noise[r] = l.noise[r];
    a.kk[r] = l.kk[r];
    a.kkp = l.nroles == 2 ? l.kkp : nullptr;
    a.out[r] = l.out[r];
  }
  return true;
}

}  // namespace

hipError_t cuot_launch_lpn(const LpnLaunch& l, hipStream_t st) {
  if (l.n <= 0) return hipSuccess;
  LpnKArgs a;
  if (!fill_args(l, a)) return hipErrorInvalidValue;
  int64_t g = (a.units_total + 255) / 256;
  const int64_t cap = cuot_knobs().lpn_wgs > 0 ? cuot_knobs().lpn_wgs : 8192;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  // algorithmic bytes per role and row: noise in, d gathers, row out
  cuot_rec_begin(st, "lpn_kernel", 16.0 * (D + 2) * (double)l.n * l.nroles);
  if (l.nroles == 2)
    hipLaunchKernelGGL(lpn_kernel<2>, dim3((unsigned)g), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(lpn_kernel<1>, dim3((unsigned)g), dim3(256), 0, st, a);
  cuot_rec_end(st);
  return hipGetLastError();
}

size_t cuot_lpn_args_bytes() { return sizeof(LpnKArgs); }

int64_t cuot_lpn_args_fill(const LpnLaunch& l, void* out) {
  LpnKArgs a;
  if (!fill_args(l, a)) return -1;
  memcpy(out, &a, sizeof(a));
  return (a.units_total + 255) / 256;
}

hipError_t cuot_launch_lpn_batch(const void* tabs_d, const int64_t* wg_start_d, int ninst, int64_t total_wgs,
                                 int nroles, hipStream_t st) {
  if (ninst < 1 || total_wgs < 1 || total_wgs > 0x7fffffff) return hipErrorInvalidValue;
  if (nroles == 2)
    hipLaunchKernelGGL(lpn_batch_kernel<2>, dim3((unsigned)total_wgs), dim3(256), 0, st, (const LpnKArgs*)tabs_d,
                       wg_start_d, ninst);
  else
    hipLaunchKernelGGL(lpn_batch_kernel<1>, dim3((unsigned)total_wgs), dim3(256), 0, st, (const LpnKArgs*)tabs_d,
                       wg_start_d, ninst);
  return hipGetLastError();
}
