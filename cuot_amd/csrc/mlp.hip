// mlp.hip -- the autotuner's cost model: a small fully connected network
// (F inputs -> H tanh -> H tanh -> 1 linear) that predicts log kernel time
// from instance sizes and execution knobs, trained on measured timings.
// It never touches correlation data.
//
// Parameters are one float32 vector, layout (row-major weights):
//   W1[H][F], b1[H], W2[H][H], b2[H], w3[H], b3
// Training is deterministic: per-sample activations and deltas are written by
// one kernel (thread per sample), per-parameter gradients are summed over
// samples in index order by a second kernel (thread per parameter).
#include "ctx.h"
#include "kernels.h"

namespace {

constexpr int MAXF = 32;
constexpr int MAXH = 64;

struct Dims {
  int F, H;
  int oW1, ob1, oW2, ob2, ow3, ob3, P;
};

Dims dims(int F, int H) {
  Dims d;
  d.F = F;
  d.H = H;
  d.oW1 = 0;
  d.ob1 = d.oW1 + H * F;
  d.oW2 = d.ob1 + H;
  d.ob2 = d.oW2 + H * H;
  d.ow3 = d.ob2 + H;
  d.ob3 = d.ow3 + H;
  d.P = d.ob3 + 1;
  return d;
}

// forward for one sample; a1/a2 receive the tanh activations
__device__ float fwd_one(const float* __restrict__ p, const Dims& d, const float* __restrict__ x, float* a1,
                         float* a2) {
  for (int i = 0; i < d.H; i++) {
    float s = p[d.ob1 + i];
    for (int j = 0; j < d.F; j++) s += p[d.oW1 + i * d.F + j] * x[j];
    a1[i] = tanhf(s);
  }
  for (int i = 0; i < d.H; i++) {
    float s = p[d.ob2 + i];
    for (int j = 0; j < d.H; j++) s += p[d.oW2 + i * d.H + j] * a1[j];
    a2[i] = tanhf(s);
  }
  float y = p[d.ob3];
  for (int j = 0; j < d.H; j++) y += p[d.ow3 + j] * a2[j];
  return y;
}

__global__ void mlp_forward_kernel(const float* __restrict__ params, Dims d, const float* __restrict__ x, int64_t N,
                                   float* __restrict__ y) {
  float a1[MAXH], a2[MAXH];
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < N; s += (int64_t)gridDim.x * blockDim.x)
    y[s] = fwd_one(params, d, x + s * d.F, a1, a2);
}

// per sample: activations a1, a2 and deltas g1 = dL/dz1, g2 = dL/dz2, gy = dL/dy
// for L = mean (y - t)^2.  act: [N][4H + 1]
__global__ void mlp_sample_kernel(const float* __restrict__ params, Dims d, const float* __restrict__ x,
                                  const float* __restrict__ target, int64_t N, float* __restrict__ act,
                                  float* __restrict__ err) {
  float a1[MAXH], a2[MAXH], g2[MAXH];
  const int R = 4 * d.H + 1;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < N; s += (int64_t)gridDim.x * blockDim.x) {
    const float y = fwd_one(params, d, x + s * d.F, a1, a2);
    const float e = y - target[s];
    err[s] = e * e;
    const float gy = 2.0f * e / (float)N;
    float* o = act + s * R;
    for (int i = 0; i < d.H; i++) {
      g2[i] = gy * params[d.ow3 + i] * (1.0f - a2[i] * a2[i]);
      o[i] = a1[i];
      o[d.H + i] = a2[i];
      o[3 * d.H + i] = g2[i];
    }
    for (int j = 0; j < d.H; j++) {
      float s1 = 0.0f;
      for (int i = 0; i < d.H; i++) s1 += params[d.oW2 + i * d.H + j] * g2[i];
      o[2 * d.H + j] = s1 * (1.0f - a1[j] * a1[j]);
    }
    o[4 * d.H] = gy;
  }
}

// thread per parameter: sum over samples in index order; thread P sums the loss
__global__ void mlp_grad_kernel(Dims d, const float* __restrict__ x, const float* __restrict__ act,
                                const float* __restrict__ err, int64_t N, float* __restrict__ grad,
                                float* __restrict__ loss) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int R = 4 * d.H + 1;
  if (q > d.P) return;
  float g = 0.0f;
  if (q == d.P) {
    for (int64_t s = 0; s < N; s++) g += err[s];
    *loss = g / (float)N;
    return;
  }
  if (q < d.ob1) {  // W1[i][j]: g1[i] * x[j]
    const int i = q / d.F, j = q % d.F;
    for (int64_t s = 0; s < N; s++) g += act[s * R + 2 * d.H + i] * x[s * d.F + j];
  } else if (q < d.oW2) {
    const int i = q - d.ob1;
    for (int64_t s = 0; s < N; s++) g += act[s * R + 2 * d.H + i];
  } else if (q < d.ob2) {  // W2[i][j]: g2[i] * a1[j]
    const int i = (q - d.oW2) / d.H, j = (q - d.oW2) % d.H;
    for (int64_t s = 0; s < N; s++) g += act[s * R + 3 * d.H + i] * act[s * R + j];
  } else if (q < d.ow3) {
    const int i = q - d.ob2;
    for (int64_t s = 0; s < N; s++) g += act[s * R + 3 * d.H + i];
  } else if (q < d.ob3) {  // w3[j]: gy * a2[j]
    const int j = q - d.ow3;
    for (int64_t s = 0; s < N; s++) g += act[s * R + 4 * d.H] * act[s * R + d.H + j];
  } else {
    for (int64_t s = 0; s < N; s++) g += act[s * R + 4 * d.H];
  }
  grad[q] = g;
}

__global__ void mlp_adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                float* __restrict__ v, int P, float lr, float b1, float b2, float eps, float c1,
                                float c2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P) return;
  const float gq = g[q];
  const float mq = b1 * m[q] + (1.0f - b1) * gq;
  const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
  m[q] = mq;
  v[q] = vq;
  p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

unsigned grid_of(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (unsigned)g;
}

}  // namespace

static int mlp_nparams(int F, int H) {
  if (F < 1 || F > MAXF || H < 1 || H > MAXH) return -1;
  return dims(F, H).P;
}

static hipError_t launch_forward(const float* params, int F, int H, const float* x, int64_t N, float* y,
                                   hipStream_t st) {
  if (mlp_nparams(F, H) < 0 || N < 0) return hipErrorInvalidValue;
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(grid_of(N)), dim3(256), 0, st, params, dims(F, H), x, N, y);
  return hipGetLastError();
}

static hipError_t launch_grad(const float* params, int F, int H, const float* x, const float* target, int64_t N,
                                float* act, float* err, float* grad, float* loss, hipStream_t st) {
  if (mlp_nparams(F, H) < 0 || N < 1) return hipErrorInvalidValue;
  const Dims d = dims(F, H);
  hipLaunchKernelGGL(mlp_sample_kernel, dim3(grid_of(N)), dim3(256), 0, st, params, d, x, target, N, act, err);
  hipLaunchKernelGGL(mlp_grad_kernel, dim3((d.P + 1 + 255) / 256), dim3(256), 0, st, d, x, act, err, N, grad, loss);
  return hipGetLastError();
}

static hipError_t launch_adam(float* params, const float* grad, float* m, float* v, int P, int step, float lr,
                                hipStream_t st) {
  if (P < 1 || step < 1) return hipErrorInvalidValue;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((P + 255) / 256), dim3(256), 0, st, params, grad, m, v, P, lr, b1, b2, eps,
                     c1, c2);
  return hipGetLastError();
}

// ---- C ABI ------------------------------------------------------------------
extern "C" int cuot_mlp_param_count(int F, int H) { return mlp_nparams(F, H); }

extern "C" size_t cuot_mlp_workspace_bytes(int F, int H, int64_t N) {
  if (mlp_nparams(F, H) < 0 || N < 1) return 0;
  return (size_t)N * (size_t)(4 * H + 2) * sizeof(float) + 256;
}

extern "C" int cuot_mlp_forward(cuot_ctx* ctx, const float* params_d, int F, int H, const float* x_d, int64_t N,
                                float* y_d, void* stream) {
  CUOT_CHECK_ARG(ctx && params_d && (N == 0 || (x_d && y_d)), "NULL argument");
  CUOT_CHECK_ARG(mlp_nparams(F, H) > 0, "F must be in [1, 32] and H in [1, 64]");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  CUOT_HIP_TRY(launch_forward(params_d, F, H, x_d, N, y_d, ctx->st(stream)));
  return CUOT_OK;
}

extern "C" int cuot_mlp_grad(cuot_ctx* ctx, const float* params_d, int F, int H, const float* x_d,
                             const float* target_d, int64_t N, void* ws_d, size_t ws_bytes, float* grad_d,
                             float* loss_d, void* stream) {
  CUOT_CHECK_ARG(ctx && params_d && x_d && target_d && ws_d && grad_d && loss_d, "NULL argument");
  CUOT_CHECK_ARG(mlp_nparams(F, H) > 0 && N >= 1, "bad dimensions");
  CUOT_CHECK_ARG(ws_bytes >= cuot_mlp_workspace_bytes(F, H, N), "workspace too small (cuot_mlp_workspace_bytes)");
  float* act = (float*)ws_d;
  float* err = act + (size_t)N * (4 * H + 1);
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  CUOT_HIP_TRY(launch_grad(params_d, F, H, x_d, target_d, N, act, err, grad_d, loss_d, ctx->st(stream)));
  return CUOT_OK;
}

extern "C" int cuot_mlp_adam(cuot_ctx* ctx, float* params_d, const float* grad_d, float* m_d, float* v_d, int P,
                             int step, float lr, void* stream) {
  CUOT_CHECK_ARG(ctx && params_d && grad_d && m_d && v_d, "NULL argument");
  CUOT_CHECK_ARG(P >= 1 && step >= 1, "P and step must be >= 1");
  CUOT_HIP_TRY(hipSetDevice(ctx->device));
  CUOT_HIP_TRY(launch_adam(params_d, grad_d, m_d, v_d, P, step, lr, ctx->st(stream)));
  return CUOT_OK;
}
