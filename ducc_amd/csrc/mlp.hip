// mlp.hip -- the parameter chooser's network and its training, as gfx950
// kernels: 6 -> 64 -> 64 -> 64 -> 3, ReLU, mean-squared error, Adam.
//
// Parameters are one float32 vector: W1 (64x6, row = output neuron), b1,
// W2 (64x64), b2, W3 (64x64), b3, W4 (3x64), b4.  Forward: one 64-thread
// workgroup per sample, layer inputs in LDS, pre-activations kept for the
// backward pass.  Backward: per-sample deltas, then one thread per weight
// sums delta*activation over the batch in sample order (deterministic).
#include <cmath>
#include <vector>

#include "dmx.h"
#include "mlp.h"

namespace dmx {

namespace {
constexpr int NI = MLP_IN, NH = MLP_HID, NO = MLP_OUT;
// parameter offsets
constexpr int oW1 = 0, ob1 = oW1 + NH * NI, oW2 = ob1 + NH,
              ob2 = oW2 + NH * NH, oW3 = ob2 + NH, ob3 = oW3 + NH * NH,
              oW4 = ob3 + NH, ob4 = oW4 + NO * NH, oEnd = ob4 + NO;
static_assert(oEnd == MLP_NPARAM, "parameter count");
// per-sample activation record: z1,z2,z3 (pre-ReLU), y
constexpr int ACT = 3 * NH + NO;

__global__ __launch_bounds__(64) void k_fwd(const float *__restrict__ P,
                                            const float *__restrict__ X,
                                            int B, float *__restrict__ act,
                                            float *__restrict__ Y) {
  const int b = blockIdx.x, o = threadIdx.x;
  if (b >= B) return;
  __shared__ float x[NI], a1[NH], a2[NH], a3[NH];
  if (o < NI) x[o] = X[b * NI + o];
  __syncthreads();
  float z = P[ob1 + o];
  for (int i = 0; i < NI; ++i) z += P[oW1 + o * NI + i] * x[i];
  if (act) act[b * ACT + o] = z;
  a1[o] = z > 0.f ? z : 0.f;
  __syncthreads();
  z = P[ob2 + o];
  for (int i = 0; i < NH; ++i) z += P[oW2 + o * NH + i] * a1[i];
  if (act) act[b * ACT + NH + o] = z;
  a2[o] = z > 0.f ? z : 0.f;
  __syncthreads();
  z = P[ob3 + o];
  for (int i = 0; i < NH; ++i) z += P[oW3 + o * NH + i] * a2[i];
  if (act) act[b * ACT + 2 * NH + o] = z;
  a3[o] = z > 0.f ? z : 0.f;
  __syncthreads();
  if (o < NO) {
    float y = P[ob4 + o];
    for (int i = 0; i < NH; ++i) y += P[oW4 + o * NH + i] * a3[i];
    if (act) act[b * ACT + 3 * NH + o] = y;
    Y[b * NO + o] = y;
  }
}

// per-sample deltas dL/dz for every layer; loss = sum_b sum_k
// wk (y_bk - t_bk)^2 / (B * NO) with output weights wk
__global__ __launch_bounds__(64) void k_bwd(const float *__restrict__ P,
                                            const float *__restrict__ act,
                                            const float *__restrict__ T,
                                            const float *__restrict__ wout,
                                            int B, float *__restrict__ dz,
                                            float *__restrict__ lossb) {
  const int b = blockIdx.x, o = threadIdx.x;
  if (b >= B) return;
  __shared__ float d4[NO], d3[NH], d2[NH];
  const float *A = act + (size_t)b * ACT;
  float *D = dz + (size_t)b * ACT;
  if (o < NO) {
    float r = A[3 * NH + o] - T[b * NO + o];
    float d = 2.f * wout[o] * r / (float)(B * NO);
    d4[o] = d;
    D[3 * NH + o] = d;
    lossb[b * NO + o] = wout[o] * r * r / (float)(B * NO);
  }
  __syncthreads();
  float s = 0.f;
  for (int k = 0; k < NO; ++k) s += P[oW4 + k * NH + o] * d4[k];
  s = A[2 * NH + o] > 0.f ? s : 0.f;
  d3[o] = s;
  D[2 * NH + o] = s;
  __syncthreads();
  s = 0.f;
  for (int k = 0; k < NH; ++k) s += P[oW3 + k * NH + o] * d3[k];
  s = A[NH + o] > 0.f ? s : 0.f;
  d2[o] = s;
  D[NH + o] = s;
  __syncthreads();
  s = 0.f;
  for (int k = 0; k < NH; ++k) s += P[oW2 + k * NH + o] * d2[k];
  s = A[o] > 0.f ? s : 0.f;
  D[o] = s;
}

// one thread per parameter: sum over samples (in order) of delta * input
__global__ void k_wgrad(const float *__restrict__ X,
                        const float *__restrict__ act,
                        const float *__restrict__ dz, int B,
                        float *__restrict__ G) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= oEnd) return;
  int layer, o, i, nin;
  bool bias;
  if (p < ob1) { layer = 0; nin = NI; o = p / NI; i = p % NI; bias = false; }
  else if (p < oW2) { layer = 0; o = p - ob1; i = 0; nin = NI; bias = true; }
  else if (p < ob2) { layer = 1; nin = NH; o = (p - oW2) / NH; i = (p - oW2) % NH; bias = false; }
  else if (p < oW3) { layer = 1; o = p - ob2; i = 0; nin = NH; bias = true; }
  else if (p < ob3) { layer = 2; nin = NH; o = (p - oW3) / NH; i = (p - oW3) % NH; bias = false; }
  else if (p < oW4) { layer = 2; o = p - ob3; i = 0; nin = NH; bias = true; }
  else if (p < ob4) { layer = 3; nin = NH; o = (p - oW4) / NH; i = (p - oW4) % NH; bias = false; }
  else { layer = 3; o = p - ob4; i = 0; nin = NH; bias = true; }
  (void)nin;
  float s = 0.f;
  for (int b = 0; b < B; ++b) {
    const float *A = act + (size_t)b * ACT;
    const float d = dz[(size_t)b * ACT + layer * NH + o];
    float in;
    if (bias) in = 1.f;
    else if (layer == 0) in = X[b * NI + i];
    else {
      const float zz = A[(layer - 1) * NH + i];
      in = zz > 0.f ? zz : 0.f;
    }
    s += d * in;
  }
  G[p] = s;
}

__global__ void k_adam(float *P, const float *G, float *m, float *v,
                       float lr, float b1, float b2, float eps, float c1,
                       float c2, float wd) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= oEnd) return;
  const float g = G[p] + wd * P[p];
  const float mm = b1 * m[p] + (1.f - b1) * g;
  const float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

__global__ void k_sum(const float *x, int n, float *out) {
  // single workgroup: deterministic tree sum
  __shared__ float sh[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) s += x[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}
}  // namespace

void mlp_forward(const float *P, const float *X, int B, float *act, float *Y,
                 hipStream_t s) {
  if (B <= 0) return;
  hipLaunchKernelGGL(k_fwd, dim3(B), dim3(64), 0, s, P, X, B, act, Y);
  DMX_HIP(hipGetLastError());
}

void mlp_grad(const float *P, const float *X, const float *T,
              const float *wout, int B, float *act, float *dz, float *lossb,
              float *G, float *loss, hipStream_t s) {
  hipLaunchKernelGGL(k_fwd, dim3(B), dim3(64), 0, s, P, X, B, act,
                     (float *)(dz + (size_t)B * ACT));  // Y scratch after dz
  hipLaunchKernelGGL(k_bwd, dim3(B), dim3(64), 0, s, P, act, T, wout, B, dz,
                     lossb);
  hipLaunchKernelGGL(k_wgrad, dim3((oEnd + 255) / 256), dim3(256), 0, s, X,
                     (const float *)act, (const float *)dz, B, G);
  if (loss)
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, (const float *)lossb,
                       B * NO, loss);
  DMX_HIP(hipGetLastError());
}

void mlp_adam_step(float *P, const float *G, float *m, float *v, float lr,
                   int t, float wd, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)t), c2 = 1.f - std::pow(b2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3((oEnd + 255) / 256), dim3(256), 0, s, P, G,
                     m, v, lr, b1, b2, 1e-8f, c1, c2, wd);
  DMX_HIP(hipGetLastError());
}

size_t mlp_act_floats(int B) { return (size_t)B * ACT; }

}  // namespace dmx
