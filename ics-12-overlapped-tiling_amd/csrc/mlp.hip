// mlp.hip — the learned tile-configuration cost model: a 28 -> 64 -> 64 -> 64
// -> 1 ReLU perceptron predicting log(seconds per grid point per time step),
// with forward, backward, Adam and inference kernels.
//
// No reference counterpart in kernel form: the reference's learned ranking is a
// WEKA IBk regressor run in a JVM (scripts/run-weka.py:74-149) over 17
// features of the driver output (scripts/extract-features.py:40-209). The
// model is small (10241 parameters), so every kernel is a plain one-thread-per-
// output-element loop; reductions run in a fixed order (no atomics), which
// makes training and inference bit-reproducible for a given seed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "ot/ot.h"
#include "ot/ot_tune.h"
#include "internal.hpp"

#pragma clang fp contract(off)

namespace ot {
namespace mlp {

constexpr int F = OT_FEATURES;
constexpr int H = OT_MLP_HIDDEN;
// flat parameter layout (row-major W[in][out])
constexpr int oW1 = 0, ob1 = oW1 + F * H, oW2 = ob1 + H, ob2 = oW2 + H * H, oW3 = ob2 + H, ob3 = oW3 + H * H,
              oW4 = ob3 + H, ob4 = oW4 + H, NP = ob4 + 1;
constexpr int kThreads = 256;

__global__ void k_gather_norm(const float* __restrict__ X, const int* __restrict__ idx, int B,
                              const float* __restrict__ mean, const float* __restrict__ istd,
                              float* __restrict__ X0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * F) return;
  const int b = t / F, f = t % F;
  const int row = idx ? idx[b] : b;
  X0[t] = (X[(size_t)row * F + f] - mean[f]) * istd[f];
}

// Z[b][j] = bias[j] + sum_i A[b][i] W[i][j]; Aout = relu(Z) (or Z)
__global__ void k_dense_fwd(const float* __restrict__ A, int in, const float* __restrict__ W,
                            const float* __restrict__ bias, int out, int B, int relu, float* __restrict__ Z,
                            float* __restrict__ Aout) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * out) return;
  const int b = t / out, j = t % out;
  float z = bias[j];
  const float* a = A + (size_t)b * in;
  for (int i = 0; i < in; ++i) z += a[i] * W[i * out + j];
  Z[t] = z;
  Aout[t] = relu ? fmaxf(z, 0.0f) : z;
}

// dOut[b] = scale * 2 (out[b] - y[row]); sq[b] = (out[b] - y[row])^2
__global__ void k_loss_grad(const float* __restrict__ out, const float* __restrict__ y, const int* __restrict__ idx,
                            int B, float scale, float* __restrict__ dOut, float* __restrict__ sq) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float diff = out[b] - y[idx ? idx[b] : b];
  dOut[b] = scale * 2.0f * diff;
  sq[b] = diff * diff;
}

// acc[0] += sum of v[0..n) in a fixed tree order (one workgroup)
__global__ void k_reduce_add(const float* __restrict__ v, int n, float* __restrict__ acc) {
  __shared__ float s[kThreads];
  float a = 0.0f;
  for (int i = threadIdx.x; i < n; i += kThreads) a += v[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[0] += s[0];
}

// dW[i][j] (+)= sum_b A[b][i] dZ[b][j]; row i == in: db[j] (+)= sum_b dZ[b][j]
__global__ void k_dense_bwd_w(const float* __restrict__ A, int in, const float* __restrict__ dZ, int out, int B,
                              int accumulate, float* __restrict__ dW, float* __restrict__ db) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (in + 1) * out) return;
  const int i = t / out, j = t % out;
  float g = 0.0f;
  if (i < in) {
    for (int b = 0; b < B; ++b) g += A[(size_t)b * in + i] * dZ[(size_t)b * out + j];
    dW[i * out + j] = accumulate ? dW[i * out + j] + g : g;
  } else {
    for (int b = 0; b < B; ++b) g += dZ[(size_t)b * out + j];
    db[j] = accumulate ? db[j] + g : g;
  }
}

// dZin[b][i] = (Zin[b][i] > 0) * sum_j dZ[b][j] W[i][j]
__global__ void k_dense_bwd_x(const float* __restrict__ dZ, int out, const float* __restrict__ W, int in,
                              const float* __restrict__ Zin, int B, float* __restrict__ dZin) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * in) return;
  const int b = t / in, i = t % in;
  float g = 0.0f;
  const float* dz = dZ + (size_t)b * out;
  for (int j = 0; j < out; ++j) g += dz[j] * W[i * out + j];
  dZin[t] = Zin[t] > 0.0f ? g : 0.0f;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
  m[i] = mi;
  v[i] = vi;
  p[i] = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

inline int nblk(int n) { return (n + kThreads - 1) / kThreads; }

}  // namespace mlp
}  // namespace ot

using namespace ot::mlp;

struct ot_model {
  float* params = nullptr;  // NP
  float* grad = nullptr;    // NP
  float* m = nullptr;       // Adam moments
  float* v = nullptr;
  float* mean = nullptr;    // F
  float* istd = nullptr;    // F
  int cap = 0;              // batch capacity of the activation buffers
  float *X0 = nullptr, *Z1 = nullptr, *A1 = nullptr, *Z2 = nullptr, *A2 = nullptr, *Z3 = nullptr,
        *A3 = nullptr, *out = nullptr, *Zo = nullptr, *dOut = nullptr, *dZ3 = nullptr, *dZ2 = nullptr,
        *dZ1 = nullptr, *sq = nullptr, *acc = nullptr;
  int64_t adam_t = 0;
};

namespace {

#define MLP_HIP(expr)                                                                                      \
  do {                                                                                                     \
    hipError_t e_ = (expr);                                                                                \
    if (e_ != hipSuccess) return ot::set_error(OT_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_));   \
  } while (0)

void free_acts(ot_model* M) {
  float** bufs[] = {&M->X0, &M->Z1, &M->A1, &M->Z2, &M->A2, &M->Z3, &M->A3, &M->out,
                    &M->Zo, &M->dOut, &M->dZ3, &M->dZ2, &M->dZ1, &M->sq};
  for (float** b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  M->cap = 0;
}

ot_status ensure_cap(ot_model* M, int B) {
  if (B <= M->cap) return OT_OK;
  free_acts(M);
  const size_t h = (size_t)B * H * sizeof(float);
  MLP_HIP(hipMalloc(&M->X0, (size_t)B * F * sizeof(float)));
  float** hb[] = {&M->Z1, &M->A1, &M->Z2, &M->A2, &M->Z3, &M->A3, &M->dZ3, &M->dZ2, &M->dZ1};
  for (float** b : hb) MLP_HIP(hipMalloc(b, h));
  float** ob[] = {&M->out, &M->Zo, &M->dOut, &M->sq};
  for (float** b : ob) MLP_HIP(hipMalloc(b, (size_t)B * sizeof(float)));
  M->cap = B;
  return OT_OK;
}

// forward pass of B rows of X (device) selected by idx (device, may be null)
void forward(ot_model* M, const float* X, const int* idx, int B) {
  const float* P = M->params;
  hipLaunchKernelGGL(k_gather_norm, dim3(nblk(B * F)), dim3(kThreads), 0, 0, X, idx, B, M->mean, M->istd, M->X0);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(B * H)), dim3(kThreads), 0, 0, M->X0, F, P + oW1, P + ob1, H, B, 1,
                     M->Z1, M->A1);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(B * H)), dim3(kThreads), 0, 0, M->A1, H, P + oW2, P + ob2, H, B, 1,
                     M->Z2, M->A2);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(B * H)), dim3(kThreads), 0, 0, M->A2, H, P + oW3, P + ob3, H, B, 1,
                     M->Z3, M->A3);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(B)), dim3(kThreads), 0, 0, M->A3, H, P + oW4, P + ob4, 1, B, 0, M->Zo,
                     M->out);
}

// gradient of scale * sum_b (out_b - y_b)^2 into M->grad (accumulate or
// overwrite); the squared errors are added to M->acc[0]
void backward(ot_model* M, const float* y, const int* idx, int B, float scale, int accumulate) {
  const float* P = M->params;
  float* G = M->grad;
  hipLaunchKernelGGL(k_loss_grad, dim3(nblk(B)), dim3(kThreads), 0, 0, M->out, y, idx, B, scale, M->dOut, M->sq);
  hipLaunchKernelGGL(k_reduce_add, dim3(1), dim3(kThreads), 0, 0, M->sq, B, M->acc);
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblk((H + 1) * 1)), dim3(kThreads), 0, 0, M->A3, H, M->dOut, 1, B,
                     accumulate, G + oW4, G + ob4);
  hipLaunchKernelGGL(k_dense_bwd_x, dim3(nblk(B * H)), dim3(kThreads), 0, 0, M->dOut, 1, P + oW4, H, M->Z3, B,
                     M->dZ3);
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblk((H + 1) * H)), dim3(kThreads), 0, 0, M->A2, H, M->dZ3, H, B,
                     accumulate, G + oW3, G + ob3);
  hipLaunchKernelGGL(k_dense_bwd_x, dim3(nblk(B * H)), dim3(kThreads), 0, 0, M->dZ3, H, P + oW3, H, M->Z2, B,
                     M->dZ2);
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblk((H + 1) * H)), dim3(kThreads), 0, 0, M->A1, H, M->dZ2, H, B,
                     accumulate, G + oW2, G + ob2);
  hipLaunchKernelGGL(k_dense_bwd_x, dim3(nblk(B * H)), dim3(kThreads), 0, 0, M->dZ2, H, P + oW2, H, M->Z1, B,
                     M->dZ1);
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblk((F + 1) * H)), dim3(kThreads), 0, 0, M->X0, F, M->dZ1, H, B,
                     accumulate, G + oW1, G + ob1);
}

struct DevSet {
  float* X = nullptr;
  float* y = nullptr;
  int* idx = nullptr;
  ~DevSet() {
    if (X) (void)hipFree(X);
    if (y) (void)hipFree(y);
    if (idx) (void)hipFree(idx);
  }
};

ot_status upload_set(DevSet& s, const float* X, const float* y, int n) {
  MLP_HIP(hipMalloc(&s.X, (size_t)n * F * sizeof(float)));
  MLP_HIP(hipMemcpy(s.X, X, (size_t)n * F * sizeof(float), hipMemcpyHostToDevice));
  if (y) {
    MLP_HIP(hipMalloc(&s.y, (size_t)n * sizeof(float)));
    MLP_HIP(hipMemcpy(s.y, y, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
  }
  return OT_OK;
}

// mean loss and (optionally) gradient over the whole device set, in chunks
ot_status full_loss_grad(ot_model* M, const DevSet& s, int n, float* loss, bool want_grad) {
  constexpr int kChunk = 4096;
  ot_status st = ensure_cap(M, std::min(n, kChunk));
  if (st != OT_OK) return st;
  MLP_HIP(hipMemset(M->acc, 0, sizeof(float)));
  for (int s0 = 0; s0 < n; s0 += kChunk) {
    const int B = std::min(kChunk, n - s0);
    forward(M, s.X + (size_t)s0 * F, nullptr, B);
    backward(M, s.y + s0, nullptr, B, 1.0f / (float)n, want_grad && s0 > 0 ? 1 : 0);
  }
  MLP_HIP(hipGetLastError());
  float acc = 0.0f;
  MLP_HIP(hipMemcpy(&acc, M->acc, sizeof(float), hipMemcpyDeviceToHost));
  if (loss) *loss = acc / (float)n;
  return OT_OK;
}

}  // namespace

extern "C" {

ot_status ot_model_create(uint64_t seed, ot_model** model) {
  if (!model) return ot::set_error(OT_EINVAL, "ot_model_create: null argument");
  ot_model* M = new ot_model();
  float** pb[] = {&M->params, &M->grad, &M->m, &M->v};
  for (float** b : pb) {
    if (hipMalloc(b, NP * sizeof(float)) != hipSuccess) {
      ot_model_destroy(M);
      return ot::set_error(OT_EHIP, "ot_model_create: device allocation failed");
    }
  }
  if (hipMalloc(&M->mean, F * sizeof(float)) != hipSuccess || hipMalloc(&M->istd, F * sizeof(float)) != hipSuccess ||
      hipMalloc(&M->acc, sizeof(float)) != hipSuccess) {
    ot_model_destroy(M);
    return ot::set_error(OT_EHIP, "ot_model_create: device allocation failed");
  }
  // He-uniform weights U(-sqrt(6/fan_in), sqrt(6/fan_in)), zero biases
  std::vector<float> p(NP, 0.0f);
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  auto fill = [&](int off, int fan_in, int count) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int i = 0; i < count; ++i) p[off + i] = (float)(a * u(rng));
  };
  fill(oW1, F, F * H);
  fill(oW2, H, H * H);
  fill(oW3, H, H * H);
  fill(oW4, H, H);
  std::vector<float> zeros(NP, 0.0f), mean(F, 0.0f), istd(F, 1.0f);
  ot_status st = ot_model_set_params(M, p.data(), NP);
  if (st == OT_OK) st = ot_model_set_normalization(M, mean.data(), istd.data());
  if (st == OT_OK && (hipMemcpy(M->m, zeros.data(), NP * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
                      hipMemcpy(M->v, zeros.data(), NP * sizeof(float), hipMemcpyHostToDevice) != hipSuccess))
    st = ot::set_error(OT_EHIP, "ot_model_create: copy failed");
  if (st != OT_OK) {
    ot_model_destroy(M);
    return st;
  }
  *model = M;
  return OT_OK;
}

void ot_model_destroy(ot_model* M) {
  if (!M) return;
  free_acts(M);
  float* bufs[] = {M->params, M->grad, M->m, M->v, M->mean, M->istd, M->acc};
  for (float* b : bufs)
    if (b) (void)hipFree(b);
  delete M;
}

ot_status ot_model_param_count(const ot_model* M, int32_t* count) {
  if (!M || !count) return ot::set_error(OT_EINVAL, "ot_model_param_count: null argument");
  *count = NP;
  return OT_OK;
}

ot_status ot_model_get_params(const ot_model* M, float* host, int32_t count) {
  if (!M || !host || count != NP) return ot::set_error(OT_EINVAL, "ot_model_get_params: need %d floats", NP);
  MLP_HIP(hipMemcpy(host, M->params, NP * sizeof(float), hipMemcpyDeviceToHost));
  return OT_OK;
}

ot_status ot_model_set_params(ot_model* M, const float* host, int32_t count) {
  if (!M || !host || count != NP) return ot::set_error(OT_EINVAL, "ot_model_set_params: need %d floats", NP);
  MLP_HIP(hipMemcpy(M->params, host, NP * sizeof(float), hipMemcpyHostToDevice));
  M->adam_t = 0;
  std::vector<float> zeros(NP, 0.0f);
  MLP_HIP(hipMemcpy(M->m, zeros.data(), NP * sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP(hipMemcpy(M->v, zeros.data(), NP * sizeof(float), hipMemcpyHostToDevice));
  return OT_OK;
}

ot_status ot_model_set_normalization(ot_model* M, const float* mean, const float* istd) {
  if (!M || !mean || !istd) return ot::set_error(OT_EINVAL, "ot_model_set_normalization: null argument");
  MLP_HIP(hipMemcpy(M->mean, mean, F * sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP(hipMemcpy(M->istd, istd, F * sizeof(float), hipMemcpyHostToDevice));
  return OT_OK;
}

ot_status ot_model_fit_normalization(ot_model* M, const float* X, int32_t n) {
  if (!M || !X || n < 1) return ot::set_error(OT_EINVAL, "ot_model_fit_normalization: bad argument");
  std::vector<float> mean(F), istd(F);
  for (int f = 0; f < F; ++f) {
    double s = 0.0, s2 = 0.0;
    for (int i = 0; i < n; ++i) s += X[(size_t)i * F + f];
    const double mu = s / n;
    for (int i = 0; i < n; ++i) {
      const double d = X[(size_t)i * F + f] - mu;
      s2 += d * d;
    }
    const double sd = std::sqrt(s2 / n);
    mean[f] = (float)mu;
    istd[f] = sd > 1e-6 ? (float)(1.0 / sd) : 1.0f;  // constant column: centred only
  }
  return ot_model_set_normalization(M, mean.data(), istd.data());
}

ot_status ot_model_loss_grad(ot_model* M, const float* X, const float* y, int32_t n, float* loss, float* grad) {
  if (!M || !X || !y || n < 1) return ot::set_error(OT_EINVAL, "ot_model_loss_grad: bad argument");
  DevSet s;
  ot_status st = upload_set(s, X, y, n);
  if (st != OT_OK) return st;
  st = full_loss_grad(M, s, n, loss, grad != nullptr);
  if (st != OT_OK) return st;
  if (grad) MLP_HIP(hipMemcpy(grad, M->grad, NP * sizeof(float), hipMemcpyDeviceToHost));
  return OT_OK;
}

ot_status ot_model_train(ot_model* M, const float* X, const float* y, int32_t n, int32_t epochs, int32_t batch,
                         float lr, uint64_t seed, float* final_loss) {
  if (!M || !X || !y || n < 1 || epochs < 0 || batch < 1 || !(lr > 0.0f))
    return ot::set_error(OT_EINVAL, "ot_model_train: bad argument");
  DevSet s;
  ot_status st = upload_set(s, X, y, n);
  if (st != OT_OK) return st;
  MLP_HIP(hipMalloc(&s.idx, (size_t)n * sizeof(int)));
  st = ensure_cap(M, std::min(batch, n));
  if (st != OT_OK) return st;
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int i = n - 1; i > 0; --i) {  // Fisher-Yates with the seeded engine
      const int j = (int)(rng() % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    MLP_HIP(hipMemcpy(s.idx, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    for (int s0 = 0; s0 < n; s0 += batch) {
      const int B = std::min(batch, n - s0);
      forward(M, s.X, s.idx + s0, B);
      backward(M, s.y, s.idx + s0, B, 1.0f / (float)B, 0);
      ++M->adam_t;
      const float bc1 = 1.0f - (float)std::pow(0.9, (double)M->adam_t);
      const float bc2 = 1.0f - (float)std::pow(0.999, (double)M->adam_t);
      hipLaunchKernelGGL(k_adam, dim3(nblk(NP)), dim3(kThreads), 0, 0, M->params, M->grad, M->m, M->v, NP, lr, b1,
                         b2, eps, bc1, bc2);
    }
    MLP_HIP(hipGetLastError());
  }
  MLP_HIP(hipDeviceSynchronize());
  if (final_loss) return full_loss_grad(M, s, n, final_loss, false);
  return OT_OK;
}

ot_status ot_model_predict(ot_model* M, const float* X, int32_t n, float* y) {
  if (!M || !X || !y || n < 1) return ot::set_error(OT_EINVAL, "ot_model_predict: bad argument");
  DevSet s;
  ot_status st = upload_set(s, X, nullptr, n);
  if (st != OT_OK) return st;
  constexpr int kChunk = 4096;
  st = ensure_cap(M, std::min(n, kChunk));
  if (st != OT_OK) return st;
  for (int s0 = 0; s0 < n; s0 += kChunk) {
    const int B = std::min(kChunk, n - s0);
    forward(M, s.X + (size_t)s0 * F, nullptr, B);
    MLP_HIP(hipMemcpy(y + s0, M->out, (size_t)B * sizeof(float), hipMemcpyDeviceToHost));
  }
  return OT_OK;
}

ot_status ot_predict(ot_model* M, const ot_params* p, float* log_time) {
  if (!M || !p || !log_time) return ot::set_error(OT_EINVAL, "ot_predict: null argument");
  float f[F];
  const ot_status st = ot_features(p, f);
  if (st != OT_OK) return st;
  return ot_model_predict(M, f, 1, log_time);
}

static const char kMagic[8] = {'O', 'T', 'M', 'L', 'P', '1', 0, 0};

ot_status ot_model_save(const ot_model* M, const char* path) {
  if (!M || !path) return ot::set_error(OT_EINVAL, "ot_model_save: null argument");
  std::vector<float> p(NP), mean(F), istd(F);
  MLP_HIP(hipMemcpy(p.data(), M->params, NP * sizeof(float), hipMemcpyDeviceToHost));
  MLP_HIP(hipMemcpy(mean.data(), M->mean, F * sizeof(float), hipMemcpyDeviceToHost));
  MLP_HIP(hipMemcpy(istd.data(), M->istd, F * sizeof(float), hipMemcpyDeviceToHost));
  FILE* f = std::fopen(path, "wb");
  if (!f) return ot::set_error(OT_EINVAL, "ot_model_save: cannot open %s", path);
  const int32_t hdr[3] = {F, H, NP};
  bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(hdr, 4, 3, f) == 3 &&
            std::fwrite(p.data(), 4, NP, f) == (size_t)NP && std::fwrite(mean.data(), 4, F, f) == (size_t)F &&
            std::fwrite(istd.data(), 4, F, f) == (size_t)F;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? OT_OK : ot::set_error(OT_EINVAL, "ot_model_save: write to %s failed", path);
}

ot_status ot_model_load(const char* path, ot_model** model) {
  if (!path || !model) return ot::set_error(OT_EINVAL, "ot_model_load: null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return ot::set_error(OT_EINVAL, "ot_model_load: cannot open %s", path);
  char magic[8];
  int32_t hdr[3];
  std::vector<float> p(NP), mean(F), istd(F);
  const bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
                  std::fread(hdr, 4, 3, f) == 3 && hdr[0] == F && hdr[1] == H && hdr[2] == NP &&
                  std::fread(p.data(), 4, NP, f) == (size_t)NP && std::fread(mean.data(), 4, F, f) == (size_t)F &&
                  std::fread(istd.data(), 4, F, f) == (size_t)F;
  std::fclose(f);
  if (!ok) return ot::set_error(OT_EINVAL, "ot_model_load: %s is not a %dx%d model file", path, F, H);
  ot_model* M = nullptr;
  ot_status st = ot_model_create(0, &M);
  if (st == OT_OK) st = ot_model_set_params(M, p.data(), NP);
  if (st == OT_OK) st = ot_model_set_normalization(M, mean.data(), istd.data());
  if (st != OT_OK) {
    ot_model_destroy(M);
    return st;
  }
  *model = M;
  return OT_OK;
}

}  // extern "C"
