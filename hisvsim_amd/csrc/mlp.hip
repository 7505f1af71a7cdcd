// Cost-model MLP on the GPU: a 3-layer perceptron (in -> H -> H -> 1, ReLU,
// FP32) trained with mean-squared error and Adam by this file's own forward
// and backward kernels. It maps features of a part to its predicted kernel
// time; the planner uses it to choose the part width.
//
// Parameter layout (one flat FP32 array, as hsv_mlp_get_params returns it):
//   W1[in][H] | b1[H] | W2[H][H] | b2[H] | W3[H] | b3
// Forward:  a1 = relu(X W1 + b1); a2 = relu(a1 W2 + b2); y = a2 W3 + b3.
// Every reduction runs in a fixed order, so training is deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hsv.h"

void hsv_set_last_error(const std::string& msg);

namespace {

int mfail(int code, const std::string& m) {
  hsv_set_last_error(m);
  return code;
}
#define MCHK(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return mfail(HSV_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Y[n][o] = act(sum_i X[n][i] W[i][o] + b[o]); one thread per output element.
__global__ void k_linear_fwd(const float* __restrict__ X, const float* __restrict__ W, const float* __restrict__ b,
                             float* __restrict__ Y, int N, int I, int O, int relu) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * O) return;
  const int n = t / O, o = t % O;
  float acc = b[o];
  for (int i = 0; i < I; ++i) acc = fmaf(X[n * I + i], W[i * O + o], acc);
  Y[t] = relu ? fmaxf(acc, 0.0f) : acc;
}

// dX[n][i] = (sum_o dY[n][o] W[i][o]) * (A[n][i] > 0), A = the layer input's
// post-ReLU activation (mask = 1 when A is the network input).
__global__ void k_linear_bwd_in(const float* __restrict__ dY, const float* __restrict__ W,
                                const float* __restrict__ A, float* __restrict__ dX, int N, int I, int O,
                                int relu_mask) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * I) return;
  const int n = t / I, i = t % I;
  float acc = 0.0f;
  for (int o = 0; o < O; ++o) acc = fmaf(dY[n * O + o], W[i * O + o], acc);
  dX[t] = (relu_mask && !(A[t] > 0.0f)) ? 0.0f : acc;
}

// dW[i][o] = sum_n X[n][i] dY[n][o]; db[o] = sum_n dY[n][o]. One thread per
// weight (plus one row of threads for the bias), summing over n in order.
__global__ void k_linear_bwd_w(const float* __restrict__ X, const float* __restrict__ dY, float* __restrict__ dW,
                               float* __restrict__ db, int N, int I, int O) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < I * O) {
    const int i = t / O, o = t % O;
    float acc = 0.0f;
    for (int n = 0; n < N; ++n) acc = fmaf(X[n * I + i], dY[n * O + o], acc);
    dW[t] = acc;
  } else if (t < I * O + O) {
    const int o = t - I * O;
    float acc = 0.0f;
    for (int n = 0; n < N; ++n) acc += dY[n * O + o];
    db[o] = acc;
  }
}

// d = 2 (y - target) / N and the per-sample squared error
__global__ void k_mse_grad(const float* __restrict__ y, const float* __restrict__ target, float* __restrict__ d,
                           float* __restrict__ sq, int N) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const float e = y[n] - target[n];
  d[n] = 2.0f * e / (float)N;
  sq[n] = e * e;
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int P, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P) return;
  const float gi = g[t];
  m[t] = b1 * m[t] + (1.0f - b1) * gi;
  v[t] = b2 * v[t] + (1.0f - b2) * gi * gi;
  const float mh = m[t] / c1, vh = v[t] / c2;
  w[t] -= lr * mh / (sqrtf(vh) + eps);
}

inline int blocks(long long n, int bs = 256) { return (int)((n + bs - 1) / bs); }

}  // namespace

struct hsv_mlp {
  int in = 0, H = 0;
  size_t P = 0;
  float* w = nullptr;   // parameters
  float* g = nullptr;   // gradients (same layout)
  float* m = nullptr;   // Adam moments
  float* v = nullptr;
  int step = 0;
  // activations / scratch for up to cap samples
  int cap = 0;
  float *X = nullptr, *T = nullptr, *a1 = nullptr, *a2 = nullptr, *y = nullptr;
  float *d1 = nullptr, *d2 = nullptr, *d3 = nullptr, *sq = nullptr;
  hipStream_t stream = nullptr;
  size_t off_b1() const { return (size_t)in * H; }
  size_t off_W2() const { return off_b1() + H; }
  size_t off_b2() const { return off_W2() + (size_t)H * H; }
  size_t off_W3() const { return off_b2() + H; }
  size_t off_b3() const { return off_W3() + H; }
};

namespace {

void free_scratch(hsv_mlp* m) {
  float** bufs[] = {&m->X, &m->T, &m->a1, &m->a2, &m->y, &m->d1, &m->d2, &m->d3, &m->sq};
  for (float** b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  m->cap = 0;
}

int ensure_cap(hsv_mlp* m, int n) {
  if (n <= m->cap) return HSV_OK;
  free_scratch(m);
  const size_t N = (size_t)n;
  struct { float** p; size_t cnt; } want[] = {{&m->X, N * m->in}, {&m->T, N}, {&m->a1, N * m->H}, {&m->a2, N * m->H},
                                              {&m->y, N}, {&m->d1, N * m->H}, {&m->d2, N * m->H}, {&m->d3, N},
                                              {&m->sq, N}};
  for (auto& w : want)
    if (hipMalloc(reinterpret_cast<void**>(w.p), std::max<size_t>(w.cnt, 1) * sizeof(float)) != hipSuccess)
      return mfail(HSV_ERR_ALLOC, "MLP scratch allocation failed");
  m->cap = n;
  return HSV_OK;
}

void forward(hsv_mlp* m, int n, const float* Xd) {
  const float* w = m->w;
  hipLaunchKernelGGL(k_linear_fwd, blocks((long long)n * m->H), 256, 0, m->stream, Xd, w, w + m->off_b1(), m->a1,
                     n, m->in, m->H, 1);
  hipLaunchKernelGGL(k_linear_fwd, blocks((long long)n * m->H), 256, 0, m->stream, m->a1, w + m->off_W2(),
                     w + m->off_b2(), m->a2, n, m->H, m->H, 1);
  hipLaunchKernelGGL(k_linear_fwd, blocks(n), 256, 0, m->stream, m->a2, w + m->off_W3(), w + m->off_b3(), m->y, n,
                     m->H, 1, 0);
}

void backward(hsv_mlp* m, int n, const float* Xd, const float* Td) {
  const float* w = m->w;
  float* g = m->g;
  hipLaunchKernelGGL(k_mse_grad, blocks(n), 256, 0, m->stream, m->y, Td, m->d3, m->sq, n);
  // layer 3: y = a2 W3 + b3
  hipLaunchKernelGGL(k_linear_bwd_w, blocks((long long)m->H + 1), 256, 0, m->stream, m->a2, m->d3, g + m->off_W3(),
                     g + m->off_b3(), n, m->H, 1);
  hipLaunchKernelGGL(k_linear_bwd_in, blocks((long long)n * m->H), 256, 0, m->stream, m->d3, w + m->off_W3(), m->a2,
                     m->d2, n, m->H, 1, 1);
  // layer 2: a2 = relu(a1 W2 + b2)
  hipLaunchKernelGGL(k_linear_bwd_w, blocks((long long)m->H * m->H + m->H), 256, 0, m->stream, m->a1, m->d2,
                     g + m->off_W2(), g + m->off_b2(), n, m->H, m->H);
  hipLaunchKernelGGL(k_linear_bwd_in, blocks((long long)n * m->H), 256, 0, m->stream, m->d2, w + m->off_W2(), m->a1,
                     m->d1, n, m->H, m->H, 1);
  // layer 1: a1 = relu(X W1 + b1)
  hipLaunchKernelGGL(k_linear_bwd_w, blocks((long long)m->in * m->H + m->H), 256, 0, m->stream, Xd, m->d1, g,
                     g + m->off_b1(), n, m->in, m->H);
}

}  // namespace

extern "C" {

int hsv_mlp_create(int in_features, int hidden, uint64_t seed, hsv_mlp** out) {
  if (!out || in_features < 1 || in_features > 256 || hidden < 1 || hidden > 1024)
    return mfail(HSV_ERR_ARG, "bad MLP shape");
  *out = nullptr;
  auto* m = new hsv_mlp();
  m->in = in_features;
  m->H = hidden;
  m->P = (size_t)in_features * hidden + hidden + (size_t)hidden * hidden + hidden + hidden + 1;
  if (hipStreamCreate(&m->stream) != hipSuccess) { delete m; return mfail(HSV_ERR_DEVICE, "stream"); }
  float** bufs[] = {&m->w, &m->g, &m->m, &m->v};
  for (float** b : bufs)
    if (hipMalloc(reinterpret_cast<void**>(b), m->P * sizeof(float)) != hipSuccess) {
      hsv_mlp_destroy(m);
      return mfail(HSV_ERR_ALLOC, "MLP parameter allocation failed");
    }
  // He-uniform initialisation from a splitmix64 stream, zero biases
  std::vector<float> h(m->P, 0.0f);
  uint64_t s = seed ? seed : 0x9E3779B97F4A7C15ull;
  auto uni = [&]() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) / 9007199254740992.0 * 2.0 - 1.0;
  };
  const double a1 = std::sqrt(6.0 / in_features), a2 = std::sqrt(6.0 / hidden);
  for (size_t i = 0; i < (size_t)in_features * hidden; ++i) h[i] = (float)(a1 * uni());
  for (size_t i = 0; i < (size_t)hidden * hidden; ++i) h[m->off_W2() + i] = (float)(a2 * uni());
  for (int i = 0; i < hidden; ++i) h[m->off_W3() + i] = (float)(a2 * uni());
  MCHK(hipMemcpy(m->w, h.data(), m->P * sizeof(float), hipMemcpyHostToDevice));
  std::vector<float> z(m->P, 0.0f);
  MCHK(hipMemcpy(m->m, z.data(), m->P * sizeof(float), hipMemcpyHostToDevice));
  MCHK(hipMemcpy(m->v, z.data(), m->P * sizeof(float), hipMemcpyHostToDevice));
  MCHK(hipMemcpy(m->g, z.data(), m->P * sizeof(float), hipMemcpyHostToDevice));
  *out = m;
  return HSV_OK;
}

int hsv_mlp_destroy(hsv_mlp* m) {
  if (!m) return HSV_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  free_scratch(m);
  float** bufs[] = {&m->w, &m->g, &m->m, &m->v};
  for (float** b : bufs)
    if (*b) (void)hipFree(*b);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return HSV_OK;
}

int hsv_mlp_num_params(hsv_mlp* m, uint64_t* count) {
  if (!m || !count) return mfail(HSV_ERR_ARG, "null argument");
  *count = m->P;
  return HSV_OK;
}

int hsv_mlp_get_params(hsv_mlp* m, float* out) {
  if (!m || !out) return mfail(HSV_ERR_ARG, "null argument");
  MCHK(hipMemcpyAsync(out, m->w, m->P * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  return HSV_OK;
}

int hsv_mlp_set_params(hsv_mlp* m, const float* in) {
  if (!m || !in) return mfail(HSV_ERR_ARG, "null argument");
  MCHK(hipMemcpyAsync(m->w, in, m->P * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  m->step = 0;
  return HSV_OK;
}

int hsv_mlp_forward(hsv_mlp* m, const float* X, int n, float* y_out) {
  if (!m || n < 0 || (n && (!X || !y_out))) return mfail(HSV_ERR_ARG, "bad arguments");
  if (!n) return HSV_OK;
  if (int rc = ensure_cap(m, n)) return rc;
  MCHK(hipMemcpyAsync(m->X, X, sizeof(float) * n * m->in, hipMemcpyHostToDevice, m->stream));
  forward(m, n, m->X);
  MCHK(hipGetLastError());
  MCHK(hipMemcpyAsync(y_out, m->y, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  return HSV_OK;
}

// Forward + backward on (X, target); the mean squared error in *loss and,
// when grads_out is not null, the gradient of every parameter.
int hsv_mlp_gradient(hsv_mlp* m, const float* X, const float* target, int n, float* loss, float* grads_out) {
  if (!m || n <= 0 || !X || !target) return mfail(HSV_ERR_ARG, "bad arguments");
  if (int rc = ensure_cap(m, n)) return rc;
  MCHK(hipMemcpyAsync(m->X, X, sizeof(float) * n * m->in, hipMemcpyHostToDevice, m->stream));
  MCHK(hipMemcpyAsync(m->T, target, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
  forward(m, n, m->X);
  backward(m, n, m->X, m->T);
  MCHK(hipGetLastError());
  if (loss) {
    std::vector<float> sq(n);
    MCHK(hipMemcpyAsync(sq.data(), m->sq, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    double s = 0;
    for (float x : sq) s += x;
    *loss = (float)(s / n);
  }
  if (grads_out) {
    MCHK(hipMemcpyAsync(grads_out, m->g, m->P * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
  }
  return HSV_OK;
}

// Full-batch Adam for `epochs` steps over mini-batches of `batch` samples
// taken in order (the caller shuffles). Final mean squared error in *loss.
int hsv_mlp_train(hsv_mlp* m, const float* X, const float* target, int n, int epochs, int batch, float lr,
                  float* loss) {
  if (!m || n <= 0 || !X || !target || epochs < 0 || batch <= 0 || !(lr > 0))
    return mfail(HSV_ERR_ARG, "bad arguments");
  const int nb = std::min(batch, n);
  if (int rc = ensure_cap(m, nb)) return rc;
  // the whole training set goes to the device once
  float *Xall = nullptr, *Tall = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&Xall), sizeof(float) * (size_t)n * m->in) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&Tall), sizeof(float) * (size_t)n) != hipSuccess) {
    if (Xall) (void)hipFree(Xall);
    return mfail(HSV_ERR_ALLOC, "training set allocation failed");
  }
  MCHK(hipMemcpyAsync(Xall, X, sizeof(float) * (size_t)n * m->in, hipMemcpyHostToDevice, m->stream));
  MCHK(hipMemcpyAsync(Tall, target, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, m->stream));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int e = 0; e < epochs; ++e) {
    for (int s0 = 0; s0 < n; s0 += nb) {
      const int cnt = std::min(nb, n - s0);
      forward(m, cnt, Xall + (size_t)s0 * m->in);
      backward(m, cnt, Xall + (size_t)s0 * m->in, Tall + s0);
      ++m->step;
      const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
      hipLaunchKernelGGL(k_adam, blocks((long long)m->P), 256, 0, m->stream, m->w, m->g, m->m, m->v, (int)m->P, lr,
                         b1, b2, eps, c1, c2);
    }
  }
  (void)hipStreamSynchronize(m->stream);
  (void)hipFree(Xall);
  (void)hipFree(Tall);
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(m->stream));
  if (loss) {
    float l = 0;
    if (int rc = hsv_mlp_gradient(m, X, target, std::min(n, 65536), &l, nullptr)) return rc;
    *loss = l;
  }
  return HSV_OK;
}

}  // extern "C"
