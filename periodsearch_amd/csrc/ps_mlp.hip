// The proposal network: a fully connected FP32 network n_in -> H -> H ->
// n_out with ReLU on the hidden layers, trained on the GPU with its own
// forward, backward and Adam kernels.  It has no counterpart in the
// reference: it proposes a starting pole and low-order shape for the LM fit
// of a trial period (the reference starts every period from the same ten
// poles and the ellipsoid shape, period_search_BOINC.cpp:819-844).
//
// Layout: row-major matrices; a batch is [batch][width].  The parameters
// live in one flat FP32 array W1 [n_in][H], b1 [H], W2 [H][H], b2 [H],
// W3 [H][n_out], b3 [n_out]; gradients and Adam moments use the same layout.
//
// Every product runs through one LDS-tiled FP32 kernel (mlp_gemm): 64x64
// output tiles, 256 threads with 4x4 outputs each, k staged 16 at a time;
// the sum over k runs in a fixed order, so results are deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ps_proposal.h"

namespace {

constexpr int kTile = 64, kKT = 16;

enum Epilogue { EPI_NONE = 0, EPI_BIAS = 1, EPI_BIAS_RELU = 2, EPI_MASK = 3 };

// C[m][n] = sum_k A(m,k) B(k,n), with A(m,k) = TA ? A[k*lda+m] : A[m*lda+k]
// and B(k,n) = TB ? B[n*ldb+k] : B[k*ldb+n]; epilogue: + bias[n], ReLU, or
// times (mask[m*ldc+n] > 0) (the ReLU derivative of the layer output).
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) mlp_gemm(const float *__restrict__ A, const float *__restrict__ B,
                                                const float *__restrict__ aux, float *__restrict__ C, int M, int N,
                                                int K, int lda, int ldb, int ldc) {
  __shared__ float As[kKT][kTile + 1];
  __shared__ float Bs[kKT][kTile + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += kKT) {
    for (int e = threadIdx.x; e < kKT * kTile; e += 256) {
      // A tile: element (kk, mm); B tile: element (kk, nn)
      int kk, mm;
      if (TA) {  // consecutive threads along m
        kk = e / kTile;
        mm = e % kTile;
      } else {  // consecutive threads along k
        mm = e / kKT;
        kk = e % kKT;
      }
      const int m = m0 + mm, k = k0 + kk;
      As[kk][mm] = (m < M && k < K) ? (TA ? A[(size_t)k * lda + m] : A[(size_t)m * lda + k]) : 0.0f;
      int kb, nn;
      if (TB) {
        nn = e / kKT;
        kb = e % kKT;
      } else {
        kb = e / kTile;
        nn = e % kTile;
      }
      const int n = n0 + nn, kq = k0 + kb;
      Bs[kb][nn] = (n < N && kq < K) ? (TB ? B[(size_t)n * ldb + kq] : B[(size_t)kq * ldb + n]) : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kKT; kk++) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int m = m0 + ty + 16 * i;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int n = n0 + tx + 16 * j;
      if (n >= N) continue;
      float v = acc[i][j];
      if (EPI == EPI_BIAS || EPI == EPI_BIAS_RELU) v += aux[n];
      if (EPI == EPI_BIAS_RELU) v = v > 0.0f ? v : 0.0f;
      if (EPI == EPI_MASK) v = aux[(size_t)m * ldc + n] > 0.0f ? v : 0.0f;
      C[(size_t)m * ldc + n] = v;
    }
  }
}

// dY = 2 w_k (Y - T) / batch; per-row loss sum_k w_k (Y - T)^2.
__global__ void mlp_loss_grad(const float *Y, const float *T, const float *w, float *dY, float *row_loss, int batch,
                              int n_out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  float s = 0.0f;
  for (int k = 0; k < n_out; k++) {
    const float d = Y[(size_t)b * n_out + k] - T[(size_t)b * n_out + k];
    s += w[k] * d * d;
    dY[(size_t)b * n_out + k] = 2.0f * w[k] * d / (float)batch;
  }
  row_loss[b] = s;
}

// db[n] = sum_b dZ[b][n], b in order.
__global__ void mlp_colsum(const float *dZ, float *db, int batch, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float s = 0.0f;
  for (int b = 0; b < batch; b++) s += dZ[(size_t)b * n + j];
  db[j] = s;
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void mlp_adam(float *p, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2,
                         float eps, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace

struct ps_mlp {
  int device = 0, n_in = 0, hid = 0, n_out = 0, max_batch = 0;
  size_t n_par = 0;
  long long step = 0;
  hipStream_t stream = nullptr;
  float *par = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
  float *x = nullptr, *h1 = nullptr, *h2 = nullptr, *y = nullptr, *t = nullptr, *w = nullptr;
  float *dy = nullptr, *d2 = nullptr, *d1 = nullptr, *rl = nullptr;
  std::vector<float> host_rl;
  std::string err;
  // offsets into the flat parameter array
  size_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
};

namespace {

#define MLP_TRY(x)                                                \
  do {                                                            \
    hipError_t err_ = (x);                                        \
    if (err_ != hipSuccess) {                                     \
      m->err = std::string(#x) + ": " + hipGetErrorString(err_);  \
      return PS_ERR_HIP;                                          \
    }                                                             \
  } while (0)

template <bool TA, bool TB, int EPI>
void gemm(ps_mlp *m, const float *A, const float *B, const float *aux, float *C, int M, int N, int K, int lda, int ldb,
          int ldc) {
  dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
  hipLaunchKernelGGL((mlp_gemm<TA, TB, EPI>), grid, dim3(256), 0, m->stream, A, B, aux, C, M, N, K, lda, ldb, ldc);
}

// Forward pass of `batch` rows already in m->x: h1, h2, y.
void forward(ps_mlp *m, int batch) {
  const float *P = m->par;
  gemm<false, false, EPI_BIAS_RELU>(m, m->x, P + m->oW1, P + m->ob1, m->h1, batch, m->hid, m->n_in, m->n_in, m->hid,
                                    m->hid);
  gemm<false, false, EPI_BIAS_RELU>(m, m->h1, P + m->oW2, P + m->ob2, m->h2, batch, m->hid, m->hid, m->hid, m->hid,
                                    m->hid);
  gemm<false, false, EPI_BIAS>(m, m->h2, P + m->oW3, P + m->ob3, m->y, batch, m->n_out, m->hid, m->hid, m->n_out,
                               m->n_out);
}

// Backward pass after forward(): loss gradient and all parameter gradients.
void backward(ps_mlp *m, int batch) {
  const float *P = m->par;
  float *G = m->grad;
  const int H = m->hid, O = m->n_out, D = m->n_in;
  hipLaunchKernelGGL(mlp_loss_grad, dim3((batch + 255) / 256), dim3(256), 0, m->stream, m->y, m->t, m->w, m->dy,
                     m->rl, batch, O);
  // layer 3: dW3 = h2^T dy, db3 = colsum dy, d2 = (dy W3^T) * [h2 > 0]
  gemm<true, false, EPI_NONE>(m, m->h2, m->dy, nullptr, G + m->oW3, H, O, batch, H, O, O);
  hipLaunchKernelGGL(mlp_colsum, dim3((O + 255) / 256), dim3(256), 0, m->stream, m->dy, G + m->ob3, batch, O);
  gemm<false, true, EPI_MASK>(m, m->dy, P + m->oW3, m->h2, m->d2, batch, H, O, O, O, H);
  // layer 2
  gemm<true, false, EPI_NONE>(m, m->h1, m->d2, nullptr, G + m->oW2, H, H, batch, H, H, H);
  hipLaunchKernelGGL(mlp_colsum, dim3((H + 255) / 256), dim3(256), 0, m->stream, m->d2, G + m->ob2, batch, H);
  gemm<false, true, EPI_MASK>(m, m->d2, P + m->oW2, m->h1, m->d1, batch, H, H, H, H, H);
  // layer 1
  gemm<true, false, EPI_NONE>(m, m->x, m->d1, nullptr, G + m->oW1, D, H, batch, D, H, H);
  hipLaunchKernelGGL(mlp_colsum, dim3((H + 255) / 256), dim3(256), 0, m->stream, m->d1, G + m->ob1, batch, H);
}

int check_batch(ps_mlp *m, int batch) {
  if (batch < 1 || batch > m->max_batch) {
    m->err = "batch outside 1..max_batch";
    return PS_ERR_ARG;
  }
  MLP_TRY(hipSetDevice(m->device));
  return PS_OK;
}

int load_xt(ps_mlp *m, const float *x, const float *t, const float *w, int batch) {
  MLP_TRY(hipMemcpyAsync(m->x, x, (size_t)batch * m->n_in * sizeof(float), hipMemcpyHostToDevice, m->stream));
  if (t) MLP_TRY(hipMemcpyAsync(m->t, t, (size_t)batch * m->n_out * sizeof(float), hipMemcpyHostToDevice, m->stream));
  if (w) MLP_TRY(hipMemcpyAsync(m->w, w, (size_t)m->n_out * sizeof(float), hipMemcpyHostToDevice, m->stream));
  return PS_OK;
}

int finish_loss(ps_mlp *m, int batch, float *loss) {
  MLP_TRY(hipMemcpyAsync(m->host_rl.data(), m->rl, (size_t)batch * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_TRY(hipStreamSynchronize(m->stream));
  if (loss) {
    double s = 0;
    for (int b = 0; b < batch; b++) s += m->host_rl[b];
    *loss = (float)(s / batch);
  }
  return PS_OK;
}

}  // namespace

extern "C" {

int ps_mlp_create(int device, int n_in, int n_hidden, int n_out, int max_batch, ps_mlp **out) {
  if (!out || n_in < 1 || n_hidden < 1 || n_out < 1 || max_batch < 1) return PS_ERR_ARG;
  *out = nullptr;
  auto *m = new ps_mlp();
  m->device = device;
  m->n_in = n_in;
  m->hid = n_hidden;
  m->n_out = n_out;
  m->max_batch = max_batch;
  size_t o = 0;
  m->oW1 = o;
  o += (size_t)n_in * n_hidden;
  m->ob1 = o;
  o += n_hidden;
  m->oW2 = o;
  o += (size_t)n_hidden * n_hidden;
  m->ob2 = o;
  o += n_hidden;
  m->oW3 = o;
  o += (size_t)n_hidden * n_out;
  m->ob3 = o;
  o += n_out;
  m->n_par = o;
  m->host_rl.resize(max_batch);
  const size_t B = max_batch;
  float **bufs[] = {&m->par, &m->grad, &m->m1, &m->m2};
  bool ok = hipSetDevice(device) == hipSuccess && hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess;
  for (float **b : bufs) ok = ok && hipMalloc(b, m->n_par * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&m->x, B * n_in * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&m->h1, B * n_hidden * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&m->h2, B * n_hidden * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&m->d1, B * n_hidden * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&m->d2, B * n_hidden * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&m->y, B * n_out * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&m->t, B * n_out * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&m->dy, B * n_out * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&m->w, n_out * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&m->rl, B * sizeof(float)) == hipSuccess;
  ok = ok && hipMemset(m->par, 0, m->n_par * sizeof(float)) == hipSuccess;
  ok = ok && hipMemset(m->m1, 0, m->n_par * sizeof(float)) == hipSuccess;
  ok = ok && hipMemset(m->m2, 0, m->n_par * sizeof(float)) == hipSuccess;
  if (!ok) {
    ps_mlp_destroy(m);
    return PS_ERR_HIP;
  }
  *out = m;
  return PS_OK;
}

int ps_mlp_destroy(ps_mlp *m) {
  if (!m) return PS_ERR_ARG;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  float *bufs[] = {m->par, m->grad, m->m1, m->m2, m->x, m->h1, m->h2, m->y, m->t, m->w, m->dy, m->d2, m->d1, m->rl};
  for (float *b : bufs)
    if (b) (void)hipFree(b);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return PS_OK;
}

const char *ps_mlp_last_error(const ps_mlp *m) { return m ? m->err.c_str() : "null handle"; }

long long ps_mlp_param_count(const ps_mlp *m) { return m ? (long long)m->n_par : -1; }

int ps_mlp_set_params(ps_mlp *m, const float *params, long long n) {
  if (!m || !params || n != (long long)m->n_par) return PS_ERR_ARG;
  MLP_TRY(hipSetDevice(m->device));
  MLP_TRY(hipMemcpyAsync(m->par, params, m->n_par * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLP_TRY(hipMemsetAsync(m->m1, 0, m->n_par * sizeof(float), m->stream));
  MLP_TRY(hipMemsetAsync(m->m2, 0, m->n_par * sizeof(float), m->stream));
  MLP_TRY(hipStreamSynchronize(m->stream));
  m->step = 0;
  return PS_OK;
}

int ps_mlp_get_params(ps_mlp *m, float *params, long long n) {
  if (!m || !params || n != (long long)m->n_par) return PS_ERR_ARG;
  MLP_TRY(hipSetDevice(m->device));
  MLP_TRY(hipMemcpyAsync(params, m->par, m->n_par * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_TRY(hipStreamSynchronize(m->stream));
  return PS_OK;
}

int ps_mlp_forward(ps_mlp *m, const float *x, int batch, float *y) {
  if (!m || !x || !y) return PS_ERR_ARG;
  int rc = check_batch(m, batch);
  if (rc) return rc;
  if ((rc = load_xt(m, x, nullptr, nullptr, batch))) return rc;
  forward(m, batch);
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipMemcpyAsync(y, m->y, (size_t)batch * m->n_out * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_TRY(hipStreamSynchronize(m->stream));
  return PS_OK;
}

int ps_mlp_gradients(ps_mlp *m, const float *x, const float *t, const float *w_out, int batch, float *grad,
                     float *loss) {
  if (!m || !x || !t || !w_out || !grad) return PS_ERR_ARG;
  int rc = check_batch(m, batch);
  if (rc) return rc;
  if ((rc = load_xt(m, x, t, w_out, batch))) return rc;
  forward(m, batch);
  backward(m, batch);
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipMemcpyAsync(grad, m->grad, m->n_par * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  return finish_loss(m, batch, loss);
}

int ps_mlp_train_step(ps_mlp *m, const float *x, const float *t, const float *w_out, int batch, float lr,
                      float *loss) {
  if (!m || !x || !t || !w_out || !(lr > 0)) return PS_ERR_ARG;
  int rc = check_batch(m, batch);
  if (rc) return rc;
  if ((rc = load_xt(m, x, t, w_out, batch))) return rc;
  forward(m, batch);
  backward(m, batch);
  m->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)m->step), c2 = 1.0f - powf(b2, (float)m->step);
  hipLaunchKernelGGL(mlp_adam, dim3((unsigned)((m->n_par + 255) / 256)), dim3(256), 0, m->stream, m->par, m->grad,
                     m->m1, m->m2, m->n_par, lr, b1, b2, 1e-8f, c1, c2);
  MLP_TRY(hipGetLastError());
  return finish_loss(m, batch, loss);
}

}  // extern "C"
