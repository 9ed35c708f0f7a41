// qd_mlp.hip -- the J-interpolation network: a 3 x 128 tanh MLP in FP32 with
// hand-written forward, backward and Adam kernels.  No counterpart in the
// reference (which sums P^J(E) over propagated J only); it predicts P^J(E) for
// J that are not propagated, from (J, E, reduced centrifugal barrier).
//
// Parameters: one flat FP32 vector W1[o][i], b1, W2[o][i], b2, W3[o][i], b3,
// w4[i], b4 (width 128, n_in inputs).  Loss: mean squared error over the
// training set.  One training iteration is
//   k_mlp_grad   one work group per tile of TS samples: forward, backward, and
//                the tile's partial gradient written to its own row
//   k_mlp_adam   one thread per parameter: sum of the rows (a fixed order, so
//                the result does not depend on scheduling) and the Adam update
//   k_mlp_tick   the step counter used for Adam's bias correction
// captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qd.h"

namespace qdm {

constexpr int W = 128;        // width
constexpr int TS = 16;        // samples per work group
constexpr int THREADS = 256;

struct Layout {
  int n_in = 3;
  long long oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0, ow4 = 0, ob4 = 0, n = 0;
  Layout() = default;
  explicit Layout(int nin) : n_in(nin) {
    oW1 = 0;
    ob1 = oW1 + (long long)W * nin;
    oW2 = ob1 + W;
    ob2 = oW2 + (long long)W * W;
    oW3 = ob2 + W;
    ob3 = oW3 + (long long)W * W;
    ow4 = ob3 + W;
    ob4 = ow4 + W;
    n = ob4 + 1;
  }
};

struct GradArgs {
  const float *p;      // parameters
  const float *X;      // [n][n_in]
  const float *T;      // [n] targets
  int n, n_in;
  float *partial;      // [tiles][n_params] partial gradients
  float *partial_loss; // [tiles] sum of squared residuals
  Layout L;
};

// dense layer: out[s][o] = tanh(sum_i Wm[o][i] in[s][i] + b[o]) for the tile
__device__ __forceinline__ void dense_tanh(const float *Wm, const float *b, const float *in, int nin, float *out) {
  const int t = threadIdx.x;
  const int o = t % W, sg = t / W;  // 2 sample groups
  float acc[TS / 2];
#pragma unroll
  for (int k = 0; k < TS / 2; k++) acc[k] = b[o];
  for (int i = 0; i < nin; i++) {
    const float w = Wm[(size_t)o * nin + i];
#pragma unroll
    for (int k = 0; k < TS / 2; k++) acc[k] = fmaf(w, in[(sg + 2 * k) * nin + i], acc[k]);
  }
#pragma unroll
  for (int k = 0; k < TS / 2; k++) out[(sg + 2 * k) * W + o] = tanhf(acc[k]);
}

// gW[o][i] = sum_s d[s][o] in[s][i] (tile partial), gb[o] = sum_s d[s][o]
__device__ __forceinline__ void dense_grad(const float *d, const float *in, int nin, float *gW, float *gb) {
  for (int e = threadIdx.x; e < W * nin; e += THREADS) {
    const int o = e / nin, i = e % nin;
    float acc = 0.f;
#pragma unroll 8
    for (int s = 0; s < TS; s++) acc = fmaf(d[s * W + o], in[s * nin + i], acc);
    gW[e] = acc;
  }
  for (int o = threadIdx.x; o < W; o += THREADS) {
    float acc = 0.f;
    for (int s = 0; s < TS; s++) acc += d[s * W + o];
    gb[o] = acc;
  }
}

// dprev[s][i] = (sum_o d[s][o] Wm[o][i]) (1 - h[s][i]^2)
__device__ __forceinline__ void back_tanh(const float *d, const float *Wm, const float *h, float *dprev) {
  const int t = threadIdx.x;
  const int i = t % W, sg = t / W;
  float acc[TS / 2];
#pragma unroll
  for (int k = 0; k < TS / 2; k++) acc[k] = 0.f;
  for (int o = 0; o < W; o++) {
    const float w = Wm[(size_t)o * W + i];
#pragma unroll
    for (int k = 0; k < TS / 2; k++) acc[k] = fmaf(d[(sg + 2 * k) * W + o], w, acc[k]);
  }
#pragma unroll
  for (int k = 0; k < TS / 2; k++) {
    const float hv = h[(sg + 2 * k) * W + i];
    dprev[(sg + 2 * k) * W + i] = acc[k] * (1.f - hv * hv);
  }
}

__global__ __launch_bounds__(THREADS) void k_mlp_grad(GradArgs A) {
  __shared__ float x[TS * 4];
  __shared__ float h1[TS * W], h2[TS * W], h3[TS * W], da[TS * W], db[TS * W];
  __shared__ float dy[TS], r2[TS];
  const int t = threadIdx.x;
  const int s0 = blockIdx.x * TS;
  const int nin = A.n_in;
  const Layout &L = A.L;
  const float *p = A.p;
  for (int e = t; e < TS * nin; e += THREADS) {
    const int s = e / nin;
    x[e] = s0 + s < A.n ? A.X[(size_t)(s0 + s) * nin + e % nin] : 0.f;
  }
  __syncthreads();
  dense_tanh(p + L.oW1, p + L.ob1, x, nin, h1);
  __syncthreads();
  dense_tanh(p + L.oW2, p + L.ob2, h1, W, h2);
  __syncthreads();
  dense_tanh(p + L.oW3, p + L.ob3, h2, W, h3);
  __syncthreads();
  // output, residual, dL/dy = 2 r / n (zero for padding samples)
  if (t < TS) {
    float y = p[L.ob4];
    for (int i = 0; i < W; i++) y = fmaf(p[L.ow4 + i], h3[t * W + i], y);
    const bool in = s0 + t < A.n;
    const float r = in ? y - A.T[s0 + t] : 0.f;
    dy[t] = 2.f * r / (float)A.n;
    r2[t] = r * r;
  }
  __syncthreads();
  float *g = A.partial + (size_t)blockIdx.x * L.n;
  if (t == 0) {
    float a = 0.f;
    for (int s = 0; s < TS; s++) a += r2[s];
    A.partial_loss[blockIdx.x] = a;
  }
  // layer 4
  for (int i = t; i < W; i += THREADS) {
    float a = 0.f;
    for (int s = 0; s < TS; s++) a = fmaf(dy[s], h3[s * W + i], a);
    g[L.ow4 + i] = a;
  }
  if (t == 0) {
    float a = 0.f;
    for (int s = 0; s < TS; s++) a += dy[s];
    g[L.ob4] = a;
  }
  // d3 = dy w4 (1 - h3^2)
  for (int e = t; e < TS * W; e += THREADS) {
    const int s = e / W, o = e % W;
    const float hv = h3[e];
    da[e] = dy[s] * p[L.ow4 + o] * (1.f - hv * hv);
  }
  __syncthreads();
  dense_grad(da, h2, W, g + L.oW3, g + L.ob3);
  back_tanh(da, p + L.oW3, h2, db);  // d2
  __syncthreads();
  dense_grad(db, h1, W, g + L.oW2, g + L.ob2);
  back_tanh(db, p + L.oW2, h1, da);  // d1
  __syncthreads();
  dense_grad(da, x, nin, g + L.oW1, g + L.ob1);
}

struct AdamArgs {
  float *p, *m, *v;
  const float *partial;
  const float *partial_loss;
  int tiles;
  long long n_params;
  float lr, b1, b2, eps;
  const int *step;   // completed steps before this one
  float *loss_out;   // mean squared error of the current parameters
  int n;
};

__global__ __launch_bounds__(256) void k_mlp_adam(AdamArgs A) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q == 0) {
    float a = 0.f;
    for (int w = 0; w < A.tiles; w++) a += A.partial_loss[w];
    *A.loss_out = a / (float)A.n;
  }
  if (q >= A.n_params) return;
  float g = 0.f;
  for (int w = 0; w < A.tiles; w++) g += A.partial[(size_t)w * A.n_params + q];
  const int k = *A.step + 1;
  const float m = A.b1 * A.m[q] + (1.f - A.b1) * g;
  const float v = A.b2 * A.v[q] + (1.f - A.b2) * g * g;
  A.m[q] = m;
  A.v[q] = v;
  const float mh = m / (1.f - powf(A.b1, (float)k));
  const float vh = v / (1.f - powf(A.b2, (float)k));
  A.p[q] -= A.lr * mh / (sqrtf(vh) + A.eps);
}

__global__ void k_mlp_tick(int *step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

__global__ __launch_bounds__(THREADS) void k_mlp_forward(const float *p, const float *X, int n, int nin, Layout L,
                                                         float *Y) {
  __shared__ float x[TS * 4];
  __shared__ float h1[TS * W], h2[TS * W], h3[TS * W];
  const int t = threadIdx.x;
  const int s0 = blockIdx.x * TS;
  for (int e = t; e < TS * nin; e += THREADS) {
    const int s = e / nin;
    x[e] = s0 + s < n ? X[(size_t)(s0 + s) * nin + e % nin] : 0.f;
  }
  __syncthreads();
  dense_tanh(p + L.oW1, p + L.ob1, x, nin, h1);
  __syncthreads();
  dense_tanh(p + L.oW2, p + L.ob2, h1, W, h2);
  __syncthreads();
  dense_tanh(p + L.oW3, p + L.ob3, h2, W, h3);
  __syncthreads();
  if (t < TS && s0 + t < n) {
    float y = p[L.ob4];
    for (int i = 0; i < W; i++) y = fmaf(p[L.ow4 + i], h3[t * W + i], y);
    Y[s0 + t] = y;
  }
}

}  // namespace qdm

using namespace qdm;

struct qd_mlp {
  int dev = 0;
  hipStream_t stream = nullptr;
  Layout L{3};
  float *p = nullptr, *m = nullptr, *v = nullptr;
  float *X = nullptr, *T = nullptr, *partial = nullptr, *ploss = nullptr, *loss = nullptr;
  int *step = nullptr;
  int n = 0, tiles = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  float lr = 1e-3f;
};

// ---- staging for qd_mlp_predict_noalloc --------------------------------------
// Inputs X[n][n_in] and outputs Y[n] of one call, at most QD_MLP_NOALLOC_MAX
// samples: 5 floats per sample, 1.25 MiB.  Each GPU the library runs on holds
// its own copy until the process ends.  In host code `scratch` names a host
// stand-in, not device memory: only the address that hipGetSymbolAddress
// returns for the current device may be passed to kernels and copies.
constexpr int NOALLOC_FLOATS = QD_MLP_NOALLOC_MAX * 5;
static __device__ __align__(256) float scratch[NOALLOC_FLOATS];

namespace {
// Error text in a fixed buffer, so that reporting an error allocates nothing.
thread_local char g_mlp_err[512];
void set_err(const char *a, const char *b = nullptr) {
  snprintf(g_mlp_err, sizeof(g_mlp_err), "%s%s%s", a, b ? ": " : "", b ? b : "");
}
int mfail(int code, const std::string &msg) {
  set_err(msg.c_str());
  return code;
}
int mfail_c(int code, const char *msg, const char *detail = nullptr) {
  set_err(msg, detail);
  return code;
}

// Device address of `scratch` per device ordinal, looked up once per device
// (with that device current) by qd_mlp_create or the first no-allocation call.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[MAX_DEVICES] = {};

int scratch_for_device(int dev, float **out) {
  if (dev < 0 || dev >= MAX_DEVICES) return mfail_c(QD_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return mfail_c(QD_ERR_DEVICE, "hipSetDevice", hipGetErrorString(e));
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return mfail_c(QD_ERR_DEVICE, "hipGetSymbolAddress(scratch)", hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<float *>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return QD_OK;
}
#define QM_HIP(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) return mfail(QD_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

void drop_graph(qd_mlp *h) {
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->exec = nullptr;
  h->graph = nullptr;
}

void free_data(qd_mlp *h) {
  for (void *b : {(void *)h->X, (void *)h->T, (void *)h->partial, (void *)h->ploss})
    if (b) (void)hipFree(b);
  h->X = h->T = h->partial = h->ploss = nullptr;
}

int launch_iteration(qd_mlp *h) {
  GradArgs g;
  g.p = h->p;
  g.X = h->X;
  g.T = h->T;
  g.n = h->n;
  g.n_in = h->L.n_in;
  g.partial = h->partial;
  g.partial_loss = h->ploss;
  g.L = h->L;
  hipLaunchKernelGGL(k_mlp_grad, dim3(h->tiles), dim3(THREADS), 0, h->stream, g);
  AdamArgs a;
  a.p = h->p;
  a.m = h->m;
  a.v = h->v;
  a.partial = h->partial;
  a.partial_loss = h->ploss;
  a.tiles = h->tiles;
  a.n_params = h->L.n;
  a.lr = h->lr;
  a.b1 = 0.9f;
  a.b2 = 0.999f;
  a.eps = 1e-8f;
  a.step = h->step;
  a.loss_out = h->loss;
  a.n = h->n;
  hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)((h->L.n + 255) / 256)), dim3(256), 0, h->stream, a);
  hipLaunchKernelGGL(k_mlp_tick, dim3(1), dim3(64), 0, h->stream, h->step);
  QM_HIP(hipGetLastError());
  return QD_OK;
}
}  // namespace

extern "C" {

const char *qd_mlp_last_error(void) { return g_mlp_err; }

int qd_mlp_n_params(int32_t n_in, int64_t *n_out) {
  if (!n_out) return mfail(QD_ERR_NULL, "null output");
  if (n_in < 1 || n_in > 4) return mfail(QD_ERR_ARG, "n_in must be 1..4");
  *n_out = Layout(n_in).n;
  return QD_OK;
}

int qd_mlp_create(int32_t n_in, const float *params, int32_t device, qd_mlp **out) {
  if (!params || !out) return mfail(QD_ERR_NULL, "null argument");
  if (n_in < 1 || n_in > 4) return mfail(QD_ERR_ARG, "n_in must be 1..4");
  qd_mlp *h = new qd_mlp();
  h->dev = device;
  h->L = Layout(n_in);
  auto bad = [&](int rc) {
    qd_mlp_destroy(h);
    return rc;
  };
  if (hipSetDevice(device) != hipSuccess) return bad(mfail(QD_ERR_DEVICE, "hipSetDevice failed"));
  float *staging = nullptr;  // resolves this device's no-allocation staging address once
  if (const int rc = scratch_for_device(device, &staging)) return bad(rc);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return bad(mfail(QD_ERR_DEVICE, "stream"));
  const size_t nb = sizeof(float) * h->L.n;
  if (hipMalloc((void **)&h->p, nb) != hipSuccess || hipMalloc((void **)&h->m, nb) != hipSuccess ||
      hipMalloc((void **)&h->v, nb) != hipSuccess || hipMalloc((void **)&h->step, sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&h->loss, sizeof(float)) != hipSuccess)
    return bad(mfail(QD_ERR_NOMEM, "device memory"));
  std::vector<float> zero(h->L.n, 0.f);
  int z = 0;
  if (hipMemcpy(h->p, params, nb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->m, zero.data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->v, zero.data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->step, &z, sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    return bad(mfail(QD_ERR_DEVICE, "upload"));
  *out = h;
  return QD_OK;
}

int qd_mlp_set_data(qd_mlp *h, const float *X, const float *T, int32_t n) {
  if (!h || !X || !T) return mfail(QD_ERR_NULL, "null argument");
  if (n < 1) return mfail(QD_ERR_ARG, "no samples");
  QM_HIP(hipSetDevice(h->dev));
  drop_graph(h);
  free_data(h);
  h->n = n;
  h->tiles = (n + TS - 1) / TS;
  QM_HIP(hipMalloc((void **)&h->X, sizeof(float) * (size_t)n * h->L.n_in));
  QM_HIP(hipMalloc((void **)&h->T, sizeof(float) * (size_t)n));
  QM_HIP(hipMalloc((void **)&h->partial, sizeof(float) * (size_t)h->tiles * h->L.n));
  QM_HIP(hipMalloc((void **)&h->ploss, sizeof(float) * (size_t)h->tiles));
  QM_HIP(hipMemcpy(h->X, X, sizeof(float) * (size_t)n * h->L.n_in, hipMemcpyHostToDevice));
  QM_HIP(hipMemcpy(h->T, T, sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
  return QD_OK;
}

// loss and full gradient of the current parameters on the training set (no update)
int qd_mlp_loss_grad(qd_mlp *h, float *loss, float *grad) {
  if (!h || !loss || !grad) return mfail(QD_ERR_NULL, "null argument");
  if (!h->X) return mfail(QD_ERR_ARG, "no training data");
  QM_HIP(hipSetDevice(h->dev));
  GradArgs g;
  g.p = h->p;
  g.X = h->X;
  g.T = h->T;
  g.n = h->n;
  g.n_in = h->L.n_in;
  g.partial = h->partial;
  g.partial_loss = h->ploss;
  g.L = h->L;
  hipLaunchKernelGGL(k_mlp_grad, dim3(h->tiles), dim3(THREADS), 0, h->stream, g);
  QM_HIP(hipGetLastError());
  std::vector<float> part((size_t)h->tiles * h->L.n), pl(h->tiles);
  QM_HIP(hipMemcpyAsync(part.data(), h->partial, sizeof(float) * part.size(), hipMemcpyDeviceToHost, h->stream));
  QM_HIP(hipMemcpyAsync(pl.data(), h->ploss, sizeof(float) * pl.size(), hipMemcpyDeviceToHost, h->stream));
  QM_HIP(hipStreamSynchronize(h->stream));
  for (long long q = 0; q < h->L.n; q++) {
    float a = 0.f;
    for (int w = 0; w < h->tiles; w++) a += part[(size_t)w * h->L.n + q];
    grad[q] = a;
  }
  float a = 0.f;
  for (int w = 0; w < h->tiles; w++) a += pl[w];
  *loss = a / (float)h->n;
  return QD_OK;
}

// iters Adam steps at learning rate lr on the training set; the loop is one
// captured iteration replayed as a graph (use_graph = 1) or plain launches (0).
int qd_mlp_train(qd_mlp *h, int32_t iters, float lr, int32_t use_graph, float *last_loss) {
  if (!h) return mfail(QD_ERR_NULL, "null handle");
  if (!h->X) return mfail(QD_ERR_ARG, "no training data");
  if (iters < 0) return mfail(QD_ERR_ARG, "iters < 0");
  QM_HIP(hipSetDevice(h->dev));
  if (lr != h->lr) drop_graph(h);
  h->lr = lr;
  if (use_graph) {
    if (!h->exec) {
      QM_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      const int rc = launch_iteration(h);
      hipGraph_t gph = nullptr;
      const hipError_t e = hipStreamEndCapture(h->stream, &gph);
      if (rc != QD_OK) return rc;
      QM_HIP(e);
      h->graph = gph;
      QM_HIP(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
    }
    for (int k = 0; k < iters; k++) QM_HIP(hipGraphLaunch(h->exec, h->stream));
  } else {
    for (int k = 0; k < iters; k++) {
      const int rc = launch_iteration(h);
      if (rc != QD_OK) return rc;
    }
  }
  if (last_loss) QM_HIP(hipMemcpyAsync(last_loss, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
  QM_HIP(hipStreamSynchronize(h->stream));
  return QD_OK;
}

int qd_mlp_predict(qd_mlp *h, const float *X, int32_t n, float *Y) {
  if (!h || !X || !Y) return mfail(QD_ERR_NULL, "null argument");
  if (n < 1) return QD_OK;
  QM_HIP(hipSetDevice(h->dev));
  float *dx = nullptr, *dy = nullptr;
  QM_HIP(hipMalloc((void **)&dx, sizeof(float) * (size_t)n * h->L.n_in));
  QM_HIP(hipMalloc((void **)&dy, sizeof(float) * (size_t)n));
  QM_HIP(hipMemcpyAsync(dx, X, sizeof(float) * (size_t)n * h->L.n_in, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_mlp_forward, dim3((n + TS - 1) / TS), dim3(THREADS), 0, h->stream, h->p, dx, n, h->L.n_in,
                     h->L, dy);
  QM_HIP(hipGetLastError());
  QM_HIP(hipMemcpyAsync(Y, dy, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, h->stream));
  QM_HIP(hipStreamSynchronize(h->stream));
  QM_HIP(hipFree(dx));
  QM_HIP(hipFree(dy));
  return QD_OK;
}

// qd_mlp_predict without allocation: the inputs and outputs are staged in the
// device's `scratch` array instead of per-call device buffers.  Calls that
// overlap on one GPU share that array and race.
int qd_mlp_predict_noalloc(qd_mlp *h, const float *X, int32_t n, float *Y) {
  if (!h || !X || !Y) return mfail_c(QD_ERR_NULL, "null argument");
  if (n < 0 || n > QD_MLP_NOALLOC_MAX)
    return mfail_c(QD_ERR_ARG, "qd_mlp_predict_noalloc: n must be in [0, QD_MLP_NOALLOC_MAX = 65536]");
  if (n == 0) return QD_OK;
  float *ptr = nullptr;
  if (const int rc = scratch_for_device(h->dev, &ptr)) return rc;
  hipError_t e = hipSetDevice(h->dev);
  if (e != hipSuccess) return mfail_c(QD_ERR_DEVICE, "hipSetDevice", hipGetErrorString(e));
  const int nin = h->L.n_in;
  float *dx = ptr;                                   // [n][n_in]
  float *dy = ptr + (size_t)QD_MLP_NOALLOC_MAX * 4;  // [n]
  e = hipMemcpyAsync(dx, X, sizeof(float) * (size_t)n * nin, hipMemcpyHostToDevice, h->stream);
  if (e != hipSuccess) return mfail_c(QD_ERR_DEVICE, "staging X", hipGetErrorString(e));
  hipLaunchKernelGGL(k_mlp_forward, dim3((n + TS - 1) / TS), dim3(THREADS), 0, h->stream, h->p, dx, n, nin, h->L,
                     dy);
  e = hipGetLastError();
  if (e != hipSuccess) return mfail_c(QD_ERR_DEVICE, "k_mlp_forward", hipGetErrorString(e));
  e = hipMemcpyAsync(Y, dy, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, h->stream);
  if (e != hipSuccess) return mfail_c(QD_ERR_DEVICE, "reading Y", hipGetErrorString(e));
  e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return mfail_c(QD_ERR_DEVICE, "hipStreamSynchronize", hipGetErrorString(e));
  return QD_OK;
}

int qd_mlp_get_params(qd_mlp *h, float *params) {
  if (!h || !params) return mfail(QD_ERR_NULL, "null argument");
  QM_HIP(hipSetDevice(h->dev));
  QM_HIP(hipMemcpy(params, h->p, sizeof(float) * h->L.n, hipMemcpyDeviceToHost));
  return QD_OK;
}

int qd_mlp_destroy(qd_mlp *h) {
  if (!h) return QD_OK;
  (void)hipSetDevice(h->dev);
  drop_graph(h);
  free_data(h);
  for (void *b : {(void *)h->p, (void *)h->m, (void *)h->v, (void *)h->step, (void *)h->loss})
    if (b) (void)hipFree(b);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return QD_OK;
}

}  // extern "C"
