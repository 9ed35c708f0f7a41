// qd_api.hip -- host side of the C ABI declared in include/qd.h.
//
// The handle owns every device buffer and the stream; nothing is global, so
// handles on different threads or devices are independent (the reference's
// CUDA path keeps module-global __constant__ state, evolutionCUDA2.cu:19-24).
//
// Per step (fused pipeline, CPU-propagator semantics, timeEvol.C:657-684):
//   k_theta_*<V prologue>       psi -> lpsi              (DVR -> FBR)
//   k_fft_lines<R>              rot/2, FFT_R, T1, IFFT_R
//   k_fft_lines<r>              FFT_r, T2, IFFT_r, rot/2
//   k_theta_*<dump epilogue>    lpsi -> psi              (FBR -> DVR)
//   k_flux_extract              psi_s, dpsi_s -> time-block slot
// and every flux_block steps k_flux_accumulate folds the block into the
// energy-resolved accumulators (CRPAux.F:4-31).  P(E) is evaluated on demand.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/qd.h"
#include "qd_kernels.hpp"

using namespace qd;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define QD_HIP(call)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return fail(QD_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));         \
  } while (0)

enum { TK_STEP = 0, TK_THETA = 1, TK_FFTR1 = 2, TK_FFTR2 = 3, TK_FLUX = 4, TK_OTHER = 5, TK_N = 6 };

struct FftCfg {
  FftPlanArg plan;
  int lpb = 1;
  int nblocks = 0;
  size_t lds = 0;
};

// bf16 round-to-nearest-even of a float, as the device conversion does
uint16_t bf16_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf/nan
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}
double bf16_value(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return (double)f;
}

bool factor_plan(int n, FftPlanArg &p) {
  p.n = n;
  p.npass = 0;
  int m = n;
  while (m > 1) {
    int f = 0;
    for (int r : {4, 2, 3, 5})
      if (m % r == 0) {
        f = r;
        break;
      }
    if (!f || p.npass >= FFT_MAX_PASS) return false;
    p.radix[p.npass++] = f;
    m /= f;
  }
  return true;
}

}  // namespace

struct qd_handle {
  // configuration
  int dev = 0;
  int prec = QD_PREC_FP64;
  int pipe = QD_PIPE_FUSED;
  int variant = QD_VARIANT_CPU;
  int tblock = 64;
  int crp_every = 0;
  int profile = 0;
  int nsplit = 3;
  int n1 = 0, n2 = 0, nth = 0, M = 0, ne = 0;
  size_t n12 = 0, npt = 0, nlp = 0, nx = 0;
  double dt = 0, dr1 = 0, dr2 = 0, mu2 = 0;
  bool has_dump = false;
  bool do_crp = false;
  int surf = 0;  // 0-based surface row
  StencilArg stencil{};
  // state
  int32_t steps = 0;          // time.steps
  bool pending_prephase = false;
  long long acc_steps = 0;    // steps folded into fai/dfai
  int slot = 0;               // filled slots of the current time block
  qd_step_cb cb = nullptr;
  void *cb_user = nullptr;
  hipStream_t stream = nullptr;
  // device buffers (FP64 path uses *64, mixed uses *32)
  double2 *psi64 = nullptr, *lpsi64 = nullptr;
  float2 *psi32 = nullptr, *lpsi32 = nullptr;
  double *V = nullptr;              // FP64 potential
  float2 *expV32 = nullptr;         // mixed: exp(-i V dt) rounded to FP32
  double *d1 = nullptr, *d2 = nullptr;
  float *d1f = nullptr, *d2f = nullptr;
  void *rotR = nullptr, *rotr = nullptr, *kinR = nullptr, *kinr = nullptr, *twR = nullptr, *twr = nullptr;
  double *Wf64 = nullptr, *Wb64 = nullptr;   // packed [Kpad][Npad]
  __bf16 *Wf16 = nullptr, *Wb16 = nullptr;   // split [NS][Npad][Kpad]
  int Wf_Npad = 0, Wb_Npad = 0, Wf_Kpad = 0, Wb_Kpad = 0;
  double2 *bpsi = nullptr, *bdpsi = nullptr, *fai = nullptr, *dfai = nullptr;
  double *energies = nullptr, *eta_sq = nullptr, *w = nullptr, *crp = nullptr, *partial = nullptr;
  int npartial = 256;
  FftCfg fR, fr;
  // timing
  hipEvent_t ev[2] = {};
  std::vector<hipEvent_t> marks;
  std::vector<int> mark_cls;
  int nmarks = 0;
  double t_ms[TK_N] = {};
  long long t_n[TK_N] = {};
  double last_step_ms = 0;
  std::vector<double> host_crp;
};

namespace {

template <typename T>
int upload(T **dst, const T *src, size_t n) {
  QD_HIP(hipMalloc((void **)dst, n * sizeof(T)));
  QD_HIP(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return QD_OK;
}

int fft_config(int n, int axis, size_t elem, int nlines, FftCfg &c) {
  if (!factor_plan(n, c.plan)) return fail(QD_ERR_ARG, "FFT length " + std::to_string(n) + " has a prime factor > 5");
  int maxb = 0;
  for (int p = 0; p < c.plan.npass; p++) maxb = std::max(maxb, n / c.plan.radix[p]);
  int lpb;
  if (axis == 0) {
    lpb = 1;
    while (lpb * 2 <= 16 && lpb * 2 * n <= 1024) lpb *= 2;
  } else {
    lpb = 16;
  }
  while (lpb > 1 && ((256 / lpb) * FFT_BPT < maxb || (size_t)lpb * n * elem > 65536)) lpb /= 2;
  if ((256 / lpb) * FFT_BPT < maxb || (size_t)lpb * n * elem > 65536)
    return fail(QD_ERR_ARG, "FFT length " + std::to_string(n) + " too large");
  c.lpb = lpb;
  c.nblocks = (nlines + lpb - 1) / lpb;
  c.lds = (size_t)lpb * n * elem;
  return QD_OK;
}

template <typename T>
int launch_fft(qd_handle *h, int axis, int flags, typename cx_of<T>::type *data) {
  FftArgs<T> a;
  a.data = data;
  const FftCfg &c = axis == 0 ? h->fR : h->fr;
  a.tw = (const typename cx_of<T>::type *)(axis == 0 ? h->twR : h->twr);
  a.kin = (const typename cx_of<T>::type *)(axis == 0 ? h->kinR : h->kinr);
  a.rotR = (const typename cx_of<T>::type *)h->rotR;
  a.rotr = (const typename cx_of<T>::type *)h->rotr;
  a.n1 = h->n1;
  a.n2 = h->n2;
  a.m = h->M;
  a.lpb = c.lpb;
  a.nlines = axis == 0 ? h->n2 * h->M : h->n1 * h->M;
  a.flags = flags;
  a.plan = c.plan;
  if (axis == 0)
    hipLaunchKernelGGL((k_fft_lines<T, 0>), dim3(c.nblocks), dim3(256), c.lds, h->stream, a);
  else
    hipLaunchKernelGGL((k_fft_lines<T, 1>), dim3(c.nblocks), dim3(256), c.lds, h->stream, a);
  QD_HIP(hipGetLastError());
  return QD_OK;
}

// theta transform: fwd (psi -> lpsi, K = nth, N = M) or bwd (lpsi -> psi, K = M, N = nth)
int launch_theta(qd_handle *h, bool fwd, int pro, int epi) {
  const int R = (int)h->n12;
  const int K = fwd ? h->nth : h->M;
  const int N = fwd ? h->M : h->nth;
  dim3 grid((R + TG_ROWS - 1) / TG_ROWS, (N + TG_COLS - 1) / TG_COLS);
  if (h->prec == QD_PREC_FP64) {
    const double2 *A = fwd ? h->psi64 : h->lpsi64;
    double2 *Cm = fwd ? h->lpsi64 : h->psi64;
    const double *B = fwd ? h->Wf64 : h->Wb64;
    const int Npad = fwd ? h->Wf_Npad : h->Wb_Npad;
#define QD_T64(P, E) \
  hipLaunchKernelGGL((k_theta_f64<P, E>), grid, dim3(256), 0, h->stream, A, Cm, B, R, K, N, Npad, h->V, h->dt, h->d1, h->d2, h->n1)
    if (pro == 0 && epi == 0) QD_T64(0, 0);
    else if (pro == 1 && epi == 0) QD_T64(1, 0);
    else if (pro == 2 && epi == 0) QD_T64(2, 0);
    else if (pro == 0 && epi == 1) QD_T64(0, 1);
    else return fail(QD_ERR_ARG, "bad theta variant");
#undef QD_T64
  } else {
    const float2 *A = fwd ? h->psi32 : h->lpsi32;
    float2 *Cm = fwd ? h->lpsi32 : h->psi32;
    const __bf16 *B = fwd ? h->Wf16 : h->Wb16;
    const int Npad = fwd ? h->Wf_Npad : h->Wb_Npad;
    const int Kpad = fwd ? h->Wf_Kpad : h->Wb_Kpad;
#define QD_T16(P, E)                                                                                     \
  do {                                                                                                   \
    if (h->nsplit == 2) {                                                                                \
      constexpr int NSPL = 2;                                                                            \
      hipLaunchKernelGGL((k_theta_bf16<NSPL, P, E>), grid, dim3(256), 0, h->stream, A, Cm, B, R, K, N, Npad, \
                         Kpad, h->expV32, h->d1f, h->d2f, h->n1);                                         \
    } else {                                                                                             \
      constexpr int NSPL = 3;                                                                            \
      hipLaunchKernelGGL((k_theta_bf16<NSPL, P, E>), grid, dim3(256), 0, h->stream, A, Cm, B, R, K, N, Npad, \
                         Kpad, h->expV32, h->d1f, h->d2f, h->n1);                                         \
    }                                                                                                    \
  } while (0)
    if (pro == 0 && epi == 0) QD_T16(0, 0);
    else if (pro == 1 && epi == 0) QD_T16(1, 0);
    else if (pro == 2 && epi == 0) QD_T16(2, 0);
    else if (pro == 0 && epi == 1) QD_T16(0, 1);
    else return fail(QD_ERR_ARG, "bad theta variant");
#undef QD_T16
  }
  QD_HIP(hipGetLastError());
  return QD_OK;
}

inline dim3 grid1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

#define QD_RC0(x)                 \
  do {                            \
    int rc0_ = (x);               \
    if (rc0_ != QD_OK) return rc0_; \
  } while (0)

// Profiling: a chain of events, one before every kernel, read back in
// batches so that the host never waits inside a step.  The segment that
// starts at event i is charged to the kernel class recorded with it.
int drain_marks(qd_handle *h) {
  const int n = h->nmarks;
  if (n < 2) return QD_OK;
  QD_HIP(hipEventSynchronize(h->marks[n - 1]));
  for (int i = 0; i + 1 < n; i++) {
    float ms = 0;
    QD_HIP(hipEventElapsedTime(&ms, h->marks[i], h->marks[i + 1]));
    const int c = h->mark_cls[i];
    if (c >= 0) {
      h->t_ms[c] += ms;
      h->t_n[c] += 1;
    }
  }
  // the newest event starts the next chain
  std::swap(h->marks[0], h->marks[n - 1]);
  h->mark_cls[0] = h->mark_cls[n - 1];
  h->nmarks = 1;
  return QD_OK;
}
int mark(qd_handle *h, int cls) {
  if (!h->profile) return QD_OK;
  QD_HIP(hipEventRecord(h->marks[h->nmarks], h->stream));
  h->mark_cls[h->nmarks] = cls;
  h->nmarks++;
  if (h->nmarks == (int)h->marks.size()) QD_RC0(drain_marks(h));
  return QD_OK;
}
int tick(qd_handle *h, int k, bool) { return mark(h, k); }
int tock(qd_handle *, int) { return QD_OK; }
#define QD_RC(x)               \
  do {                         \
    int rc_ = (x);             \
    if (rc_ != QD_OK) return rc_; \
  } while (0)

int flush_flux(qd_handle *h) {
  if (!h->do_crp || h->slot == 0) return QD_OK;
  QD_RC(tick(h, TK_FLUX, true));
  dim3 grid((unsigned)((h->nx + FX_X - 1) / FX_X), (unsigned)((h->ne + FX_E - 1) / FX_E));
  hipLaunchKernelGGL(k_flux_accumulate, grid, dim3(256), 0, h->stream, h->bpsi, h->bdpsi, h->fai, h->dfai,
                     h->energies, (int)h->nx, h->ne, h->slot, h->acc_steps, h->dt);
  QD_HIP(hipGetLastError());
  QD_RC(tock(h, TK_FLUX));
  h->acc_steps += h->slot;
  h->slot = 0;
  return QD_OK;
}

int compute_crp(qd_handle *h) {
  QD_RC(flush_flux(h));
  const double sign = h->variant == QD_VARIANT_CUDA ? -1.0 : 1.0;
  hipLaunchKernelGGL(k_crp, dim3(h->ne), dim3(256), 0, h->stream, h->fai, h->dfai, h->w, h->eta_sq, h->crp, h->n1,
                     h->nth, h->dr1 / h->mu2, sign);
  QD_HIP(hipGetLastError());
  h->host_crp.resize(h->ne);
  QD_HIP(hipMemcpyAsync(h->host_crp.data(), h->crp, sizeof(double) * h->ne, hipMemcpyDeviceToHost, h->stream));
  QD_HIP(hipStreamSynchronize(h->stream));
  return QD_OK;
}

int apply_prephase(qd_handle *h) {
  // runs on FP64 psi; the mixed path keeps an FP64 copy until the first step
  double2 *p64 = h->psi64;
  bool tmp = false;
  if (h->prec == QD_PREC_MIXED) {
    QD_HIP(hipMalloc((void **)&p64, sizeof(double2) * h->npt));
    tmp = true;
    hipLaunchKernelGGL(k_f32_to_f64, grid1(h->npt), dim3(256), 0, h->stream, h->psi32, p64, h->npt);
  }
  if (h->variant == QD_VARIANT_CUDA)
    hipLaunchKernelGGL(k_prephase_cuda, grid1(h->npt), dim3(256), 0, h->stream, p64, h->V, h->dt, h->npt);
  else
    hipLaunchKernelGGL(k_prephase_cpu, grid1(h->npt), dim3(256), 0, h->stream, p64, h->V, h->dt, h->npt);
  QD_HIP(hipGetLastError());
  if (tmp) {
    hipLaunchKernelGGL(k_f64_to_f32, grid1(h->npt), dim3(256), 0, h->stream, p64, h->psi32, h->npt);
    QD_HIP(hipStreamSynchronize(h->stream));
    QD_HIP(hipFree(p64));
  }
  return QD_OK;
}

template <typename T>
int step_impl(qd_handle *h) {
  typedef typename cx_of<T>::type C;
  C *psi = (C *)(h->prec == QD_PREC_FP64 ? (void *)h->psi64 : (void *)h->psi32);
  C *lpsi = (C *)(h->prec == QD_PREC_FP64 ? (void *)h->lpsi64 : (void *)h->lpsi32);
  const bool cuda = h->variant == QD_VARIANT_CUDA;
  const bool dump = h->has_dump;
  if (h->pipe == QD_PIPE_FUSED) {
    // CUDA propagator: dump at the start of the step (evolutionCUDA2.cu:503),
    // fused with the potential phase; CPU propagator: dump after the step.
    QD_RC(tick(h, TK_THETA, true));
    QD_RC(launch_theta(h, true, cuda && dump ? 2 : 1, 0));
    QD_RC(tock(h, TK_THETA));
    QD_RC(tick(h, TK_FFTR1, true));
    QD_RC(launch_fft<T>(h, 0, FFT_PRE_ROT | FFT_FWD | FFT_KIN | FFT_INV, lpsi));
    QD_RC(tock(h, TK_FFTR1));
    QD_RC(tick(h, TK_FFTR2, true));
    QD_RC(launch_fft<T>(h, 1, FFT_FWD | FFT_KIN | FFT_INV | FFT_POST_ROT, lpsi));
    QD_RC(tock(h, TK_FFTR2));
    QD_RC(tick(h, TK_THETA, true));
    QD_RC(launch_theta(h, false, 0, !cuda && dump ? 1 : 0));
    QD_RC(tock(h, TK_THETA));
  } else {
    // one kernel per reference operator (timeEvol.C:619-647, dump_psi :713-735)
    QD_RC(tick(h, TK_OTHER, true));
    if (cuda && dump)
      hipLaunchKernelGGL(k_dump<T>, grid1(h->npt), dim3(256), 0, h->stream, psi, h->d1, h->d2, h->n1, h->n2, h->npt);
    hipLaunchKernelGGL(k_vphase<T>, grid1(h->npt), dim3(256), 0, h->stream, psi,
                       h->prec == QD_PREC_FP64 ? h->V : nullptr, (const C *)h->expV32, h->dt, h->npt);
    QD_RC(tock(h, TK_OTHER));
    QD_RC(tick(h, TK_THETA, true));
    QD_RC(launch_theta(h, true, 0, 0));
    QD_RC(tock(h, TK_THETA));
    QD_RC(tick(h, TK_OTHER, true));
    hipLaunchKernelGGL(k_rot<T>, grid1(h->nlp), dim3(256), 0, h->stream, lpsi, (const C *)h->rotR, (const C *)h->rotr,
                       h->n1, h->n2, h->nlp);
    QD_RC(tock(h, TK_OTHER));
    QD_RC(tick(h, TK_FFTR1, true));
    QD_RC(launch_fft<T>(h, 0, FFT_FWD, lpsi));
    QD_RC(tock(h, TK_FFTR1));
    QD_RC(tick(h, TK_FFTR2, true));
    QD_RC(launch_fft<T>(h, 1, FFT_FWD, lpsi));
    QD_RC(tock(h, TK_FFTR2));
    QD_RC(tick(h, TK_OTHER, true));
    hipLaunchKernelGGL(k_kin2d<T>, grid1(h->nlp), dim3(256), 0, h->stream, lpsi, (const C *)h->kinR,
                       (const C *)h->kinr, h->n1, h->n2, h->nlp);
    QD_RC(tock(h, TK_OTHER));
    QD_RC(tick(h, TK_FFTR2, true));
    QD_RC(launch_fft<T>(h, 1, FFT_INV, lpsi));
    QD_RC(tock(h, TK_FFTR2));
    QD_RC(tick(h, TK_FFTR1, true));
    QD_RC(launch_fft<T>(h, 0, FFT_INV, lpsi));
    QD_RC(tock(h, TK_FFTR1));
    QD_RC(tick(h, TK_OTHER, true));
    hipLaunchKernelGGL(k_rot<T>, grid1(h->nlp), dim3(256), 0, h->stream, lpsi, (const C *)h->rotR, (const C *)h->rotr,
                       h->n1, h->n2, h->nlp);
    QD_RC(tock(h, TK_OTHER));
    QD_RC(tick(h, TK_THETA, true));
    QD_RC(launch_theta(h, false, 0, 0));
    QD_RC(tock(h, TK_THETA));
    if (!cuda && dump) {
      QD_RC(tick(h, TK_OTHER, true));
      hipLaunchKernelGGL(k_dump<T>, grid1(h->npt), dim3(256), 0, h->stream, psi, h->d1, h->d2, h->n1, h->n2, h->npt);
      QD_RC(tock(h, TK_OTHER));
    }
  }
  QD_HIP(hipGetLastError());
  if (h->do_crp) {
    QD_RC(tick(h, TK_FLUX, true));
    hipLaunchKernelGGL(k_flux_extract<T>, grid1(h->nx), dim3(256), 0, h->stream, (const C *)psi, h->bpsi, h->bdpsi,
                       h->n1, h->n2, h->nth, h->surf, h->dr2, h->stencil, h->slot);
    QD_HIP(hipGetLastError());
    QD_RC(tock(h, TK_FLUX));
    h->slot++;
    if (h->slot == h->tblock) QD_RC(flush_flux(h));
  }
  return QD_OK;
}

int build_tables(qd_handle *h, const qd_problem *p) {
  const int n1 = h->n1, n2 = h->n2, M = h->M, nth = h->nth;
  const bool cuda = h->variant == QD_VARIANT_CUDA;
  const double dt = h->dt;
  // radial tables (MatlabStructures.C:17-36, fftwinterface.C:110-125)
  auto momentum = [](int nm, double xl, std::vector<double> &pm) {
    pm.resize(nm);
    const double two_Pi_xl = 2.0 * M_PI / xl;
    const int half = nm / 2;
    double p1 = 0;
    for (int i = 0; i <= half; i++) {
      pm[i] = p1;
      p1 += two_Pi_xl;
    }
    p1 = two_Pi_xl * (-half + 1);
    for (int i = half + 1; i < nm; i++) {
      pm[i] = p1;
      p1 += two_Pi_xl;
    }
  };
  std::vector<double> p1v, p2v;
  momentum(n1, n1 * p->r1.dr, p1v);
  momentum(n2, n2 * p->r2.dr, p2v);
  std::vector<double> I1(n1), I2(n2);
  for (int i = 0; i < n1; i++) {
    const double r = cuda ? p->r1.r[0] + i * p->r1.dr : p->r1.r[i];
    I1[i] = 1.0 / ((p->r1.mass + p->r1.mass) * r * r);
  }
  for (int j = 0; j < n2; j++) {
    const double r = cuda ? p->r2.r[0] + j * p->r2.dr : p->r2.r[j];
    I2[j] = 1.0 / ((p->r2.mass + p->r2.mass) * r * r);
  }
  // complex tables in FP64
  std::vector<double> rotR(2 * (size_t)M * n1), rotr(2 * (size_t)M * n2), kinR(2 * n1), kinr(2 * n2), twR(2 * n1),
      twr(2 * n2);
  for (int l = 0; l < M; l++) {
    for (int i = 0; i < n1; i++) {
      const double a = (-dt / 2) * (l * (l + 1) * I1[i]);
      rotR[2 * ((size_t)l * n1 + i)] = cos(a);
      rotR[2 * ((size_t)l * n1 + i) + 1] = sin(a);
    }
    for (int j = 0; j < n2; j++) {
      const double a = (-dt / 2) * (l * (l + 1) * I2[j]);
      rotr[2 * ((size_t)l * n2 + j)] = cos(a);
      rotr[2 * ((size_t)l * n2 + j) + 1] = sin(a);
    }
  }
  for (int i = 0; i < n1; i++) {
    const double T = p1v[i] * p1v[i] / (2 * p->r1.mass);
    kinR[2 * i] = cos(-dt * T) / n1;
    kinR[2 * i + 1] = sin(-dt * T) / n1;
    const long double ang = -2.0L * 3.141592653589793238462643383279502884L * i / n1;
    twR[2 * i] = (double)cosl(ang);
    twR[2 * i + 1] = (double)sinl(ang);
  }
  for (int j = 0; j < n2; j++) {
    const double T = p2v[j] * p2v[j] / (2 * p->r2.mass);
    kinr[2 * j] = cos(-dt * T) / n2;
    kinr[2 * j + 1] = sin(-dt * T) / n2;
    const long double ang = -2.0L * 3.141592653589793238462643383279502884L * j / n2;
    twr[2 * j] = (double)cosl(ang);
    twr[2 * j + 1] = (double)sinl(ang);
  }
  auto put = [&](void **dst, const std::vector<double> &v) -> int {
    if (h->prec == QD_PREC_FP64) {
      double *d = nullptr;
      QD_RC(upload(&d, v.data(), v.size()));
      *dst = d;
    } else {
      std::vector<float> f(v.begin(), v.end());
      float *d = nullptr;
      QD_RC(upload(&d, f.data(), f.size()));
      *dst = d;
    }
    return QD_OK;
  };
  QD_RC(put(&h->rotR, rotR));
  QD_RC(put(&h->rotr, rotr));
  QD_RC(put(&h->kinR, kinR));
  QD_RC(put(&h->kinr, kinr));
  QD_RC(put(&h->twR, twR));
  QD_RC(put(&h->twr, twr));

  // theta matrices (setup_weight_legendre, timeEvol.C:737-760; legendre, MatlabStructures.C:49)
  const double *leg = p->theta.legendre;
  const double *w = p->theta.w;
  std::vector<double> Wf((size_t)nth * M), Wb((size_t)M * nth);  // Wf[k][l], Wb[l][k]
  for (int l = 0; l < M; l++) {
    const double f = l + 0.5;
    for (int k = 0; k < nth; k++) {
      Wf[(size_t)k * M + l] = f * w[k] * leg[l + (size_t)M * k];
      Wb[(size_t)l * nth + k] = leg[l + (size_t)M * k];
    }
  }
  auto pack64 = [&](const std::vector<double> &B, int K, int N, double **dst, int &Npad) -> int {
    Npad = (N + TG_COLS - 1) / TG_COLS * TG_COLS;
    const int Kpad = (K + TG_KC64 - 1) / TG_KC64 * TG_KC64;
    std::vector<double> P((size_t)Kpad * Npad, 0.0);
    for (int k = 0; k < K; k++)
      for (int n = 0; n < N; n++) P[(size_t)k * Npad + n] = B[(size_t)k * N + n];
    return upload(dst, P.data(), P.size());
  };
  auto pack16 = [&](const std::vector<double> &B, int K, int N, __bf16 **dst, int &Npad, int &Kpad) -> int {
    Npad = (N + TG_COLS - 1) / TG_COLS * TG_COLS;
    Kpad = (K + TG_KC16 - 1) / TG_KC16 * TG_KC16;
    const int NS = 3;  // pieces stored; a 2-piece kernel reads the first two
    std::vector<uint16_t> P((size_t)NS * Npad * Kpad, 0);
    for (int k = 0; k < K; k++)
      for (int n = 0; n < N; n++) {
        double r = B[(size_t)k * N + n];
        for (int s = 0; s < NS; s++) {
          const uint16_t b = bf16_bits((float)r);
          P[((size_t)s * Npad + n) * Kpad + k] = b;
          r -= bf16_value(b);
        }
      }
    QD_HIP(hipMalloc((void **)dst, P.size() * 2));
    QD_HIP(hipMemcpy(*dst, P.data(), P.size() * 2, hipMemcpyHostToDevice));
    return QD_OK;
  };
  if (h->prec == QD_PREC_FP64) {
    QD_RC(pack64(Wf, nth, M, &h->Wf64, h->Wf_Npad));
    QD_RC(pack64(Wb, M, nth, &h->Wb64, h->Wb_Npad));
  } else {
    QD_RC(pack16(Wf, nth, M, &h->Wf16, h->Wf_Npad, h->Wf_Kpad));
    QD_RC(pack16(Wb, M, nth, &h->Wb16, h->Wb_Npad, h->Wb_Kpad));
  }
  return QD_OK;
}

int validate(const qd_problem *p) {
  if (!p) return fail(QD_ERR_NULL, "problem is NULL");
  if (!p->pot || !p->psi || !p->r1.r || !p->r2.r || !p->theta.x || !p->theta.w || !p->theta.legendre)
    return fail(QD_ERR_NULL, "a required array (pot, psi, r1.r, r2.r, theta.x/w/legendre) is NULL");
  if (p->r1.n <= 0 || p->r2.n <= 0 || p->theta.n <= 0 || p->theta.m < 0)
    return fail(QD_ERR_ARG, "grid sizes must be positive");
  if (p->r1.n % 2 || p->r2.n % 2) return fail(QD_ERR_ODD_N, "r1.n and r2.n must be even (FFT momentum grid)");
  if (p->time.total_steps < 0) return fail(QD_ERR_ARG, "total_steps < 0");
  if (!(p->r1.mass > 0) || !(p->r2.mass > 0) || !(p->r1.dr > 0) || !(p->r2.dr > 0))
    return fail(QD_ERR_ARG, "masses and spacings must be positive");
  return QD_OK;
}

void release(qd_handle *h) {
  if (!h) return;
  void *bufs[] = {h->psi64, h->lpsi64, h->psi32, h->lpsi32, h->V, h->expV32, h->d1, h->d2, h->d1f, h->d2f,
                  h->rotR, h->rotr, h->kinR, h->kinr, h->twR, h->twr, h->Wf64, h->Wb64, h->Wf16, h->Wb16,
                  h->bpsi, h->bdpsi, h->fai, h->dfai, h->energies, h->eta_sq, h->w, h->crp, h->partial};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (auto &e : h->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : h->marks)
    if (e) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
}

int create_impl(const qd_problem *p, const qd_options *o, qd_handle *h) {
  QD_RC(validate(p));
  qd_options opt{};
  if (o) opt = *o;
  if (opt.precision != QD_PREC_FP64 && opt.precision != QD_PREC_MIXED) return fail(QD_ERR_ARG, "bad precision");
  if (opt.pipeline != QD_PIPE_FUSED && opt.pipeline != QD_PIPE_PLAIN) return fail(QD_ERR_ARG, "bad pipeline");
  if (opt.variant != QD_VARIANT_CPU && opt.variant != QD_VARIANT_CUDA) return fail(QD_ERR_ARG, "bad variant");
  h->dev = opt.device;
  h->prec = opt.precision;
  h->pipe = opt.pipeline;
  h->variant = opt.variant;
  h->tblock = opt.flux_block > 0 ? opt.flux_block : 64;
  h->crp_every = opt.crp_every;
  h->profile = opt.profile;
  h->nsplit = opt.theta_split == 2 ? 2 : 3;
  if (opt.theta_split != 0 && opt.theta_split != 2 && opt.theta_split != 3)
    return fail(QD_ERR_ARG, "theta_split must be 0, 2 or 3");
  h->n1 = p->r1.n;
  h->n2 = p->r2.n;
  h->nth = p->theta.n;
  h->M = p->theta.m + 1;
  h->n12 = (size_t)h->n1 * h->n2;
  h->npt = h->n12 * h->nth;
  h->nlp = h->n12 * h->M;
  h->nx = (size_t)h->n1 * h->nth;
  h->dt = p->time.time_step;
  h->dr1 = p->r1.dr;
  h->dr2 = p->r2.dr;
  h->mu2 = p->r2.mass;
  h->steps = p->time.steps;
  h->pending_prephase = p->time.steps == 0;
  h->has_dump = p->dump1 && p->dump2;
  h->do_crp = p->crp.calculate_CRP != 0;
  h->ne = h->do_crp ? p->crp.n_energies : 0;
  if (h->n12 > (size_t)INT32_MAX / 2) return fail(QD_ERR_ARG, "grid too large");

  const size_t elem = h->prec == QD_PREC_FP64 ? sizeof(double2) : sizeof(float2);
  QD_RC(fft_config(h->n1, 0, elem, h->n2 * h->M, h->fR));
  QD_RC(fft_config(h->n2, 1, elem, h->n1 * h->M, h->fr));

  if (h->do_crp) {
    if (!p->crp.energies || !p->crp.eta_sq || !p->crp.CRP) return fail(QD_ERR_NULL, "CRP arrays are NULL");
    if (h->ne <= 0) return fail(QD_ERR_ARG, "n_energies must be positive");
    const int np = p->crp.n_gradient_points;
    if (np < 3 || np > 61 || np % 2 == 0) return fail(QD_ERR_STENCIL, "NPoints error");
    StencilArg &st = h->stencil;
    st.h = np / 2;
    static const double t3[] = {1}, t5[] = {8, -1}, t7[] = {45, -9, 1}, t9[] = {672, -168, 32, -3},
                        t11[] = {2100, -600, 150, -25, 2};
    const double *tab = np == 3 ? t3 : np == 5 ? t5 : np == 7 ? t7 : np == 9 ? t9 : np == 11 ? t11 : nullptr;
    const double den = np == 3 ? 2 : np == 5 ? 12 : np == 7 ? 60 : np == 9 ? 840 : np == 11 ? 2520 : 1;
    if (tab) {
      for (int j = 0; j < st.h; j++) st.num[j] = tab[j];
      st.den = den;
    } else {  // c_j = (-1)^(j+1) (h!)^2 / (j (h-j)! (h+j)!)  (gradients2.cu:18-35)
      for (int j = 1; j <= st.h; j++) {
        long double c = 1.0L;
        for (int i = 1; i <= j; i++) c *= (long double)(st.h - j + i) / (long double)(st.h + i);
        c /= j;
        st.num[j - 1] = (double)((j % 2) ? c : -c);
      }
      st.den = 1;
    }
    h->surf = p->crp.n_dividing_surface - (h->variant == QD_VARIANT_CUDA ? 0 : 1);
    if (h->surf - st.h < 0 || h->surf + st.h >= h->n2) return fail(QD_ERR_SURFACE, "surface stencil outside r grid");
  }

  // ---- everything above is host-side validation; device work starts here ----
  QD_HIP(hipSetDevice(h->dev));
  QD_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  for (auto &e : h->ev) QD_HIP(hipEventCreate(&e));
  if (h->profile) {
    h->marks.assign(512, nullptr);
    h->mark_cls.assign(512, -1);
    for (auto &e : h->marks) QD_HIP(hipEventCreate(&e));
  }

  const double *psi_in = p->psi;
  if (h->prec == QD_PREC_FP64) {
    QD_RC(upload(&h->psi64, reinterpret_cast<const double2 *>(psi_in), h->npt));
    QD_HIP(hipMalloc((void **)&h->lpsi64, sizeof(double2) * h->nlp));
  } else {
    double2 *tmp = nullptr;
    QD_RC(upload(&tmp, reinterpret_cast<const double2 *>(psi_in), h->npt));
    QD_HIP(hipMalloc((void **)&h->psi32, sizeof(float2) * h->npt));
    QD_HIP(hipMalloc((void **)&h->lpsi32, sizeof(float2) * h->nlp));
    hipLaunchKernelGGL(k_f64_to_f32, grid1(h->npt), dim3(256), 0, h->stream, tmp, h->psi32, h->npt);
    QD_HIP(hipStreamSynchronize(h->stream));
    QD_HIP(hipFree(tmp));
  }
  QD_RC(upload(&h->V, p->pot, h->npt));
  if (h->prec == QD_PREC_MIXED) {
    std::vector<float> e(2 * h->npt);
    for (size_t x = 0; x < h->npt; x++) {
      const double a = -h->dt * p->pot[x];
      e[2 * x] = (float)cos(a);
      e[2 * x + 1] = (float)sin(a);
    }
    float *d = nullptr;
    QD_RC(upload(&d, e.data(), e.size()));
    h->expV32 = reinterpret_cast<float2 *>(d);
  }
  if (h->has_dump) {
    QD_RC(upload(&h->d1, p->dump1, h->n1));
    QD_RC(upload(&h->d2, p->dump2, h->n2));
    std::vector<float> f1(p->dump1, p->dump1 + h->n1), f2(p->dump2, p->dump2 + h->n2);
    QD_RC(upload(&h->d1f, f1.data(), f1.size()));
    QD_RC(upload(&h->d2f, f2.data(), f2.size()));
  }
  QD_RC(upload(&h->w, p->theta.w, h->nth));
  QD_RC(build_tables(h, p));
  if (h->do_crp) {
    QD_RC(upload(&h->energies, p->crp.energies, h->ne));
    QD_RC(upload(&h->eta_sq, p->crp.eta_sq, h->ne));
    QD_HIP(hipMalloc((void **)&h->crp, sizeof(double) * h->ne));
    QD_HIP(hipMalloc((void **)&h->bpsi, sizeof(double2) * h->nx * h->tblock));
    QD_HIP(hipMalloc((void **)&h->bdpsi, sizeof(double2) * h->nx * h->tblock));
    QD_HIP(hipMalloc((void **)&h->fai, sizeof(double2) * h->nx * h->ne));
    QD_HIP(hipMalloc((void **)&h->dfai, sizeof(double2) * h->nx * h->ne));
    QD_HIP(hipMemsetAsync(h->fai, 0, sizeof(double2) * h->nx * h->ne, h->stream));
    QD_HIP(hipMemsetAsync(h->dfai, 0, sizeof(double2) * h->nx * h->ne, h->stream));
  }
  QD_HIP(hipMalloc((void **)&h->partial, sizeof(double) * h->npartial));
  QD_HIP(hipStreamSynchronize(h->stream));
  return QD_OK;
}

int propagate_impl(qd_handle *h, int32_t nsteps) {
  QD_HIP(hipSetDevice(h->dev));
  if (nsteps < 0) return fail(QD_ERR_ARG, "n_steps < 0");
  if (nsteps == 0) return QD_OK;
  for (int k = 0; k < TK_N; k++) {
    h->t_ms[k] = 0;
    h->t_n[k] = 0;
  }
  if (h->pending_prephase) {
    QD_RC(apply_prephase(h));
    h->pending_prephase = false;
  }
  hipEvent_t e0 = h->ev[0], e1 = h->ev[1];
  h->nmarks = 0;
  QD_HIP(hipEventRecord(e0, h->stream));
  for (int s = 0; s < nsteps; s++) {
    if (h->prec == QD_PREC_FP64) QD_RC(step_impl<double>(h));
    else QD_RC(step_impl<float>(h));
    h->steps++;
    const bool every = h->crp_every > 0 && (h->steps % h->crp_every == 0);
    if (every && (h->cb || h->do_crp)) {
      const double *c = nullptr;
      if (h->do_crp) {
        QD_RC(compute_crp(h));
        c = h->host_crp.data();
      }
      if (h->cb && h->cb(h->steps, c, h->cb_user)) break;
    }
  }
  if (h->profile) {
    QD_RC(mark(h, -1));
    QD_RC(drain_marks(h));
    h->nmarks = 0;
  }
  QD_HIP(hipEventRecord(e1, h->stream));
  QD_HIP(hipEventSynchronize(e1));
  float ms = 0;
  QD_HIP(hipEventElapsedTime(&ms, e0, e1));
  h->last_step_ms = ms / nsteps;
  return QD_OK;
}

int get_psi_impl(qd_handle *h, double *out) {
  QD_HIP(hipSetDevice(h->dev));
  if (h->prec == QD_PREC_FP64) {
    QD_HIP(hipMemcpyAsync(out, h->psi64, sizeof(double2) * h->npt, hipMemcpyDeviceToHost, h->stream));
  } else {
    double2 *tmp = nullptr;
    QD_HIP(hipMalloc((void **)&tmp, sizeof(double2) * h->npt));
    hipLaunchKernelGGL(k_f32_to_f64, grid1(h->npt), dim3(256), 0, h->stream, h->psi32, tmp, h->npt);
    QD_HIP(hipMemcpyAsync(out, tmp, sizeof(double2) * h->npt, hipMemcpyDeviceToHost, h->stream));
    QD_HIP(hipStreamSynchronize(h->stream));
    QD_HIP(hipFree(tmp));
  }
  QD_HIP(hipStreamSynchronize(h->stream));
  return QD_OK;
}

int norm_impl(qd_handle *h, double *out) {
  QD_HIP(hipSetDevice(h->dev));
  if (h->prec == QD_PREC_FP64)
    hipLaunchKernelGGL(k_norm_partial<double>, dim3(h->npartial), dim3(256), 0, h->stream, h->psi64, h->w, h->n12,
                       h->npt, h->partial);
  else
    hipLaunchKernelGGL(k_norm_partial<float>, dim3(h->npartial), dim3(256), 0, h->stream, h->psi32, h->w, h->n12,
                       h->npt, h->partial);
  QD_HIP(hipGetLastError());
  std::vector<double> part(h->npartial);
  QD_HIP(hipMemcpyAsync(part.data(), h->partial, sizeof(double) * h->npartial, hipMemcpyDeviceToHost, h->stream));
  QD_HIP(hipStreamSynchronize(h->stream));
  double s = 0;
  for (double v : part) s += v;
  *out = s * h->dr1 * h->dr2;
  return QD_OK;
}

}  // namespace

extern "C" {

const char *qd_last_error(void) { return g_last_error.c_str(); }

int qd_create(const qd_problem *problem, const qd_options *options, qd_handle **out) {
  if (!out) return fail(QD_ERR_NULL, "out is NULL");
  *out = nullptr;
  qd_handle *h = new (std::nothrow) qd_handle;
  if (!h) return fail(QD_ERR_NOMEM, "out of host memory");
  const int rc = create_impl(problem, options, h);
  if (rc != QD_OK) {
    release(h);
    delete h;
    return rc;
  }
  *out = h;
  return QD_OK;
}

int qd_propagate(qd_handle *h, int32_t n_steps) {
  if (!h) return fail(QD_ERR_NULL, "handle is NULL");
  return propagate_impl(h, n_steps);
}

int qd_set_callback(qd_handle *h, qd_step_cb cb, void *user) {
  if (!h) return fail(QD_ERR_NULL, "handle is NULL");
  h->cb = cb;
  h->cb_user = user;
  return QD_OK;
}

int qd_get_psi(qd_handle *h, double *psi_out) {
  if (!h || !psi_out) return fail(QD_ERR_NULL, "NULL argument");
  return get_psi_impl(h, psi_out);
}

int qd_get_crp(qd_handle *h, double *crp_out) {
  if (!h || !crp_out) return fail(QD_ERR_NULL, "NULL argument");
  if (!h->do_crp) return fail(QD_ERR_ARG, "calculate_CRP was 0");
  QD_HIP(hipSetDevice(h->dev));
  QD_RC(compute_crp(h));
  memcpy(crp_out, h->host_crp.data(), sizeof(double) * h->ne);
  return QD_OK;
}

int qd_get_norm(qd_handle *h, double *norm_out) {
  if (!h || !norm_out) return fail(QD_ERR_NULL, "NULL argument");
  return norm_impl(h, norm_out);
}

int qd_get_steps(qd_handle *h, int32_t *steps_out) {
  if (!h || !steps_out) return fail(QD_ERR_NULL, "NULL argument");
  *steps_out = h->steps;
  return QD_OK;
}

int qd_get_timing(qd_handle *h, int32_t which, double *ms_out) {
  if (!h || !ms_out) return fail(QD_ERR_NULL, "NULL argument");
  if (which < 0 || which >= TK_N) return fail(QD_ERR_ARG, "bad timing selector");
  if (which == TK_STEP) {
    *ms_out = h->last_step_ms;
  } else {
    *ms_out = h->t_n[which] ? h->t_ms[which] / h->t_n[which] : 0.0;
  }
  return QD_OK;
}

int qd_destroy(qd_handle *h) {
  if (!h) return QD_OK;
  (void)hipSetDevice(h->dev);
  release(h);
  delete h;
  return QD_OK;
}

int qd_time_evolution(qd_problem *problem, const qd_options *options) {
  qd_handle *h = nullptr;
  int rc = qd_create(problem, options, &h);
  if (rc != QD_OK) return rc;
  rc = qd_propagate(h, problem->time.total_steps);
  if (rc == QD_OK) rc = qd_get_psi(h, problem->psi);
  if (rc == QD_OK && h->do_crp && problem->time.total_steps > 0) rc = qd_get_crp(h, problem->crp.CRP);
  if (rc == QD_OK) problem->time.steps = h->steps;
  qd_destroy(h);
  return rc;
}

}  // extern "C"
