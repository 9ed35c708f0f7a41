// qd_api.hip -- host side of the C ABI declared in include/qd.h.
//
// The handle owns every device buffer and the stream; nothing is global, so
// handles on different threads or devices are independent (the reference's
// CUDA path keeps module-global __constant__ state, evolutionCUDA2.cu:19-24).
//
// A handle propagates a batch of (J, K) blocks on shared R, r grids.  The
// reference's J = 0 problem (qd_create / qd_time_evolution) is a batch of one
// block; qd_create_jk builds the body-fixed K blocks of several J.
//
// Per step (fused pipeline, CPU-propagator semantics, timeEvol.C:657-684):
//   k_theta_*<V prologue>     psi -> lpsi                   (DVR -> FBR)
//   k_rpass                   K->l, D_R/2, FFT_R, T1, IFFT_R, D_R/2, l->K
//   k_rline                   D_r/2, FFT_r, T2, IFFT_r, D_r/2
//   k_theta_*<dump epilogue>  lpsi -> psi                   (FBR -> DVR)
//   k_flux_extract            psi_s, dpsi_s -> time-block slot
// and every flux_block steps k_flux_accumulate folds the block into the
// energy-resolved accumulators (CRPAux.F:4-31).  P(E) is evaluated on demand.
// With the fused theta step (R pass, r pass, theta step) the extraction of
// step n runs in the first work groups of step n + 1's register-FFT r pass,
// beside its line transforms; a fold, a read of psi and the end of a
// propagate call run it on its own first.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/qd.h"
#include "qd_fft.hpp"
#include "qd_rfft.hpp"
#include "qd_kernels.hpp"
#include "qd_mix.hpp"
#include "qd_dispatch.hpp"
#include "qd_pack.hpp"
#include "qd_records.hpp"
#include "qd_setup.hpp"
#include "qd_tables.hpp"

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
#define QD_RC(x)                  \
  do {                            \
    int rc_ = (x);                \
    if (rc_ != QD_OK) return rc_; \
  } while (0)

enum { TK_STEP = 0, TK_THETA = 1, TK_FFTR1 = 2, TK_FFTR2 = 3, TK_FLUX = 4, TK_OTHER = 5, TK_ACC = 6, TK_N = 7 };

// bf16 round-to-nearest-even of a float, as the device conversion does
uint16_t bf16_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
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

int max_butterflies(const FftPlanArg &p) {
  int m = 0;
  for (int i = 0; i < p.npass; i++) m = std::max(m, p.n / p.radix[i]);
  return m;
}

int pow2ceil(int v) {
  int p = 1;
  while (p < v) p *= 2;
  return p;
}

// Matrices of one theta basis: Wf[k][j] (nd x nf), Pb[j][k] (nf x nd), DVR weights.
struct MatSet {
  int nd = 0, nf = 0;
  std::vector<double> Wf, Pb, w;
};

// One block of the batch as the builder sees it.
struct HostBlock {
  int J = 0, K = 0;
  int mat = 0;                  // theta matrix set
  int vset = 0;                 // potential set (blocks of one K share it)
  const double *psi = nullptr;  // initial psi (interleaved), n12 * nd complex
  const double *eta_sq = nullptr;  // |eta(E)|^2 normalising this block's flux (NULL: crp.eta_sq)
};

// Mixing unit group of the R pass: FBR columns (eigen slot m <-> column of K = m) and U.
struct Group {
  std::vector<int> cols;
  int u0 = -1;
};

// Everything needed to build a handle, from either entry point.
struct BatchSpec {
  const qd_problem *base = nullptr;       // grids, time, dumps, flux parameters
  std::vector<HostBlock> blocks;
  std::vector<MatSet> mats;
  std::vector<const double *> vsets;      // potential of each set, n12 * nd
  std::vector<long long> fcol_lamval;     // lambda = (J - j)^2 eigenvalue per FBR column (l basis)
  std::vector<int> fcol_j;                // j per FBR column
  std::vector<Group> groups;
  std::vector<double> umat;               // concatenated U matrices (row K, column m)
  int jmax = 0;
};

}  // namespace

// One class of bundles of the FP32 mixing R pass (k_rpass_mx): its kernel
// shape and its device records.
struct MxSet {
  int nl = 0, gb = 0, nbundle = 0;
  long long slots = 0, slots2 = 0;  // resident work groups with one / two per CU (0: does not fit)
  size_t lds = 0;
  MxHead *head = nullptr;
  long long *off = nullptr;
  int *lam = nullptr, *unit = nullptr;
};

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
  int n1 = 0, n2 = 0, ne = 0;
  int nb = 0, ndcol = 0, nfcol = 0;
  size_t n12 = 0, npt = 0, nlp = 0, nx = 0;
  double dt = 0, dr1 = 0, dr2 = 0, mu2 = 0;
  bool has_dump = false;
  bool do_crp = false;
  bool has_mix = false;
  int surf = 0;  // 0-based surface row
  StencilArg stencil{};
  std::vector<BlockDesc> hblocks;
  // state
  int32_t steps = 0;
  bool pending_prephase = false;
  long long acc_steps = 0;            // host mirror of d_acc_steps
  long long *d_acc_steps = nullptr;   // steps already accumulated into fai (device)
  int slot = 0;
  // the last fused step's flux-surface extraction, not yet run (it goes into
  // the next r pass, or runs on its own before anything reads its slot or psi)
  bool ext_pending = false;
  int ext_slot = 0;
  bool ext_in_rline = true;  // defer the extraction into the r pass (test hook: off)
  // captured flux block (tblock steps + accumulation) for qd_options.use_graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int use_graph = 0;
  bool tuned = false;
  qd_step_cb cb = nullptr;
  void *cb_user = nullptr;
  hipStream_t stream = nullptr;
  // device buffers
  double2 *psi64 = nullptr, *lpsi64 = nullptr;
  float2 *psi32 = nullptr, *lpsi32 = nullptr;
  double *V = nullptr;
  float2 *expV32 = nullptr;
  long long *dcol_vidx = nullptr;
  double *dcol_w = nullptr;
  int *fcol_lam = nullptr, *fcol_j = nullptr;
  BlockDesc *d_blocks = nullptr;
  // the same descriptors ordered by K (potential set), then J: the matrix-stationary
  // theta kernels take their blocks in this order, so the blocks that read one
  // exp(-i V dt) set run together and share it through the caches
  BlockDesc *d_blocks_k = nullptr;
  double *d1 = nullptr, *d2 = nullptr;
  float *d1f = nullptr, *d2f = nullptr;
  void *dR = nullptr, *dr = nullptr, *kinR = nullptr, *kinr = nullptr, *twR = nullptr, *twr = nullptr;
  // shear-form tables of the FP32 register passes (rf_tab<float>); FP64: the ones above
  void *ltwR = nullptr, *ltwr = nullptr, *lkinR = nullptr, *lkinr = nullptr, *lphR = nullptr, *lphr = nullptr;
  void *Wf = nullptr, *Wb = nullptr;  // packed FP64 [mat][Kpad][Npad] or bf16 [mat][3][Npad][Kpad]
  long long mat_stride = 0;
  int mat_Kpad = 0, mat_Npad = 0, ncb = 8, f64_cb = 8;
  int col_tiles = 1;
  int *bundle_u0 = nullptr;
  RUnit *units = nullptr;
  int *unit_cols = nullptr;
  void *umat = nullptr;
  int nbundle = 0, rp_tpl = 16, rp_ml = 8;
  // matrix-stationary mixed theta kernel: k-steps (0: tiled kernel), tiles per work group
  int tmx_ks = 0, tmx_tpw = 0;
  // register-FFT passes (fused forward / kinetic / inverse with phases)
  bool rp_rf = false, rl_rf = false;
  int rp_lw = 0;
  long long rp_slots = 1, rl_slots = 1;  // resident work groups of the persistent passes
  int rl_nth = 256;                       // threads per work group of the register-FFT r pass (256 or 512)
  long long rl_slots2 = 0;                // resident 512-thread work groups of it (0: does not fit)
  int *bundle_l0 = nullptr, *line_lam = nullptr;  // mixing R pass: per-bundle line tables
  // FP32 mixing R pass with prefetch (k_rpass_mx): bundle records
  bool rp_mx = false, rp_legacy = true;  // legacy: the LDS-staged and Stockham R passes are set up
  // [0]: 32-line bundles (groups of at most 32 lines), [1]: 48-line bundles
  MxSet mxs[2];
  bool mx_two = false;                   // 32-line class: the variant sized for two work groups per CU
  int mx_chunk = MX_CHUNK;               // consecutive bundles per turn of a work group
  float *mx_umat = nullptr;              // unit matrices of both classes
  long long *line_off = nullptr;
  // split mixing R pass: K -> l mixing into a scratch array, the register FFT
  // pass without mixing there, l -> K mixing back (any group size)
  bool rp_split = false, sp_ok = false;
  int sp_nunits = 0, sp_wpu = 1, sp_lw = 0;  // units, waves per unit, lines per FFT work group
  long long sp_slots = 1;
  void *sp_tmp = nullptr;  // scratch of nlp points when psi is smaller
  size_t tmx_lds = 0;
  // fused theta step (k_theta_fused): between steps the state is the FBR array
  // holding theta-forward(V psi); psi itself is rebuilt from it when read
  bool fuse_theta = false, state_fbr = false, dvr_valid = true;
  int fuse_opt = 0, tfu_tpw = 1, tfu_nth = TMX_THREADS;  // tile-blocks per item, threads per work group
  size_t tfu_lds = 0;
  // the fused step on the parity halves of the theta transforms (k_theta_fused_par)
  bool tfu_par = false, par_ok = false, dense_ok = false;  // kernel in use; which kernels fit
  int par_ksh = 1;                                         // k-steps of 32 per half basis
  // half bases of 33..64: each tile shared by a pair of waves (k_theta_fused_pair)
  bool tfu_pair = false, pair_ok = false;
  // wave-pair kernel: the pair meets through LDS counters instead of work-group
  // barriers (on a 256 x 192 x 120 grid 47.5 -> 45.5 us per launch, the same bits)
  bool tfu_psync = true;
  void *Wpar = nullptr;
  long long par_stride = 0;
  // fused theta kernel: groups of blocks (in d_blocks_k order) sharing the theta
  // matrices and the V columns, and the work items built from them
  std::vector<std::pair<int, int>> tfu_groups;  // first block, blocks
  std::vector<TfuBlock> tfu_blocks;             // the blocks in K order, for the bounds checks
  long long nV = 0;                             // values of V (all potential sets)
  int4 *d_tfu_items = nullptr;
  int tfu_nitems = 0, tfu_items_cap = 0;

  bool rp_ct = false, rl_ct = false;
  size_t rp_lds = 0;
  FftPlanArg planR{}, planr{};
  int rl_lpb = 16, rl_blocks = 0;
  size_t rl_lds = 0;
  // the reference's per-step diagnostics (options.diagnostics_every): device
  // tables [psq1 n1][psq2 n2][I1 n1][I2 n2][1/(l+1/2) nfcol][l(l+1)/(l+1/2) nfcol]
  int diag_every = 0;
  bool diag_ok = false, diag_have = false;
  double *diag_tab = nullptr;
  double diag_last[5] = {};
  // FP64 staging for host <-> device copies of FP32 complex arrays (mixed
  // precision): STAGE_POINTS values, allocated with the handle
  double2 *stage = nullptr;
  double2 *bpsi = nullptr, *bdpsi = nullptr, *fai = nullptr, *dfai = nullptr;
  double2 *ftau = nullptr;  // tau[t][e] of the current time block (k_flux_tau)
  // partial P(E) sums of each 128-value x range, left by every fold
  // (k_flux_accumulate): used where every block starts on such a range
  double *crp_part = nullptr;
  bool part_ok = false, part_valid = true;
  double *energies = nullptr, *eta_sq = nullptr, *crp = nullptr, *partial = nullptr;
  int npartial = 256;
  // timing
  hipEvent_t ev[2] = {};
  std::vector<hipEvent_t> marks;
  std::vector<int> mark_cls;
  int nmarks = 0;
  double t_ms[TK_N] = {};
  long long t_n[TK_N] = {};
  double last_step_ms = 0;
  std::vector<double> host_crp;  // [nb][ne]
};

namespace {

template <typename T>
int upload(T **dst, const T *src, size_t n) {
  QD_HIP(hipMalloc((void **)dst, std::max<size_t>(n, 1) * sizeof(T)));
  if (n) QD_HIP(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return QD_OK;
}

inline dim3 grid1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

constexpr size_t STAGE_POINTS = size_t(1) << 22;  // 4 Mi complex values: 64 MiB of FP64 staging

// n host complex values (interleaved FP64) into the handle's complex array dst
// at element offset off, in the handle's precision.  The mixed path converts
// through the staging buffer in chunks, so no full-size temporary is made.
int put_complex(qd_handle *h, void *dst, size_t off, const double *src, size_t n) {
  if (h->prec == QD_PREC_FP64) {
    QD_HIP(hipMemcpy(static_cast<double2 *>(dst) + off, src, sizeof(double2) * n, hipMemcpyHostToDevice));
    return QD_OK;
  }
  for (size_t c = 0; c < n; c += STAGE_POINTS) {
    const size_t m = std::min(STAGE_POINTS, n - c);
    QD_HIP(hipMemcpyAsync(h->stage, src + 2 * c, sizeof(double2) * m, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_f64_to_f32, grid1(m), dim3(256), 0, h->stream, h->stage, static_cast<float2 *>(dst) + off + c,
                       m);
    QD_HIP(hipGetLastError());
  }
  QD_HIP(hipStreamSynchronize(h->stream));
  return QD_OK;
}
// n complex values of the handle's array src (from element offset off) to the
// host as interleaved FP64.
int get_complex(qd_handle *h, const void *src, size_t off, double *dst, size_t n) {
  if (h->prec == QD_PREC_FP64) {
    QD_HIP(hipMemcpyAsync(dst, static_cast<const double2 *>(src) + off, sizeof(double2) * n, hipMemcpyDeviceToHost,
                          h->stream));
  } else {
    for (size_t c = 0; c < n; c += STAGE_POINTS) {
      const size_t m = std::min(STAGE_POINTS, n - c);
      hipLaunchKernelGGL(k_f32_to_f64, grid1(m), dim3(256), 0, h->stream, static_cast<const float2 *>(src) + off + c,
                         h->stage, m);
      QD_HIP(hipGetLastError());
      QD_HIP(hipMemcpyAsync(dst + 2 * c, h->stage, sizeof(double2) * m, hipMemcpyDeviceToHost, h->stream));
    }
  }
  QD_HIP(hipStreamSynchronize(h->stream));
  return QD_OK;
}

// ---- profiling: a chain of events, one before every kernel, drained in batches ----
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
  if (h->nmarks == (int)h->marks.size()) QD_RC(drain_marks(h));
  return QD_OK;
}

constexpr int QD_RF_FLAGS = FFT_PHASE | FFT_FWD | FFT_KIN | FFT_INV;

// ---- launches ----
template <typename T>
int launch_rpass(qd_handle *h, int flags, typename cx_of<T>::type *data) {
  typedef typename cx_of<T>::type C;
  RPassArgs<T> a;
  a.data = data;
  typedef typename tab_of<T>::type TB;
  a.tw = (const TB *)h->twR;
  a.kin = (const TB *)h->kinR;
  a.dR = (const TB *)h->dR;
  constexpr bool f64 = sizeof(T) == 8;  // FP64: the register passes read the plain tables
  a.ltw = (const typename rf_tab<T>::tw *)(f64 ? h->twR : h->ltwR);
  a.lkin = (const typename rf_tab<T>::kin *)(f64 ? h->kinR : h->lkinR);
  a.lph = (const typename rf_tab<T>::ph *)(f64 ? h->dR : h->lphR);
  a.fcol_lam = h->fcol_lam;
  a.bundle_u0 = h->bundle_u0;
  a.units = h->units;
  a.unit_cols = h->unit_cols;
  a.umat = (const T *)h->umat;
  a.n1 = h->n1;
  a.n2 = h->n2;
  a.tpl = h->rp_tpl;
  a.flags = flags;
  if (h->has_mix && (flags & FFT_PHASE)) a.flags |= FFT_MIX;
  a.plan = h->planR;
  a.nfcol = (int)h->nfcol;
  a.bundle_l0 = h->bundle_l0;
  a.line_off = h->line_off;
  a.line_lam = h->line_lam;
  a.nbundle = h->nbundle;
  if (h->rp_split && (flags & QD_RF_FLAGS) == QD_RF_FLAGS) {
    C *tmp = (C *)(h->sp_tmp ? h->sp_tmp : (h->prec == QD_PREC_FP64 ? (void *)h->psi64 : (void *)h->psi32));
    const unsigned mblk = (unsigned)(((long long)h->sp_nunits * h->sp_wpu + 3) / 4);
    hipLaunchKernelGGL(k_mix_lines<T>, dim3(mblk), dim3(256), 0, h->stream, (const C *)data, tmp, h->units,
                       h->unit_cols, (const T *)h->umat, h->sp_nunits, h->n1, h->n2, h->sp_wpu, 1);
    QD_HIP(hipGetLastError());
    RPassArgs<T> b = a;
    b.data = tmp;
    b.flags = flags & ~FFT_MIX;
    size_t lds = 0;
    const void *fn = rpass_rf_kernel<T>(h->n1, false, h->sp_lw, &lds);
    void *args[] = {&b};
    const unsigned nblk =
        (unsigned)std::min(h->sp_slots, ((long long)h->nfcol * h->n2 + h->sp_lw - 1) / h->sp_lw);
    QD_HIP(hipLaunchKernel(fn, dim3(nblk), dim3(h->sp_lw * rf_n1(h->n1)), args, lds, h->stream));
    QD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mix_lines<T>, dim3(mblk), dim3(256), 0, h->stream, (const C *)tmp, data, h->units,
                       h->unit_cols, (const T *)h->umat, h->sp_nunits, h->n1, h->n2, h->sp_wpu, 0);
    QD_HIP(hipGetLastError());
    return QD_OK;
  }
  if constexpr (sizeof(T) == 4) {
    if (h->rp_mx && (flags & QD_RF_FLAGS) == QD_RF_FLAGS) {
      RMixArgs m;
      m.data = (float2 *)data;
      m.tw = (const float2 *)h->ltwR;
      m.kin = (const float4 *)h->lkinR;
      m.dR = (const float4 *)h->lphR;
      m.umat = h->mx_umat;
      m.chunk = h->mx_chunk;
      // the two bundle classes touch disjoint lines
      for (int c = 0; c < 2; c++) {
        const MxSet &M = h->mxs[c];
        if (M.nbundle == 0) continue;
        const bool two = c == 0 && h->mx_two && M.slots2 > 0;
        m.head = M.head;
        m.line_off = M.off;
        m.line_lam = M.lam;
        m.unit = M.unit;
        m.nbundle = M.nbundle;
        size_t lds0 = 0;
        const void *fn = rpass_mx_kernel(h->n1, M.nl, M.gb, two, &lds0);
        if (!fn) return fail(QD_ERR_ARG, "no mixing R-pass kernel for this bundle class");
        void *args[] = {&m};
        const unsigned nblk =
            (unsigned)std::min(two ? M.slots2 : M.slots, ((long long)M.nbundle + h->mx_chunk - 1) / h->mx_chunk);
        QD_HIP(hipLaunchKernel(fn, dim3(nblk), dim3(M.nl * rf_n1(h->n1)), args, M.lds, h->stream));
        QD_HIP(hipGetLastError());
      }
      return QD_OK;
    }
  }
  if (!h->rp_legacy) return fail(QD_ERR_ARG, "this R-pass stage needs mixing groups of at most 32 lines");
  if (h->rp_rf && (flags & QD_RF_FLAGS) == QD_RF_FLAGS) {
    size_t lds = 0;
    const void *fn = rpass_rf_kernel<T>(h->n1, h->has_mix, h->rp_lw, &lds);
    void *args[] = {&a};
    const int lw = h->rp_lw;
    const int nthreads = lw * rf_n1(h->n1);
    const unsigned nblk = (unsigned)std::min(h->rp_slots, h->has_mix ? (long long)h->nbundle
                                                                     : ((long long)h->nfcol * h->n2 + lw - 1) / lw);
    QD_HIP(hipLaunchKernel(fn, dim3(nblk), dim3(nthreads), args, lds, h->stream));
    QD_HIP(hipGetLastError());
    return QD_OK;
  }
  const void *fn = h->rp_ct ? rpass_kernel<T>(h->n1, h->rp_ml, h->has_mix) : nullptr;
  if (fn) {
    void *args[] = {&a};
    QD_HIP(hipLaunchKernel(fn, dim3(h->nbundle), dim3(256), args, h->rp_lds, h->stream));
  } else {
    hipLaunchKernelGGL((k_rpass<T>), dim3(h->nbundle), dim3(RP_THREADS), h->rp_lds, h->stream, a);
  }
  QD_HIP(hipGetLastError());
  return QD_OK;
}

// arguments of the flux-surface extraction into time-block slot `slot`
template <typename T>
FluxExtArgs<T> flux_ext_args(const qd_handle *h, int slot) {
  FluxExtArgs<T> E;
  E.psi = (const typename cx_of<T>::type *)(sizeof(T) == 8 ? (const void *)h->psi64 : (const void *)h->psi32);
  E.bpsi = h->bpsi;
  E.bdpsi = h->bdpsi;
  E.n1 = h->n1;
  E.n2 = h->n2;
  E.ncols = (int)h->ndcol;
  E.s = h->surf;
  E.slot = slot;
  E.dr = h->dr2;
  E.st = h->stencil;
  E.nwg = 0;
  return E;
}

template <typename T>
int launch_flux_extract(qd_handle *h, int slot) {
  hipLaunchKernelGGL(k_flux_extract<T>, grid1(h->nx), dim3(256), 0, h->stream, flux_ext_args<T>(h, slot));
  QD_HIP(hipGetLastError());
  return QD_OK;
}

// runs the deferred extraction of the last fused step on its own
int resolve_ext(qd_handle *h) {
  if (!h->ext_pending) return QD_OK;
  h->ext_pending = false;
  QD_RC(mark(h, TK_FLUX));
  return h->prec == QD_PREC_FP64 ? launch_flux_extract<double>(h, h->ext_slot)
                                 : launch_flux_extract<float>(h, h->ext_slot);
}

// ext_slot >= 0: the register-FFT pass also extracts the flux surface into
// that time-block slot, in work groups after its own
template <typename T>
int launch_rline(qd_handle *h, int flags, typename cx_of<T>::type *data, int ext_slot = -1) {
  typedef typename cx_of<T>::type C;
  RLineArgs<T> a;
  a.data = data;
  typedef typename tab_of<T>::type TB;
  a.tw = (const TB *)h->twr;
  a.kin = (const TB *)h->kinr;
  a.dr = (const TB *)h->dr;
  constexpr bool f64 = sizeof(T) == 8;  // FP64: the register pass reads the plain tables
  a.ltw = (const typename rf_tab<T>::tw *)(f64 ? h->twr : h->ltwr);
  a.lkin = (const typename rf_tab<T>::kin *)(f64 ? h->kinr : h->lkinr);
  a.lph = (const typename rf_tab<T>::ph *)(f64 ? h->dr : h->lphr);
  a.fcol_j = h->fcol_j;
  a.n1 = h->n1;
  a.n2 = h->n2;
  a.lpb = h->rl_lpb;
  a.nlines = h->n1 * h->nfcol;
  a.flags = flags;
  a.plan = h->planr;
  a.ext = flux_ext_args<T>(h, ext_slot);
  if (h->rl_rf && (flags & QD_RF_FLAGS) == QD_RF_FLAGS) {
    size_t lds = 0;
    int lpb = 0;
    const void *fn = rline_rf_kernel<T>(h->n2, h->rl_nth, &lds, &lpb);
    if (ext_slot >= 0) a.ext.nwg = (int)((h->nx + h->rl_nth - 1) / h->rl_nth);
    void *args[] = {&a};
    const long long slots = h->rl_nth == 512 ? h->rl_slots2 : h->rl_slots;
    const unsigned nblk = (unsigned)std::min(slots, ((long long)h->n1 * h->nfcol + lpb - 1) / lpb);
    QD_HIP(hipLaunchKernel(fn, dim3(nblk + (unsigned)a.ext.nwg), dim3(h->rl_nth), args, lds, h->stream));
    QD_HIP(hipGetLastError());
    return QD_OK;
  }
  if (ext_slot >= 0) QD_RC(launch_flux_extract<T>(h, ext_slot));
  const void *fn = h->rl_ct ? rline_kernel<T>(h->n2, h->rl_lpb) : nullptr;
  if (fn) {
    void *args[] = {&a};
    QD_HIP(hipLaunchKernel(fn, dim3(h->rl_blocks), dim3(256), args, h->rl_lds, h->stream));
  } else {
    hipLaunchKernelGGL((k_rline<T>), dim3(h->rl_blocks), dim3(256), h->rl_lds, h->stream, a);
  }
  QD_HIP(hipGetLastError());
  return QD_OK;
}

int launch_theta(qd_handle *h, bool fwd, int pro, int epi) {
  ThetaArgs P;
  P.A = h->prec == QD_PREC_FP64 ? (const void *)(fwd ? h->psi64 : h->lpsi64)
                                : (const void *)(fwd ? h->psi32 : h->lpsi32);
  P.C = h->prec == QD_PREC_FP64 ? (void *)(fwd ? h->lpsi64 : h->psi64) : (void *)(fwd ? h->lpsi32 : h->psi32);
  P.B = fwd ? h->Wf : h->Wb;
  P.mat_stride = h->mat_stride;
  P.Kpad = h->mat_Kpad;
  P.Npad = h->mat_Npad;
  P.blocks = h->d_blocks;
  P.fwd = fwd ? 1 : 0;
  P.R = (int)h->n12;
  P.V = h->V;
  P.expV = h->expV32;
  P.dcol_vidx = h->dcol_vidx;
  P.dt = h->dt;
  P.d1 = h->d1;
  P.d2 = h->d2;
  P.d1f = h->d1f;
  P.d2f = h->d2f;
  P.n1 = h->n1;
  if (h->tmx_ks) {
    P.tpw = h->tmx_tpw;
    P.blocks = h->d_blocks_k;
    const void *fn = theta_mx_kernel(h->tmx_ks, pro, epi);
    if (!fn) return fail(QD_ERR_ARG, "bad theta variant");
    const long long ntiles = (h->n12 + 15) / 16;
    dim3 grid((unsigned)((ntiles + h->tmx_tpw - 1) / h->tmx_tpw), 1, h->nb);
    void *args[] = {&P};
    QD_HIP(hipLaunchKernel(fn, grid, dim3(TMX_THREADS), args, h->tmx_lds, h->stream));
    return QD_OK;
  }
  const int rows = (int)((h->n12 + TG_ROWS - 1) / TG_ROWS);
  dim3 grid(rows, h->col_tiles, h->nb);
  if (h->prec == QD_PREC_FP64) {
#define QD_T64(PR, EP)                                                                       \
  do {                                                                                       \
    if (h->f64_cb == 4) hipLaunchKernelGGL((k_theta_f64<PR, EP, 4>), grid, dim3(256), 0, h->stream, P); \
    else hipLaunchKernelGGL((k_theta_f64<PR, EP, 8>), grid, dim3(256), 0, h->stream, P);    \
  } while (0)
    if (pro == 0 && epi == 0) QD_T64(0, 0);
    else if (pro == 1 && epi == 0) QD_T64(1, 0);
    else if (pro == 2 && epi == 0) QD_T64(2, 0);
    else if (pro == 0 && epi == 1) QD_T64(0, 1);
    else return fail(QD_ERR_ARG, "bad theta variant");
#undef QD_T64
  } else {
#define QD_T16(NS_, NCB_, PR, EP) \
  hipLaunchKernelGGL((k_theta_bf16<NS_, NCB_, PR, EP>), grid, dim3(256), 0, h->stream, P)
#define QD_T16_SEL(PR, EP)                                   \
  do {                                                       \
    if (h->nsplit == 2 && h->ncb == 8) QD_T16(2, 8, PR, EP); \
    else if (h->nsplit == 2) QD_T16(2, 4, PR, EP);           \
    else if (h->ncb == 8) QD_T16(3, 8, PR, EP);              \
    else QD_T16(3, 4, PR, EP);                               \
  } while (0)
    if (pro == 0 && epi == 0) QD_T16_SEL(0, 0);
    else if (pro == 1 && epi == 0) QD_T16_SEL(1, 0);
    else if (pro == 2 && epi == 0) QD_T16_SEL(2, 0);
    else if (pro == 0 && epi == 1) QD_T16_SEL(0, 1);
    else return fail(QD_ERR_ARG, "bad theta variant");
#undef QD_T16_SEL
#undef QD_T16
  }
  QD_HIP(hipGetLastError());
  return QD_OK;
}

// rows: the flux surface's psi rows (false leaves psi untouched, for the
// autotuner); all: every row of psi (the CUDA propagator's order, whose
// pre-dump psi the FBR state cannot give back)
int launch_theta_fused(qd_handle *h, bool rows = true, bool all = false) {
  ThetaFusedArgs F;
  F.L = h->lpsi32;
  F.psi = h->psi32;
  F.Bb = h->tfu_par ? h->Wpar : h->Wb;
  F.Bf = h->Wf;
  F.mat_stride = h->tfu_par ? h->par_stride : h->mat_stride;
  F.blocks = h->d_blocks_k;
  F.R = (int)h->n12;
  F.expV = h->expV32;
  F.dcol_vidx = h->dcol_vidx;
  F.d1f = h->has_dump ? h->d1f : nullptr;
  F.d2f = h->has_dump ? h->d2f : nullptr;
  F.n1 = h->n1;
  F.r_lo = all ? 0 : h->do_crp && rows ? h->surf - h->stencil.h : 1;
  F.r_hi = all ? h->n2 - 1 : h->do_crp && rows ? h->surf + h->stencil.h : 0;
  F.pre_dump = h->variant == QD_VARIANT_CUDA ? 1 : 0;
  F.items = h->d_tfu_items;
  size_t lds = 0;
  const void *fn = h->tfu_pair ? theta_fused_pair_kernel(h->tfu_nth, h->tfu_psync, &lds)
                   : h->tfu_par ? theta_fused_par_kernel(h->par_ksh, h->tfu_nth, &lds)
                                : theta_fused_kernel(h->tmx_ks, h->tfu_nth, &lds);
  if (!fn) return fail(QD_ERR_ARG, "no fused theta kernel for this basis size");
  if (h->tfu_nitems < 1) return fail(QD_ERR_ARG, "fused theta work items not set up");
  dim3 grid((unsigned)h->tfu_nitems, 1, 1);
  void *args[] = {&F};
  QD_HIP(hipLaunchKernel(fn, grid, dim3(h->tfu_nth), args, lds, h->stream));
  QD_HIP(hipGetLastError());
  return QD_OK;
}

// Work items of the fused theta kernel for the current tfu_tpw: each group's
// tiles in ranges of max(8, tfu_tpw / blocks) tiles, so that every item holds
// about tfu_tpw tile-blocks.  Set-up and autotuning only (a synchronous copy).
int build_tfu_items(qd_handle *h) {
  const long long ntiles = ((long long)h->n12 + 15) / 16;
  // tiles taken side by side: one per wave, or one per wave pair
  const int lanes = h->tfu_pair ? h->tfu_nth / 128 : h->tfu_nth / 64;
  const std::vector<TfuItem> items = qd::build_tfu_items(h->tfu_groups, ntiles, h->tfu_tpw, lanes);
  const std::string err = check_tfu_items(items, h->tfu_blocks, (long long)h->n12, h->ndcol, h->nfcol, h->nV);
  if (!err.empty()) return fail(QD_ERR_ARG, err);
  if ((int)items.size() > h->tfu_items_cap) {
    QD_HIP(hipStreamSynchronize(h->stream));  // no launch still reads the old items
    if (h->d_tfu_items) QD_HIP(hipFree(h->d_tfu_items));
    h->d_tfu_items = nullptr;
    h->tfu_items_cap = 0;
    QD_HIP(hipMalloc((void **)&h->d_tfu_items, items.size() * sizeof(int4)));
    h->tfu_items_cap = (int)items.size();
  }
  static_assert(sizeof(TfuItem) == sizeof(int4), "work items are read as int4");
  QD_HIP(hipStreamSynchronize(h->stream));
  QD_HIP(hipMemcpy(h->d_tfu_items, items.data(), items.size() * sizeof(TfuItem), hipMemcpyHostToDevice));
  h->tfu_nitems = (int)items.size();
  return QD_OK;
}

// Into the fused state: the FBR array becomes theta-forward(V psi).
int enter_fbr_state(qd_handle *h) {
  if (!h->fuse_theta || h->state_fbr) return QD_OK;
  // the CUDA propagator's step starts with the dump (psi itself is not changed)
  QD_RC(launch_theta(h, true, h->variant == QD_VARIANT_CUDA && h->has_dump ? 2 : 1, 0));
  h->state_fbr = true;
  return QD_OK;
}
// psi from the fused state: conj(exp(-i V dt)) times theta-backward of the
// FBR array (the forward and backward transforms are exact inverses).  In the
// CUDA propagator's order the FBR state holds the dumped psi, so there the
// fused step that ends a call, precedes a diagnostics step or lands on a
// callback point writes all of psi instead (step_impl), and psi is always
// valid where it is read.
int materialize_psi(qd_handle *h) {
  QD_RC(resolve_ext(h));  // it reads the psi rows this would overwrite
  if (!h->fuse_theta || !h->state_fbr || h->dvr_valid) return QD_OK;
  if (h->variant == QD_VARIANT_CUDA) return fail(QD_ERR_ARG, "internal: psi of the CUDA-order fused state is not held");
  QD_RC(launch_theta(h, false, 0, 2));
  h->dvr_valid = true;
  return QD_OK;
}

int flush_flux(qd_handle *h) {
  QD_RC(resolve_ext(h));
  if (!h->do_crp || h->slot == 0) return QD_OK;
  QD_RC(mark(h, TK_ACC));
  hipLaunchKernelGGL(k_flux_tau, grid1((size_t)h->slot * h->ne), dim3(256), 0, h->stream, h->ftau, h->energies, h->ne,
                     h->slot, h->d_acc_steps, h->dt);
  const long long nxb = (h->nx + FX_X - 1) / FX_X, ney = (h->ne + FX_E - 1) / FX_E;
  dim3 grid((unsigned)(((nxb + 7) / 8) * 8 * ney));
  hipLaunchKernelGGL(k_flux_accumulate, grid, dim3(256), 0, h->stream, h->bpsi, h->bdpsi, h->fai, h->dfai, h->ftau,
                     (int)h->nx, h->ne, h->slot, h->part_ok ? h->crp_part : nullptr, h->dcol_w, h->n1);
  hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(64), 0, h->stream, h->d_acc_steps, h->slot);
  QD_HIP(hipGetLastError());
  QD_RC(mark(h, -1));  // closes the accumulation's interval
  h->acc_steps += h->slot;
  h->slot = 0;
  h->part_valid = true;
  return QD_OK;
}

int compute_crp(qd_handle *h) {
  QD_RC(flush_flux(h));
  if (h->profile) QD_RC(drain_marks(h));
  const double sign = h->variant == QD_VARIANT_CUDA ? -1.0 : 1.0;
  if (h->part_ok && h->part_valid)
    hipLaunchKernelGGL(k_crp_parts, dim3(h->ne, h->nb), dim3(64), 0, h->stream, h->crp_part, h->eta_sq, h->d_blocks,
                       h->crp, h->n1, h->ne, h->dr1 / h->mu2, sign);
  else
    hipLaunchKernelGGL(k_crp, dim3(h->ne, h->nb), dim3(256), 0, h->stream, h->fai, h->dfai, h->dcol_w, h->eta_sq,
                       h->d_blocks, h->crp, h->n1, (int)h->nx, h->ne, h->dr1 / h->mu2, sign);
  QD_HIP(hipGetLastError());
  h->host_crp.resize((size_t)h->nb * h->ne);
  QD_HIP(hipMemcpyAsync(h->host_crp.data(), h->crp, sizeof(double) * h->host_crp.size(), hipMemcpyDeviceToHost,
                        h->stream));
  QD_HIP(hipStreamSynchronize(h->stream));
  return QD_OK;
}

int apply_prephase(qd_handle *h) {
  // FP64 arithmetic on the stored values of either precision, in place
  const bool cuda = h->variant == QD_VARIANT_CUDA;
  if (h->prec == QD_PREC_FP64) {
    if (cuda)
      hipLaunchKernelGGL(k_prephase_cuda<double2>, grid1(h->npt), dim3(256), 0, h->stream, h->psi64, h->V,
                         h->dcol_vidx, h->n12, h->dt, h->npt);
    else
      hipLaunchKernelGGL(k_prephase_cpu<double2>, grid1(h->npt), dim3(256), 0, h->stream, h->psi64, h->V,
                         h->dcol_vidx, h->n12, h->dt, h->npt);
  } else {
    if (cuda)
      hipLaunchKernelGGL(k_prephase_cuda<float2>, grid1(h->npt), dim3(256), 0, h->stream, h->psi32, h->V,
                         h->dcol_vidx, h->n12, h->dt, h->npt);
    else
      hipLaunchKernelGGL(k_prephase_cpu<float2>, grid1(h->npt), dim3(256), 0, h->stream, h->psi32, h->V,
                         h->dcol_vidx, h->n12, h->dt, h->npt);
  }
  QD_HIP(hipGetLastError());
  return QD_OK;
}

enum { DG_MOD = 0, DG_KIN = 1, DG_ROT = 2, DG_POT = 3 };

// One diagnostic partial sum (k_diag_sum) of a complex array of the handle's
// precision into slot `slot` of h->partial.
template <typename T>
int diag_sum(qd_handle *h, const typename cx_of<T>::type *a, size_t n, int slot) {
  const int n1 = h->n1, n2 = h->n2;
  const double *t = h->diag_tab;
  const double *psq1 = t, *psq2 = t + n1, *I1 = t + n1 + n2, *I2 = t + 2 * n1 + n2;
  const double *wkin = t + 2 * (n1 + n2), *wrot = wkin + h->nfcol;
  const double *wc = slot == DG_KIN ? wkin : slot == DG_ROT ? wrot : h->dcol_w;
  const double *fi = slot == DG_KIN ? psq1 : slot == DG_ROT ? I1 : nullptr;
  const double *fj = slot == DG_KIN ? psq2 : slot == DG_ROT ? I2 : nullptr;
  hipLaunchKernelGGL(k_diag_sum<T>, dim3(h->npartial), dim3(256), 0, h->stream, a, n, h->n12, n1, wc, fi, fj,
                     slot == DG_POT ? h->V : nullptr, h->dcol_vidx, h->partial + (size_t)slot * h->npartial);
  QD_HIP(hipGetLastError());
  return QD_OK;
}

// The reference's operator sequence, one kernel per operator (evolution_dt,
// timeEvol.C:619-647; dump_psi :713-735; the CUDA propagator dumps first,
// evolutionCUDA2.cu:503).  With diag, the four diagnostics are reduced at the
// points where the reference evaluates them.
template <typename T>
int plain_step(qd_handle *h, bool diag) {
  typedef typename cx_of<T>::type C;
  typedef typename tab_of<T>::type TB;
  C *psi = (C *)(h->prec == QD_PREC_FP64 ? (void *)h->psi64 : (void *)h->psi32);
  C *lpsi = (C *)(h->prec == QD_PREC_FP64 ? (void *)h->lpsi64 : (void *)h->lpsi32);
  const bool cuda = h->variant == QD_VARIANT_CUDA;
  const bool dump = h->has_dump;
  QD_RC(mark(h, TK_OTHER));
  if (cuda && dump)
    hipLaunchKernelGGL(k_dump<T>, grid1(h->npt), dim3(256), 0, h->stream, psi, h->d1, h->d2, h->n1, h->n2, h->npt);
  hipLaunchKernelGGL(k_vphase<T>, grid1(h->npt), dim3(256), 0, h->stream, psi,
                     h->prec == QD_PREC_FP64 ? h->V : nullptr, (const C *)h->expV32, h->dcol_vidx, h->n12, h->dt,
                     h->npt);
  QD_RC(mark(h, TK_THETA));
  QD_RC(launch_theta(h, true, 0, 0));
  QD_RC(mark(h, TK_OTHER));
  hipLaunchKernelGGL(k_rot<T>, grid1(h->nlp), dim3(256), 0, h->stream, lpsi, (const TB *)h->dR, (const TB *)h->dr,
                     h->fcol_lam, h->fcol_j, h->n1, h->n2, h->nlp);
  QD_RC(mark(h, TK_FFTR1));
  QD_RC(launch_rpass<T>(h, FFT_FWD, lpsi));
  QD_RC(mark(h, TK_FFTR2));
  QD_RC(launch_rline<T>(h, FFT_FWD, lpsi));
  QD_RC(mark(h, TK_OTHER));
  hipLaunchKernelGGL(k_kin2d<T>, grid1(h->nlp), dim3(256), 0, h->stream, lpsi, (const TB *)h->kinR,
                     (const TB *)h->kinr, h->n1, h->n2, h->nlp);
  if (diag) QD_RC(diag_sum<T>(h, lpsi, h->nlp, DG_KIN));
  QD_RC(mark(h, TK_FFTR2));
  QD_RC(launch_rline<T>(h, FFT_INV, lpsi));
  QD_RC(mark(h, TK_FFTR1));
  QD_RC(launch_rpass<T>(h, FFT_INV, lpsi));
  QD_RC(mark(h, TK_OTHER));
  hipLaunchKernelGGL(k_rot<T>, grid1(h->nlp), dim3(256), 0, h->stream, lpsi, (const TB *)h->dR, (const TB *)h->dr,
                     h->fcol_lam, h->fcol_j, h->n1, h->n2, h->nlp);
  if (diag) QD_RC(diag_sum<T>(h, lpsi, h->nlp, DG_ROT));
  QD_RC(mark(h, TK_THETA));
  QD_RC(launch_theta(h, false, 0, 0));
  if (diag) QD_RC(diag_sum<T>(h, psi, h->npt, DG_POT));
  if (!cuda && dump) {
    QD_RC(mark(h, TK_OTHER));
    hipLaunchKernelGGL(k_dump<T>, grid1(h->npt), dim3(256), 0, h->stream, psi, h->d1, h->d2, h->n1, h->n2, h->npt);
  }
  if (diag) QD_RC(diag_sum<T>(h, psi, h->npt, DG_MOD));
  QD_HIP(hipGetLastError());
  return QD_OK;
}

// Reads the four diagnostic sums of the step just enqueued (one host sync).
int finish_diag(qd_handle *h, int32_t steps_after) {
  std::vector<double> part(4 * (size_t)h->npartial);
  QD_HIP(hipMemcpyAsync(part.data(), h->partial, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h->stream));
  QD_HIP(hipStreamSynchronize(h->stream));
  double sum[4] = {0, 0, 0, 0};
  for (int k = 0; k < 4; k++)
    for (int b = 0; b < h->npartial; b++) sum[k] += part[(size_t)k * h->npartial + b];
  const double da = h->dr1 * h->dr2;
  h->diag_last[0] = steps_after;
  h->diag_last[1] = sum[DG_MOD] * da;
  h->diag_last[2] = sum[DG_KIN] * da * (double)h->n1 * h->n2;  // the kinetic table carries 1/(n1 n2)
  h->diag_last[3] = sum[DG_ROT] * da;
  h->diag_last[4] = sum[DG_POT] * da;
  h->diag_have = true;
  if (!std::isfinite(h->diag_last[1])) return fail(QD_ERR_NAN, "non-finite norm");
  return QD_OK;
}

// psi_out: psi is read after this step (the call ends, a diagnostics step or
// a callback follows); the CUDA-order fused step then writes all of it.
template <typename T>
int step_impl(qd_handle *h, bool diag = false, bool psi_out = false) {
  typedef typename cx_of<T>::type C;
  C *psi = (C *)(h->prec == QD_PREC_FP64 ? (void *)h->psi64 : (void *)h->psi32);
  C *lpsi = (C *)(h->prec == QD_PREC_FP64 ? (void *)h->lpsi64 : (void *)h->lpsi32);
  const bool cuda = h->variant == QD_VARIANT_CUDA;
  const bool dump = h->has_dump;
  if (diag) {
    // the diagnostics need the intermediate arrays: this step runs one kernel
    // per operator on psi; a fused-theta handle returns to its FBR state after it
    QD_RC(resolve_ext(h));
    QD_RC(materialize_psi(h));
    h->state_fbr = false;
    QD_RC(plain_step<T>(h, true));
  } else if (h->fuse_theta) {
    // state on entry: the FBR array holds theta-forward(V psi_n) (CUDA order:
    // of V dump psi_n); on exit the same for psi_{n+1}.  The psi rows of the
    // flux surface are written by the fused kernel, after the dump (CUDA
    // order: before it), for k_flux_extract below.
    const bool all = cuda && psi_out;
    QD_RC(mark(h, TK_FFTR1));
    QD_RC(launch_rpass<T>(h, FFT_PHASE | FFT_FWD | FFT_KIN | FFT_INV, lpsi));
    QD_RC(mark(h, TK_FFTR2));
    // the previous step's extraction (if pending) rides along with the r pass
    const int ext = h->ext_pending ? h->ext_slot : -1;
    h->ext_pending = false;
    QD_RC(launch_rline<T>(h, FFT_PHASE | FFT_FWD | FFT_KIN | FFT_INV, lpsi, ext));
    QD_RC(mark(h, TK_THETA));
    QD_RC(launch_theta_fused(h, true, all));
    h->dvr_valid = all;
  } else if (h->pipe == QD_PIPE_FUSED) {
    // CUDA propagator: dump at the start of the step (evolutionCUDA2.cu:503),
    // fused with the potential phase; CPU propagator: dump after the step.
    QD_RC(mark(h, TK_THETA));
    QD_RC(launch_theta(h, true, cuda && dump ? 2 : 1, 0));
    QD_RC(mark(h, TK_FFTR1));
    QD_RC(launch_rpass<T>(h, FFT_PHASE | FFT_FWD | FFT_KIN | FFT_INV, lpsi));
    QD_RC(mark(h, TK_FFTR2));
    QD_RC(launch_rline<T>(h, FFT_PHASE | FFT_FWD | FFT_KIN | FFT_INV, lpsi));
    QD_RC(mark(h, TK_THETA));
    QD_RC(launch_theta(h, false, 0, !cuda && dump ? 1 : 0));
  } else {
    QD_RC(plain_step<T>(h, false));
  }
  QD_HIP(hipGetLastError());
  if (h->do_crp) {
    if (h->fuse_theta && !diag && h->rl_rf && h->ext_in_rline) {
      // deferred to the next r pass (or run on its own by flush_flux,
      // materialize_psi or the end of the call)
      h->ext_pending = true;
      h->ext_slot = h->slot;
    } else {
      QD_RC(mark(h, TK_FLUX));
      QD_RC(launch_flux_extract<T>(h, h->slot));
    }
    h->slot++;
    if (h->slot == h->tblock) QD_RC(flush_flux(h));
  }
  if (diag) {
    if (h->fuse_theta) {
      QD_RC(enter_fbr_state(h));
      h->dvr_valid = true;
    }
    QD_RC(finish_diag(h, h->steps + 1));  // labelled with the step count after this step
  }
  return QD_OK;
}

// ---- construction ----
int validate_base(const qd_problem *p, bool need_theta) {
  if (!p) return fail(QD_ERR_NULL, "problem is NULL");
  if (!p->r1.r || !p->r2.r) return fail(QD_ERR_NULL, "r1.r or r2.r is NULL");
  if (need_theta && (!p->pot || !p->psi || !p->theta.x || !p->theta.w || !p->theta.legendre))
    return fail(QD_ERR_NULL, "a required array (pot, psi, theta.x/w/legendre) is NULL");
  if (p->r1.n <= 0 || p->r2.n <= 0) return fail(QD_ERR_ARG, "grid sizes must be positive");
  if (need_theta && (p->theta.n <= 0 || p->theta.m < 0)) return fail(QD_ERR_ARG, "theta sizes must be positive");
  if (p->r1.n % 2 || p->r2.n % 2) return fail(QD_ERR_ODD_N, "r1.n and r2.n must be even (FFT momentum grid)");
  if (p->time.total_steps < 0) return fail(QD_ERR_ARG, "total_steps < 0");
  if (!(p->r1.mass > 0) || !(p->r2.mass > 0) || !(p->r1.dr > 0) || !(p->r2.dr > 0))
    return fail(QD_ERR_ARG, "masses and spacings must be positive");
  return QD_OK;
}

int set_options(qd_handle *h, const qd_options *o) {
  qd_options opt{};
  if (o) opt = *o;
  if (opt.precision != QD_PREC_FP64 && opt.precision != QD_PREC_MIXED) return fail(QD_ERR_ARG, "bad precision");
  if (opt.pipeline != QD_PIPE_FUSED && opt.pipeline != QD_PIPE_PLAIN) return fail(QD_ERR_ARG, "bad pipeline");
  if (opt.variant != QD_VARIANT_CPU && opt.variant != QD_VARIANT_CUDA) return fail(QD_ERR_ARG, "bad variant");
  if (opt.theta_split != 0 && opt.theta_split != 2 && opt.theta_split != 3)
    return fail(QD_ERR_ARG, "theta_split must be 0, 2 or 3");
  if (opt.flux_block < 0) return fail(QD_ERR_ARG, "flux_block < 0");
  h->dev = opt.device;
  h->prec = opt.precision;
  h->pipe = opt.pipeline;
  h->variant = opt.variant;
  h->tblock = opt.flux_block > 0 ? opt.flux_block : 64;
  h->use_graph = opt.use_graph;
  h->crp_every = opt.crp_every;
  h->profile = opt.profile;
  h->nsplit = opt.theta_split == 2 ? 2 : 3;
  if (opt.fuse_theta != 0 && opt.fuse_theta != -1) return fail(QD_ERR_ARG, "fuse_theta must be 0 or -1");
  h->fuse_opt = opt.fuse_theta;
  if (opt.diagnostics_every < 0) return fail(QD_ERR_ARG, "diagnostics_every < 0");
  h->diag_every = opt.diagnostics_every;
  return QD_OK;
}

// host-side checks of the flux parameters; fills the stencil
int setup_flux(qd_handle *h, const qd_problem *p) {
  h->do_crp = p->crp.calculate_CRP != 0;
  h->ne = h->do_crp ? p->crp.n_energies : 0;
  if (!h->do_crp) return QD_OK;
  if (!p->crp.energies || !p->crp.eta_sq) return fail(QD_ERR_NULL, "CRP arrays are NULL");
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
  return QD_OK;
}

// Upload a complex table given in FP64: as double2 for the FP64 path, as
// float4 (hi, lo) pairs for the mixed path (see tab_of in qd_kernels.hpp),
// hi rounded to keep each factor's modulus (qdt::round_unit_complex) and
// lo = the rest.
int put_table(qd_handle *h, void **dst, const std::vector<double> &v) {
  if (h->prec == QD_PREC_FP64) {
    double *d = nullptr;
    QD_RC(upload(&d, v.data(), v.size()));
    *dst = d;
  } else {
    const size_t n = v.size() / 2;
    std::vector<float> f(4 * n);
    for (size_t i = 0; i < n; i++) {
      float hx = 0.f, hy = 0.f;
      qdt::round_unit_complex(v[2 * i], v[2 * i + 1], hx, hy);
      f[4 * i] = hx;
      f[4 * i + 1] = hy;
      f[4 * i + 2] = (float)(v[2 * i] - (double)hx);
      f[4 * i + 3] = (float)(v[2 * i + 1] - (double)hy);
    }
    float *d = nullptr;
    QD_RC(upload(&d, f.data(), f.size()));
    *dst = d;
  }
  return QD_OK;
}

// Shear-form tables of the FP32 register passes along an axis of n points
// (rf_tab<float> in qd_kernels.hpp, rf_line in qd_rfft.hpp) from the phase
// angles of the kinetic factors (kin_ang[n]) and of the D/2 rows (ph_ang, rows
// of n): the twiddles W_n^e for |e| <= n/4; the kinetic factors, with the sign
// and 1/n in a scale, in the order stage B's thread b reads them (entry
// b + N2 k1 is frequency b' + N2 k1, b' the centred b); the D/2 rows as the
// constants (t, s, 2t) of half their angle.  Nothing for a length without a
// register pass.
int put_lift_tables(qd_handle *h, int n, const std::vector<double> &kin_ang, const std::vector<double> &ph_ang,
                    void **tw, void **kin, void **ph) {
  const int N1 = rf_n1(n);
  if (N1 == 0) return QD_OK;
  const int N2 = n / N1;
  const long double pi = 3.141592653589793238462643383279502884L;
  std::vector<float> ftw(2 * (size_t)(n / 2 + 1)), fkin(4 * (size_t)n), fph(4 * ph_ang.size());
  for (int e = -n / 4; e <= n / 4; e++)
    qdt::lift_consts((double)(-2 * pi * e / n), ftw[2 * (e + n / 4)], ftw[2 * (e + n / 4) + 1]);
  for (int b = 0; b < N2; b++)
    for (int k1 = 0; k1 < N1; k1++) {
      const int bc = b < (N2 + 1) / 2 ? b : b - N2;
      const int f = ((bc + N2 * k1) % n + n) % n;
      int sg = 1;
      const double th = qdt::lift_reduce(kin_ang[f], sg);
      float *o = &fkin[4 * (size_t)(b + N2 * k1)];
      qdt::lift_consts(th, o[0], o[1]);
      const double c = (double)sg / n;
      o[2] = (float)c;
      o[3] = (float)(c - (double)o[2]);
    }
  const long double twopi = 2 * pi;
  for (size_t i = 0; i < ph_ang.size(); i++) {
    float *o = &fph[4 * i];
    qdt::lift_consts((double)(remainderl((long double)ph_ang[i], twopi) / 2), o[0], o[1]);
    o[2] = 2 * o[0];
    o[3] = 0.f;
  }
  float *d = nullptr;
  QD_RC(upload(&d, ftw.data(), ftw.size()));
  *tw = d;
  QD_RC(upload(&d, fkin.data(), fkin.size()));
  *kin = d;
  QD_RC(upload(&d, fph.data(), fph.size()));
  *ph = d;
  return QD_OK;
}

// Parity halves of the theta matrices for k_theta_fused_par: per set
// [4][3][32 ksh][32 ksh] BF16 pieces (backward even, backward odd, forward
// even, forward odd; k contiguous).  False when a half has more than 32 ksh
// functions or points, or a set is not symmetric under x -> -x with FBR parity
// (-1)^c to 1e-10 of its largest entry (the kernel relies on it).
bool pack_theta_parity(const std::vector<MatSet> &mats, int ksh, std::vector<uint16_t> &out, long long &stride) {
  const int NP = 32 * ksh, KP = 32 * ksh;
  stride = 4LL * 3 * NP * KP;
  out.assign(mats.size() * (size_t)stride, 0);
  for (size_t s = 0; s < mats.size(); s++) {
    const MatSet &M = mats[s];
    const int nd = M.nd, nf = M.nf, NH = (nd + 1) / 2, CE = (nf + 1) / 2, CO = nf / 2;
    if (NH > NP || CE > KP) return false;
    double mf = 0, mb = 0;
    for (double v : M.Wf) mf = std::max(mf, std::fabs(v));
    for (double v : M.Pb) mb = std::max(mb, std::fabs(v));
    for (int c = 0; c < nf; c++)
      for (int i = 0; i < nd; i++) {
        const double sg = (c & 1) ? -1.0 : 1.0;
        if (std::fabs(M.Pb[(size_t)c * nd + nd - 1 - i] - sg * M.Pb[(size_t)c * nd + i]) > 1e-10 * mb) return false;
        if (std::fabs(M.Wf[(size_t)(nd - 1 - i) * nf + c] - sg * M.Wf[(size_t)i * nf + c]) > 1e-10 * mf) return false;
      }
    auto put = [&](int part, int n, int k, double v) {
      for (int q = 0; q < 3; q++) {
        const uint16_t b = bf16_bits((float)v);
        out[s * (size_t)stride + (((size_t)part * 3 + q) * NP + n) * KP + k] = b;
        v -= bf16_value(b);
      }
    };
    for (int i = 0; i < NH; i++) {
      const bool mid = nd % 2 == 1 && i == NH - 1;  // the middle node, its own mirror
      for (int c = 0; c < CE; c++) put(0, i, c, M.Pb[(size_t)(2 * c) * nd + i]);
      for (int c = 0; c < CO; c++) put(1, i, c, mid ? 0.0 : M.Pb[(size_t)(2 * c + 1) * nd + i]);
      for (int c = 0; c < CE; c++) put(2, c, i, (mid ? 0.5 : 1.0) * M.Wf[(size_t)i * nf + 2 * c]);
      for (int c = 0; c < CO; c++) put(3, c, i, mid ? 0.0 : M.Wf[(size_t)i * nf + 2 * c + 1]);
    }
  }
  return true;
}

// LDS limit and work split of the matrix-stationary theta kernel.  A work
// group loads its block's matrix once and then streams tpw 16-row tiles; tpw
// minimises (work-group rounds) x (tpw + matrix load), where a round is one
// wave of resident work groups over the whole chip.
int setup_theta_mx(qd_handle *h) {
  int ncu = 0;
  QD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->dev));
  int per_cu = 1;
  for (int pro = 0; pro <= 2; pro++)
    for (int epi = 0; epi <= 2; epi++) {
      const void *fn = theta_mx_kernel(h->tmx_ks, pro, epi);
      if (!fn) continue;
      QD_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->tmx_lds));
      int nb = 0;
      QD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, TMX_THREADS, h->tmx_lds));
      if (nb < 1) return fail(QD_ERR_DEVICE, "theta kernel does not fit on a compute unit");
      per_cu = pro + epi == 0 ? nb : std::min(per_cu, nb);
    }
  const long long slots = (long long)std::max(1, ncu) * per_cu;
  const long long ntiles = (h->n12 + 15) / 16;
  const long long nb = h->nb;
  double best = 1e300;
  int best_tpw = 1;
  const double load_cost = 0.25 * h->tmx_ks;  // matrix load in tile units
  for (long long tpw = 1; tpw <= ntiles; tpw++) {
    const long long wgs = nb * ((ntiles + tpw - 1) / tpw);
    const long long rounds = (wgs + slots - 1) / slots;
    const double cost = (double)rounds * ((double)std::max<long long>(tpw, TMX_THREADS / 64 / 2) + load_cost);
    if (cost < best - 1e-9) {
      best = cost;
      best_tpw = (int)tpw;
    }
  }
  h->tmx_tpw = best_tpw;
  return QD_OK;
}

// Bundle records of the FP32 mixing R pass (k_rpass_mx): units (group, r)
// packed into bundles of at most NL lines and two distinct unit matrices
// (qdp::pack_bundles); per bundle a header and NL-strided line and unit
// entries.  Groups of at most 32 lines go to 32-line bundles, larger groups to
// 48-line bundles: two classes, each with its own kernel, so that a batch with
// a few large groups still runs most of its lines on the 32-line kernel.  The
// matrices (shared by both classes) are stored padded to g4 = 4 ceil(g / 4),
// U then U^T, rounded to FP32 with unit column lengths
// (qdt::round_orthogonal_columns).
int setup_mx(qd_handle *h, const BatchSpec &S, const std::vector<int> &fcol_lam, int nlam) {
  const int n1 = h->n1, n2 = h->n2;
  std::vector<std::vector<int>> gcols(S.groups.size());
  std::vector<int> gu0(S.groups.size());
  for (size_t i = 0; i < S.groups.size(); i++) {
    gcols[i] = S.groups[i].cols;
    gu0[i] = S.groups[i].u0;
  }
  std::vector<float> upad;
  MxRecords rec[2];
  if (!build_mx_records(gcols, gu0, S.umat, fcol_lam, n1, n2, upad, rec)) return QD_OK;  // no mixing kernel
  int ncu = 0;
  QD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->dev));
  struct HostSet {
    size_t lds = 0;
    const void *fn = nullptr;
  } hs[2];
  for (int c = 0; c < 2; c++) {
    if (rec[c].heads.empty()) continue;
    hs[c].fn = rpass_mx_kernel(n1, rec[c].nl, rec[c].gb, false, &hs[c].lds);
    if (!hs[c].fn || hs[c].lds + (size_t)rec[c].nl * 12 > 160 * 1024) return QD_OK;  // no mixing kernel for this batch
    const std::string err = check_mx_records(rec[c], h->nlp, n1, nlam, upad.size());
    if (!err.empty()) return fail(QD_ERR_ARG, err);
  }
  // both classes fit and their records are in range: upload and size the launches
  QD_RC(upload(&h->mx_umat, upad.data(), upad.size()));
  for (int c = 0; c < 2; c++) {
    const HostSet &H = hs[c];
    const MxRecords &R = rec[c];
    MxSet &M = h->mxs[c];
    if (R.heads.empty()) continue;
    QD_RC(upload(&M.head, R.heads.data(), R.heads.size()));
    QD_RC(upload(&M.off, R.loff.data(), R.loff.size()));
    QD_RC(upload(&M.lam, R.llam.data(), R.llam.size()));
    QD_RC(upload(&M.unit, R.uni.data(), R.uni.size()));
    const int nthreads = R.nl * rf_n1(n1);
    int nb = 0;
    QD_HIP(hipFuncSetAttribute(H.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)H.lds));
    QD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, H.fn, nthreads, H.lds));
    if (nb < 1) return fail(QD_ERR_ARG, "the mixing R pass does not fit a CU");
    M.slots = (long long)nb * std::max(1, ncu);
    // the variant for two work groups per CU, where both fit
    M.slots2 = 0;
    size_t lds2 = 0;
    const void *fn2 = rpass_mx_kernel(n1, R.nl, R.gb, true, &lds2);
    if (fn2 && lds2 == H.lds) {
      int nb2 = 0;
      QD_HIP(hipFuncSetAttribute(fn2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)H.lds));
      QD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb2, fn2, nthreads, H.lds));
      if (nb2 >= 2) M.slots2 = (long long)nb2 * std::max(1, ncu);
    }
    M.nl = R.nl;
    M.gb = R.gb;
    M.lds = H.lds;
    M.nbundle = (int)R.heads.size();
  }
  h->mx_two = h->mxs[0].slots2 > 0;  // measured faster where it fits; the autotuner re-checks
  h->rp_mx = true;
  return QD_OK;
}

int create_batch(qd_handle *h, const BatchSpec &S) {
  const qd_problem *p = S.base;
  const int n1 = p->r1.n, n2 = p->r2.n;
  h->n1 = n1;
  h->n2 = n2;
  h->n12 = (size_t)n1 * n2;
  h->dt = p->time.time_step;
  h->dr1 = p->r1.dr;
  h->dr2 = p->r2.dr;
  h->mu2 = p->r2.mass;
  h->steps = p->time.steps;
  h->pending_prephase = p->time.steps == 0;
  h->has_dump = p->dump1 && p->dump2;
  h->has_mix = !S.umat.empty();
  QD_RC(setup_flux(h, p));
  if (h->pipe == QD_PIPE_PLAIN && (S.blocks.size() != 1 || h->has_mix))
    return fail(QD_ERR_ARG, "the plain pipeline propagates a single J = 0 block");

  // blocks and column tables
  h->nb = (int)S.blocks.size();
  int dcol = 0, fcol = 0, maxn = 0;
  std::vector<long long> dcol_vidx;
  std::vector<double> dcol_w;
  std::vector<long long> voff(S.vsets.size(), -1);
  long long vtot = 0;
  for (const HostBlock &b : S.blocks) {
    const MatSet &M = S.mats[b.mat];
    BlockDesc d;
    d.J = b.J;
    d.K = b.K;
    d.nd = M.nd;
    d.nf = M.nf;
    d.dcol0 = dcol;
    d.fcol0 = fcol;
    d.mat = b.mat;
    d.pad = 0;
    h->hblocks.push_back(d);
    if (voff[b.vset] < 0) {
      voff[b.vset] = vtot;
      vtot += (long long)h->n12 * M.nd;
      h->nV = vtot;
    }
    for (int k = 0; k < M.nd; k++) {
      dcol_vidx.push_back(voff[b.vset] + (long long)h->n12 * k);
      dcol_w.push_back(M.w[k]);
    }
    dcol += M.nd;
    fcol += M.nf;
    maxn = std::max(maxn, std::max(M.nd, M.nf));
  }
  h->ndcol = dcol;
  h->nfcol = fcol;
  h->npt = h->n12 * dcol;
  h->nlp = h->n12 * fcol;
  h->nx = (size_t)n1 * dcol;
  if ((int)S.fcol_j.size() != fcol || (int)S.fcol_lamval.size() != fcol)
    return fail(QD_ERR_ARG, "internal: column tables do not match the blocks");

  // FFT plans and launch shapes
  if (!factor_plan(n1, h->planR)) return fail(QD_ERR_ARG, "FFT length " + std::to_string(n1) + " has a prime factor > 5");
  if (!factor_plan(n2, h->planr)) return fail(QD_ERR_ARG, "FFT length " + std::to_string(n2) + " has a prime factor > 5");
  const size_t elem = h->prec == QD_PREC_FP64 ? sizeof(double2) : sizeof(float2);
  int gmax = 1;
  for (const Group &g : S.groups) gmax = std::max(gmax, (int)g.cols.size());
  // the FP32 mixing R pass with prefetch takes mixing groups of up to 48 lines
  // on 64-, 128- and 256-point R grids, the LDS-staged and Stockham R passes up
  // to 32; the split R pass (mixing in passes of its own around the register
  // FFT pass without mixing) any number, where that FFT kernel exists
  const bool mx_cand = h->prec == QD_PREC_MIXED && h->has_mix && (n1 == 64 || n1 == 128 || n1 == 256) && gmax <= 48;
  bool sp_cand = false;
  if (h->has_mix && rf_length(n1)) {
    const int lw = 256 / rf_n1(n1);
    size_t lds = 0;
    const void *fn = h->prec == QD_PREC_FP64 ? rpass_rf_kernel<double>(n1, false, lw, &lds)
                                             : rpass_rf_kernel<float>(n1, false, lw, &lds);
    sp_cand = fn && lds <= 160 * 1024 && n2 % lw == 0;
    h->sp_lw = lw;
  }
  if (gmax > RP_GMAX && !mx_cand && !sp_cand)
    return fail(QD_ERR_ARG, "mixing groups larger than 32 lines (J > 31) need an R grid of a length with a "
                            "register FFT kernel (" + std::to_string(n1) + " has none)");
  h->rp_legacy = gmax <= RP_GMAX;
  std::vector<int> bu0, ucols;
  std::vector<RUnit> units;
  const size_t tabsz = h->prec == QD_PREC_FP64 ? sizeof(double2) : sizeof(float4);
  if (!h->rp_legacy && sp_cand) {
    // the units alone (the split pass has no bundles)
    for (const Group &g : S.groups) {
      const int c0 = (int)ucols.size();
      for (int c : g.cols) ucols.push_back(c);
      for (int r = 0; r < n2; r++) units.push_back(RUnit{c0, (int)g.cols.size(), g.u0, r});
    }
  }
  if (h->rp_legacy) {
    // R pass: bundles of whole units; about 2048 elements per work group, at
    // least one mixing group, one line per thread slot.  Lengths with a
    // compile-time kernel use 8 or 32 lines per group.
    const int mb = max_butterflies(h->planR);
    int ml = 1;
    while (ml * 2 <= 64 && (size_t)ml * 2 * n1 <= 2048) ml *= 2;
    ml = std::max(ml, pow2ceil(gmax));
    bool use_ct = false;
    if (ct_length(n1)) {
      const int mlc = (gmax > 8 || n1 <= 64) ? 32 : 8;
      const bool fits = (h->prec == QD_PREC_FP64 ? rpass_kernel<double>(n1, mlc, h->has_mix) : rpass_kernel<float>(n1, mlc, h->has_mix)) != nullptr;
      if (fits && mlc >= gmax) {
        ml = mlc;
        use_ct = true;
      }
    }
    if (ml > 64) return fail(QD_ERR_ARG, "mixing groups too large for the R pass");
    const int tpl = RP_THREADS / ml;
    if (tpl * FFT_BPT < mb) return fail(QD_ERR_ARG, "R grid too large for the R pass with these mixing groups");
    const size_t pad_stride = (size_t)n1 + n1 / 32 + 1;
    const size_t lds = 2 * (size_t)n1 * tabsz + (size_t)ml * (use_ct ? pad_stride : (size_t)n1) * elem;
    if (lds > 160 * 1024) return fail(QD_ERR_ARG, "R grid too large for the R pass (LDS)");
    h->rp_ct = use_ct;
    h->rp_ml = ml;
    h->rp_tpl = tpl;
    int lines = 0, nu = 0;
    bu0.push_back(0);
    for (const Group &g : S.groups) {
      const int gs = (int)g.cols.size();
      const int c0 = (int)ucols.size();
      for (int c : g.cols) ucols.push_back(c);
      for (int r = 0; r < n2; r++) {
        if (lines + gs > ml || nu == 64) {
          bu0.push_back((int)units.size());
          lines = 0;
          nu = 0;
        }
        RUnit u;
        u.col0 = c0;
        u.g = gs;
        u.u0 = g.u0;
        u.r = r;
        units.push_back(u);
        lines += gs;
        nu++;
      }
    }
    bu0.push_back((int)units.size());
    h->nbundle = (int)bu0.size() - 1;
    h->rp_lds = lds;
  }
  {
    // r pass: lpb consecutive i of one column (lpb divides n1)
    const int mb = max_butterflies(h->planr);
    int lpb = 16;
    auto ok = [&](int l) {
      return n1 % l == 0 && (256 / l) * FFT_BPT >= mb && 3 * (size_t)n2 * tabsz + (size_t)l * n2 * elem <= 65536;
    };
    while (lpb > 1 && !ok(lpb)) lpb /= 2;
    if (!ok(lpb)) return fail(QD_ERR_ARG, "FFT length " + std::to_string(n2) + " too large");
    h->rl_lpb = lpb;
    h->rl_ct = false;
    if (ct_length(n2)) {
      for (int l : {16, 8}) {
        const size_t ldsc = 3 * (size_t)n2 * tabsz + (size_t)n2 * (l + 1) * elem;
        const bool fits = (h->prec == QD_PREC_FP64 ? rline_kernel<double>(n2, l) : rline_kernel<float>(n2, l)) != nullptr;
        if (fits && n1 % l == 0 && ldsc <= 160 * 1024) {
          h->rl_lpb = l;
          h->rl_ct = true;
          break;
        }
      }
    }
    h->rl_blocks = (int)(((size_t)n1 * fcol + h->rl_lpb - 1) / h->rl_lpb);
    h->rl_lds = h->rl_ct ? 3 * (size_t)n2 * tabsz + (size_t)n2 * (h->rl_lpb + 1) * elem
                         : 3 * (size_t)n2 * tabsz + (size_t)h->rl_lpb * n2 * elem;
  }
  // register-FFT passes for the fused step (other flag sets keep the
  // Stockham kernels above)
  h->rp_rf = h->rl_rf = false;
  if (rf_length(n1) && h->rp_legacy) {
    const int n1f = rf_n1(n1);
    int lw = 256 / n1f;
    if (h->has_mix) lw = std::max(lw, h->rp_ml);  // a bundle holds up to rp_ml lines
    size_t lds = 0;
    const void *fn = h->prec == QD_PREC_FP64 ? rpass_rf_kernel<double>(n1, h->has_mix, lw, &lds)
                                             : rpass_rf_kernel<float>(n1, h->has_mix, lw, &lds);
    // without mixing a batch of lw lines must lie in one column
    // the mixing kernel also holds 2 x lw x 40 unit-matrix entries statically
    const size_t lds_static = h->has_mix ? 2 * (size_t)lw * 40 * (h->prec == QD_PREC_FP64 ? 8 : 4) + 1024 : 0;
    if (fn && lds + lds_static <= 160 * 1024 && lw * n1f <= 1024 && (h->has_mix || n2 % lw == 0)) {
      h->rp_rf = true;
      h->rp_lw = lw;
    }
  }
  if (rf_length(n2) && n1 % (256 / rf_n1(n2)) == 0) {
    size_t lds = 0;
    int lpb = 0;
    const void *fn = h->prec == QD_PREC_FP64 ? rline_rf_kernel<double>(n2, 256, &lds, &lpb)
                                             : rline_rf_kernel<float>(n2, 256, &lds, &lpb);
    h->rl_rf = fn && lds <= 160 * 1024;
  }
  // theta matrices: padded shapes shared by all sets
  h->ncb = (h->prec == QD_PREC_MIXED && maxn <= 64) ? 4 : 8;
  h->f64_cb = maxn <= 64 ? 4 : 8;  // FP64 theta: 64 or 128 columns per work group
  const int colw = h->prec == QD_PREC_FP64 ? 16 * h->f64_cb : h->ncb * 16;
  h->mat_Npad = (maxn + colw - 1) / colw * colw;
  h->col_tiles = h->mat_Npad / colw;
  const int kc = h->prec == QD_PREC_FP64 ? TG_KC64 : TG_KC16;
  h->mat_Kpad = (maxn + kc - 1) / kc * kc;
  // the matrix-stationary kernel holds a whole (32 ks)^2 matrix in LDS and
  // addresses one block's columns with 32-bit byte offsets
  h->tmx_ks = 0;
  if (h->prec == QD_PREC_MIXED && h->nsplit == 3 && maxn <= 128 &&
      (double)((maxn + 31) / 32 * 32) * (double)h->n12 * 8.0 < 2147483648.0) {
    h->tmx_ks = (maxn + 31) / 32;
    h->mat_Npad = h->mat_Kpad = 32 * h->tmx_ks;
    h->tmx_lds = (size_t)3 * 32 * h->tmx_ks * (32 * h->tmx_ks + 8) * 2;
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
  if (h->rp_lds > 65536) {
    const void *fn = h->rp_ct ? (h->prec == QD_PREC_FP64 ? rpass_kernel<double>(n1, h->rp_ml, h->has_mix) : rpass_kernel<float>(n1, h->rp_ml, h->has_mix))
                              : (h->prec == QD_PREC_FP64 ? (const void *)k_rpass<double> : (const void *)k_rpass<float>);
    QD_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->rp_lds));
  }
  if (h->rl_lds > 65536) {
    const void *fn = h->prec == QD_PREC_FP64 ? rline_kernel<double>(n2, h->rl_lpb) : rline_kernel<float>(n2, h->rl_lpb);
    QD_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->rl_lds));
  }
  if (h->tmx_ks) QD_RC(setup_theta_mx(h));
  int ncu = 0;
  QD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->dev));
  if (h->rp_rf) {
    size_t lds = 0;
    const void *fn = h->prec == QD_PREC_FP64 ? rpass_rf_kernel<double>(n1, h->has_mix, h->rp_lw, &lds)
                                             : rpass_rf_kernel<float>(n1, h->has_mix, h->rp_lw, &lds);
    QD_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int nb = 0;
    QD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, h->rp_lw * rf_n1(n1), lds));
    h->rp_slots = std::max(1, nb) * std::max(1, ncu);
  }
  if (h->rl_rf) {
    size_t lds = 0;
    int lpb = 0;
    const void *fn = h->prec == QD_PREC_FP64 ? rline_rf_kernel<double>(n2, 256, &lds, &lpb)
                                             : rline_rf_kernel<float>(n2, 256, &lds, &lpb);
    QD_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int nb = 0;
    QD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 256, lds));
    h->rl_slots = std::max(1, nb) * std::max(1, ncu);
    // the 512-thread variant (batches of twice the consecutive R points), where it fits
    const void *fn2 = h->prec == QD_PREC_FP64 ? rline_rf_kernel<double>(n2, 512, &lds, &lpb)
                                              : rline_rf_kernel<float>(n2, 512, &lds, &lpb);
    h->rl_slots2 = 0;
    if (fn2 && lds <= 160 * 1024 && n1 % lpb == 0) {
      QD_HIP(hipFuncSetAttribute(fn2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      int nb2 = 0;
      QD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb2, fn2, 512, lds));
      if (nb2 >= 1) h->rl_slots2 = (long long)nb2 * std::max(1, ncu);
    }
  }
  if (sp_cand && !units.empty()) {
    size_t lds = 0;
    const void *fn = h->prec == QD_PREC_FP64 ? rpass_rf_kernel<double>(n1, false, h->sp_lw, &lds)
                                             : rpass_rf_kernel<float>(n1, false, h->sp_lw, &lds);
    QD_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int nb = 0;
    QD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, h->sp_lw * rf_n1(n1), lds));
    h->sp_slots = std::max(1, nb) * std::max(1, ncu);
    h->sp_nunits = (int)units.size();
    h->sp_wpu = (n1 + 63) / 64;
    h->sp_ok = true;
    // the default where no register-FFT mixing pass in one launch takes these
    // groups (the Stockham mixing pass is slower: 40.0 against 11.7 ms on the
    // FP64 J = 0..20 batch)
    h->rp_split = !mx_cand && (!h->rp_legacy || !h->rp_rf);
  }
  QD_RC(upload(&h->d_blocks, h->hblocks.data(), h->hblocks.size()));
  {
    std::vector<BlockDesc> byk = h->hblocks;
    std::stable_sort(byk.begin(), byk.end(), [](const BlockDesc &a, const BlockDesc &b) { return a.K < b.K; });
    QD_RC(upload(&h->d_blocks_k, byk.data(), byk.size()));
    // runs of blocks sharing the theta matrices, the basis sizes and the V columns
    h->tfu_groups.clear();
    h->tfu_blocks.clear();
    for (const BlockDesc &d : byk) h->tfu_blocks.push_back({d.nd, d.nf, d.dcol0, d.fcol0, d.mat, dcol_vidx[d.dcol0]});
    for (size_t i = 0; i < byk.size();) {
      size_t j = i + 1;
      while (j < byk.size() && byk[j].mat == byk[i].mat && byk[j].nd == byk[i].nd && byk[j].nf == byk[i].nf &&
             dcol_vidx[byk[j].dcol0] == dcol_vidx[byk[i].dcol0])
        j++;
      h->tfu_groups.push_back({(int)i, (int)(j - i)});
      i = j;
    }
  }
  QD_RC(upload(&h->dcol_vidx, dcol_vidx.data(), dcol_vidx.size()));
  QD_RC(upload(&h->dcol_w, dcol_w.data(), dcol_w.size()));
  QD_RC(upload(&h->bundle_u0, bu0.data(), bu0.size()));
  QD_RC(upload(&h->units, units.data(), units.size()));
  QD_RC(upload(&h->unit_cols, ucols.data(), ucols.size()));
  if (h->prec == QD_PREC_FP64) {
    double *u = nullptr;
    QD_RC(upload(&u, S.umat.data(), S.umat.size()));
    h->umat = u;
  } else {
    std::vector<float> uf(S.umat.size());
    for (const Group &g : S.groups) {
      if (g.u0 < 0) continue;
      const int gs = (int)g.cols.size();
      std::vector<float> q;
      qdt::round_orthogonal_columns(S.umat.data() + g.u0, gs, q);
      std::copy(q.begin(), q.end(), uf.begin() + g.u0);
    }
    float *u = nullptr;
    QD_RC(upload(&u, uf.data(), uf.size()));
    h->umat = u;
  }

  // psi: blocks back to back
  if (h->prec == QD_PREC_FP64) {
    QD_HIP(hipMalloc((void **)&h->psi64, sizeof(double2) * h->npt));
    QD_HIP(hipMalloc((void **)&h->lpsi64, sizeof(double2) * h->nlp));
  } else {
    QD_HIP(hipMalloc((void **)&h->stage, sizeof(double2) * std::min(STAGE_POINTS, std::max(h->npt, h->nlp))));
    QD_HIP(hipMalloc((void **)&h->psi32, sizeof(float2) * h->npt));
    QD_HIP(hipMalloc((void **)&h->lpsi32, sizeof(float2) * h->nlp));
  }
  // the split R pass mixes into psi, which no step reads across the R pass;
  // a DVR array smaller than the FBR one takes a scratch array instead
  if (h->sp_ok && h->npt < h->nlp) QD_HIP(hipMalloc(&h->sp_tmp, elem * h->nlp));
  for (int b = 0; b < h->nb; b++) {
    const BlockDesc &d = h->hblocks[b];
    QD_RC(put_complex(h, h->prec == QD_PREC_FP64 ? (void *)h->psi64 : (void *)h->psi32, h->n12 * d.dcol0,
                      S.blocks[b].psi, h->n12 * d.nd));
  }
  // potential sets (FP64; the mixed path adds an FP32 exp(-i V dt) table)
  {
    std::vector<double> vh((size_t)vtot);
    for (size_t s = 0; s < S.vsets.size(); s++) {
      if (voff[s] < 0) continue;
      int nd = 0;
      for (const HostBlock &b : S.blocks)
        if (b.vset == (int)s) nd = S.mats[b.mat].nd;
      memcpy(vh.data() + voff[s], S.vsets[s], sizeof(double) * h->n12 * nd);
    }
    QD_RC(upload(&h->V, vh.data(), vh.size()));
    if (h->prec == QD_PREC_MIXED) {
      std::vector<float> e(2 * vh.size());
      for (size_t x = 0; x < vh.size(); x++) {
        const double a = -h->dt * vh[x];
        qdt::round_unit_complex(cos(a), sin(a), e[2 * x], e[2 * x + 1]);
      }
      float *d = nullptr;
      QD_RC(upload(&d, e.data(), e.size()));
      h->expV32 = reinterpret_cast<float2 *>(d);
    }
  }
  if (h->has_dump) {
    QD_RC(upload(&h->d1, p->dump1, n1));
    QD_RC(upload(&h->d2, p->dump2, n2));
    std::vector<float> f1(p->dump1, p->dump1 + n1), f2(p->dump2, p->dump2 + n2);
    QD_RC(upload(&h->d1f, f1.data(), f1.size()));
    QD_RC(upload(&h->d2f, f2.data(), f2.size()));
  }

  // radial tables (MatlabStructures.C:17-36, fftwinterface.C:110-125)
  const bool cuda = h->variant == QD_VARIANT_CUDA;
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
    const double r = cuda ? p->r1.r[0] + i * p->r1.dr : p->r1.r[i];  // cumath.h:153-160
    I1[i] = 1.0 / ((p->r1.mass + p->r1.mass) * r * r);
  }
  for (int j = 0; j < n2; j++) {
    const double r = cuda ? p->r2.r[0] + j * p->r2.dr : p->r2.r[j];
    I2[j] = 1.0 / ((p->r2.mass + p->r2.mass) * r * r);
  }
  const double dt = h->dt;
  // D_R/2 per distinct lambda, D_r/2 per j
  std::map<long long, int> lam_index;
  std::vector<long long> lams;
  std::vector<int> fcol_lam(fcol);
  for (int c = 0; c < fcol; c++) {
    auto it = lam_index.find(S.fcol_lamval[c]);
    if (it == lam_index.end()) {
      it = lam_index.emplace(S.fcol_lamval[c], (int)lams.size()).first;
      lams.push_back(S.fcol_lamval[c]);
    }
    fcol_lam[c] = it->second;
  }
  const int njr = S.jmax + 1;
  std::vector<double> dRt(2 * lams.size() * n1), drt(2 * (size_t)njr * n2), kinR(2 * n1), kinr(2 * n2), twR(2 * n1),
      twr(2 * n2);
  std::vector<double> dRa(lams.size() * n1), dra((size_t)njr * n2), kRa(n1), kra(n2);  // the phase angles
  for (size_t q = 0; q < lams.size(); q++)
    for (int i = 0; i < n1; i++) {
      const double a = (-dt / 2) * ((double)lams[q] * I1[i]);
      dRa[q * n1 + i] = a;
      dRt[2 * (q * n1 + i)] = cos(a);
      dRt[2 * (q * n1 + i) + 1] = sin(a);
    }
  for (int j = 0; j < njr; j++)
    for (int r = 0; r < n2; r++) {
      const double a = (-dt / 2) * ((double)j * (j + 1) * I2[r]);
      dra[(size_t)j * n2 + r] = a;
      drt[2 * ((size_t)j * n2 + r)] = cos(a);
      drt[2 * ((size_t)j * n2 + r) + 1] = sin(a);
    }
  for (int i = 0; i < n1; i++) {
    const double T = p1v[i] * p1v[i] / (2 * p->r1.mass);
    kRa[i] = -dt * T;
    kinR[2 * i] = cos(-dt * T) / n1;
    kinR[2 * i + 1] = sin(-dt * T) / n1;
    const long double ang = -2.0L * 3.141592653589793238462643383279502884L * i / n1;
    twR[2 * i] = (double)cosl(ang);
    twR[2 * i + 1] = (double)sinl(ang);
  }
  for (int j = 0; j < n2; j++) {
    const double T = p2v[j] * p2v[j] / (2 * p->r2.mass);
    kra[j] = -dt * T;
    kinr[2 * j] = cos(-dt * T) / n2;
    kinr[2 * j + 1] = sin(-dt * T) / n2;
    const long double ang = -2.0L * 3.141592653589793238462643383279502884L * j / n2;
    twr[2 * j] = (double)cosl(ang);
    twr[2 * j + 1] = (double)sinl(ang);
  }
  QD_RC(put_table(h, &h->dR, dRt));
  QD_RC(put_table(h, &h->dr, drt));
  QD_RC(put_table(h, &h->kinR, kinR));
  QD_RC(put_table(h, &h->kinr, kinr));
  QD_RC(put_table(h, &h->twR, twR));
  QD_RC(put_table(h, &h->twr, twr));
  if (h->prec == QD_PREC_MIXED) {
    QD_RC(put_lift_tables(h, n1, kRa, dRa, &h->ltwR, &h->lkinR, &h->lphR));
    QD_RC(put_lift_tables(h, n2, kra, dra, &h->ltwr, &h->lkinr, &h->lphr));
  }
  // diagnostics tables: a single J = 0 grid, FBR column c is Legendre degree l = c
  h->diag_ok = S.blocks.size() == 1 && S.blocks[0].J == 0;
  if (h->diag_every > 0 && !h->diag_ok)
    return fail(QD_ERR_ARG, "diagnostics_every needs a single J = 0 grid (qd_create)");
  if (h->diag_ok) {
    std::vector<double> t;
    for (int i = 0; i < n1; i++) t.push_back(p1v[i] * p1v[i] / (2 * p->r1.mass));
    for (int j = 0; j < n2; j++) t.push_back(p2v[j] * p2v[j] / (2 * p->r2.mass));
    t.insert(t.end(), I1.begin(), I1.end());
    t.insert(t.end(), I2.begin(), I2.end());
    for (int c = 0; c < fcol; c++) t.push_back(1.0 / (c + 0.5));
    for (int c = 0; c < fcol; c++) t.push_back((double)c * (c + 1) / (c + 0.5));
    QD_RC(upload(&h->diag_tab, t.data(), t.size()));
  }
  QD_RC(upload(&h->fcol_lam, fcol_lam.data(), fcol_lam.size()));
  if (mx_cand) QD_RC(setup_mx(h, S, fcol_lam, (int)lams.size()));
  if (h->rp_rf && h->has_mix) {
    // per-bundle line tables of the mixing R pass: element offset and D_R row per line
    std::vector<long long> loff;
    std::vector<int> llam, bl0;
    for (int b = 0; b < h->nbundle; b++) {
      bl0.push_back((int)loff.size());
      for (int u = bu0[b]; u < bu0[b + 1]; u++)
        for (int k = 0; k < units[u].g; k++) {
          const int col = ucols[units[u].col0 + k];
          loff.push_back(((long long)col * n2 + units[u].r) * n1);
          llam.push_back(fcol_lam[col]);
        }
    }
    bl0.push_back((int)loff.size());
    QD_RC(upload(&h->bundle_l0, bl0.data(), bl0.size()));
    QD_RC(upload(&h->line_off, loff.data(), loff.size()));
    QD_RC(upload(&h->line_lam, llam.data(), llam.size()));
  }
  QD_RC(upload(&h->fcol_j, S.fcol_j.data(), S.fcol_j.size()));

  // theta matrices, packed per set
  const int nmat = (int)S.mats.size();
  const int Kpad = h->mat_Kpad, Npad = h->mat_Npad;
  if (h->prec == QD_PREC_FP64) {
    h->mat_stride = (long long)Kpad * Npad;
    std::vector<double> F((size_t)nmat * h->mat_stride, 0.0), B((size_t)nmat * h->mat_stride, 0.0);
    for (int s = 0; s < nmat; s++) {
      const MatSet &M = S.mats[s];
      double *f = F.data() + (size_t)s * h->mat_stride, *bq = B.data() + (size_t)s * h->mat_stride;
      for (int k = 0; k < M.nd; k++)
        for (int j = 0; j < M.nf; j++) f[(size_t)k * Npad + j] = M.Wf[(size_t)k * M.nf + j];
      for (int j = 0; j < M.nf; j++)
        for (int k = 0; k < M.nd; k++) bq[(size_t)j * Npad + k] = M.Pb[(size_t)j * M.nd + k];
    }
    double *df = nullptr, *db = nullptr;
    QD_RC(upload(&df, F.data(), F.size()));
    QD_RC(upload(&db, B.data(), B.size()));
    h->Wf = df;
    h->Wb = db;
  } else {
    const int NS = 3;  // pieces stored; the two-piece kernel reads the first two
    h->mat_stride = (long long)NS * Npad * Kpad;
    std::vector<uint16_t> F((size_t)nmat * h->mat_stride, 0), B((size_t)nmat * h->mat_stride, 0);
    auto split = [&](std::vector<uint16_t> &dst, int s, int K, int N, const std::vector<double> &src, bool transposed) {
      for (int k = 0; k < K; k++)
        for (int n = 0; n < N; n++) {
          double r = transposed ? src[(size_t)k * N + n] : src[(size_t)k * N + n];
          for (int q = 0; q < NS; q++) {
            const uint16_t b = bf16_bits((float)r);
            dst[(size_t)s * h->mat_stride + ((size_t)q * Npad + n) * Kpad + k] = b;
            r -= bf16_value(b);
          }
        }
    };
    for (int s = 0; s < nmat; s++) {
      const MatSet &M = S.mats[s];
      split(F, s, M.nd, M.nf, M.Wf, false);  // B of the forward GEMM: [k][j]
      split(B, s, M.nf, M.nd, M.Pb, false);  // B of the backward GEMM: [j][k]
    }
    void *df = nullptr, *db = nullptr;
    QD_HIP(hipMalloc(&df, F.size() * 2));
    QD_HIP(hipMemcpy(df, F.data(), F.size() * 2, hipMemcpyHostToDevice));
    QD_HIP(hipMalloc(&db, B.size() * 2));
    QD_HIP(hipMemcpy(db, B.data(), B.size() * 2, hipMemcpyHostToDevice));
    h->Wf = df;
    h->Wb = db;
    // fused theta step, either propagator order, three pieces, both matrices
    // in LDS: the dense transforms (bases up to 64) or their parity halves
    // (halves up to 64 functions: bases up to 128), the halves by default
    if (h->fuse_opt == 0 && h->pipe == QD_PIPE_FUSED && h->nsplit == 3 && h->tmx_ks >= 1) {
      auto fits = [&](const void *fn, int nth, size_t lds) -> bool {
        if (!fn) return false;
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) return false;
        int nbk = 0;
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbk, fn, nth, lds) == hipSuccess && nbk >= 1;
      };
      size_t flds = 0;
      h->dense_ok = h->tmx_ks <= 2 && fits(theta_fused_kernel(h->tmx_ks, TMX_THREADS, &flds), TMX_THREADS, flds);
      const int ksh = ((maxn + 1) / 2 + 31) / 32;
      const int pnth = ksh == 1 ? TMX_THREADS : 256;
      size_t plds = 0;
      std::vector<uint16_t> Q;
      if (ksh <= 2 && fits(theta_fused_par_kernel(ksh, pnth, &plds), pnth, plds) &&
          pack_theta_parity(S.mats, ksh, Q, h->par_stride)) {
        QD_HIP(hipMalloc(&h->Wpar, Q.size() * 2));
        QD_HIP(hipMemcpy(h->Wpar, Q.data(), Q.size() * 2, hipMemcpyHostToDevice));
        if (ksh == 1) (void)fits(theta_fused_par_kernel(1, 768, &plds), 768, plds);
        h->par_ok = true;
        h->par_ksh = ksh;
        // (both forms of the wave-pair kernel: either meeting may be selected)
        h->pair_ok = ksh == 2 && fits(theta_fused_pair_kernel(512, false, &plds), 512, plds) &&
                     fits(theta_fused_pair_kernel(512, true, &plds), 512, plds);
      }
      if (h->dense_ok || h->par_ok) {
        // default: the parity halves, as wave pairs where they fit (half
        // bases of 33..64: 93.4 against 96.6 us per step on a 256 x 192 x 120
        // grid)
        h->fuse_theta = true;
        h->tfu_par = h->par_ok;
        h->tfu_pair = h->pair_ok;
        h->tfu_nth = h->pair_ok ? 512 : h->par_ok ? pnth : TMX_THREADS;
        h->tfu_tpw = h->tmx_tpw;
        QD_RC(build_tfu_items(h));
      }
    }
  }

  if (h->do_crp) {
    QD_RC(upload(&h->energies, p->crp.energies, h->ne));
    {
      // |eta(E)|^2 per block, [nb][ne]: a block's own array, else the problem's
      std::vector<double> eta((size_t)h->nb * h->ne);
      for (int b = 0; b < h->nb; b++) {
        const double *src = S.blocks[b].eta_sq ? S.blocks[b].eta_sq : p->crp.eta_sq;
        std::copy(src, src + h->ne, eta.begin() + (size_t)b * h->ne);
      }
      QD_RC(upload(&h->eta_sq, eta.data(), eta.size()));
    }
    QD_HIP(hipMalloc((void **)&h->crp, sizeof(double) * h->ne * h->nb));
    QD_HIP(hipMalloc((void **)&h->d_acc_steps, sizeof(long long)));
    {
      const long long z = 0;
      QD_HIP(hipMemcpy(h->d_acc_steps, &z, sizeof(z), hipMemcpyHostToDevice));
    }
    QD_HIP(hipMalloc((void **)&h->bpsi, sizeof(double2) * h->nx * h->tblock));
    QD_HIP(hipMalloc((void **)&h->bdpsi, sizeof(double2) * h->nx * h->tblock));
    QD_HIP(hipMalloc((void **)&h->ftau, sizeof(double2) * h->ne * h->tblock));
    QD_HIP(hipMalloc((void **)&h->fai, sizeof(double2) * h->nx * h->ne));
    QD_HIP(hipMalloc((void **)&h->dfai, sizeof(double2) * h->nx * h->ne));
    QD_HIP(hipMemsetAsync(h->fai, 0, sizeof(double2) * h->nx * h->ne, h->stream));
    QD_HIP(hipMemsetAsync(h->dfai, 0, sizeof(double2) * h->nx * h->ne, h->stream));
    // P(E) from the folds' partial sums where every block starts on a 128-value x range
    h->part_ok = true;
    for (const BlockDesc &d : h->hblocks)
      if (((long long)n1 * d.dcol0) % FX_X) h->part_ok = false;
    if (h->part_ok) {
      const size_t nparts = (h->nx + FX_X - 1) / FX_X * (size_t)h->ne;
      QD_HIP(hipMalloc((void **)&h->crp_part, sizeof(double) * nparts));
      QD_HIP(hipMemsetAsync(h->crp_part, 0, sizeof(double) * nparts, h->stream));
    }
  }
  QD_HIP(hipMalloc((void **)&h->partial, sizeof(double) * 4 * h->npartial));  // 4 diagnostic slots
  QD_HIP(hipStreamSynchronize(h->stream));
  return QD_OK;
}

// The reference's J = 0 problem as a batch of one block.
int spec_from_problem(const qd_problem *p, BatchSpec &S) {
  QD_RC(validate_base(p, true));
  const int nth = p->theta.n, M = p->theta.m + 1;
  S.base = p;
  S.jmax = M - 1;
  MatSet ms;
  ms.nd = nth;
  ms.nf = M;
  ms.Wf.resize((size_t)nth * M);
  ms.Pb.resize((size_t)M * nth);
  ms.w.assign(p->theta.w, p->theta.w + nth);
  const double *leg = p->theta.legendre;
  for (int l = 0; l < M; l++) {
    const double f = l + 0.5;  // setup_weight_legendre, timeEvol.C:737-760
    for (int k = 0; k < nth; k++) {
      ms.Wf[(size_t)k * M + l] = f * p->theta.w[k] * leg[l + (size_t)M * k];
      ms.Pb[(size_t)l * nth + k] = leg[l + (size_t)M * k];
    }
  }
  S.mats.push_back(ms);
  S.vsets.push_back(p->pot);
  HostBlock b;
  b.J = 0;
  b.K = 0;
  b.mat = 0;
  b.vset = 0;
  b.psi = p->psi;
  S.blocks.push_back(b);
  for (int l = 0; l < M; l++) {
    S.fcol_lamval.push_back((long long)l * (l + 1));
    S.fcol_j.push_back(l);
    Group g;
    g.cols.push_back(l);
    S.groups.push_back(g);
  }
  return QD_OK;
}

int spec_from_jk(const qd_jk_problem *q, BatchSpec &S) {
  if (!q) return fail(QD_ERR_NULL, "problem is NULL");
  QD_RC(validate_base(&q->base, false));
  if (q->jmax < 0 || q->n_blocks <= 0 || !q->blocks) return fail(QD_ERR_ARG, "jmax and blocks are required");
  S.base = &q->base;
  S.jmax = q->jmax;
  const int jmax = q->jmax;
  // theta sets: one per K (the J = 0 block reuses the reference Legendre matrices)
  std::map<int, int> mat_of_key, vset_of_K;
  for (int b = 0; b < q->n_blocks; b++) {
    const qd_jk_block &B = q->blocks[b];
    if (B.J < 0 || B.K < 0 || B.K > B.J || B.K > jmax) return fail(QD_ERR_ARG, "block needs 0 <= K <= min(J, jmax)");
    if (!B.pot || !B.psi) return fail(QD_ERR_NULL, "block pot or psi is NULL");
    const int key = B.J == 0 ? -1 : B.K;
    if (!mat_of_key.count(key)) {
      MatSet ms;
      std::vector<double> x, w, Pb;
      qdt::assoc_legendre_dvr(B.K, jmax, x, w, Pb);
      const int n = jmax - B.K + 1;
      ms.nd = ms.nf = n;
      ms.w = w;
      ms.Wf.resize((size_t)n * n);
      ms.Pb.resize((size_t)n * n);
      if (key == -1) {  // reference J = 0 matrices on the same Gauss-Legendre grid
        std::vector<double> L = qdt::legendre_p2(jmax, x);
        for (int l = 0; l < n; l++)
          for (int k = 0; k < n; k++) {
            ms.Wf[(size_t)k * n + l] = (l + 0.5) * w[k] * L[(size_t)l * n + k];
            ms.Pb[(size_t)l * n + k] = L[(size_t)l * n + k];
          }
      } else {
        for (int j = 0; j < n; j++)
          for (int k = 0; k < n; k++) {
            ms.Wf[(size_t)k * n + j] = w[k] * Pb[(size_t)j * n + k];
            ms.Pb[(size_t)j * n + k] = Pb[(size_t)j * n + k];
          }
      }
      mat_of_key[key] = (int)S.mats.size();
      S.mats.push_back(ms);
    }
    if (!vset_of_K.count(B.K)) {
      vset_of_K[B.K] = (int)S.vsets.size();
      S.vsets.push_back(B.pot);
    }
    HostBlock hb;
    hb.J = B.J;
    hb.K = B.K;
    hb.mat = mat_of_key[key];
    hb.vset = vset_of_K[B.K];
    hb.psi = B.psi;
    hb.eta_sq = B.eta_sq;
    S.blocks.push_back(hb);
  }
  // FBR columns: block (J, K), local index a <-> j = K + a
  std::map<std::pair<int, int>, int> fcol0;  // (J, K) -> first FBR column
  int fc = 0;
  for (const HostBlock &hb : S.blocks) {
    if (fcol0.count({hb.J, hb.K})) return fail(QD_ERR_ARG, "duplicate (J, K) block");
    fcol0[{hb.J, hb.K}] = fc;
    fc += S.mats[hb.mat].nf;
  }
  S.fcol_lamval.assign(fc, 0);
  S.fcol_j.assign(fc, 0);
  std::map<int, std::vector<int>> Ks_of_J;
  for (const HostBlock &hb : S.blocks) Ks_of_J[hb.J].push_back(hb.K);
  for (auto &kv : Ks_of_J) {
    const int J = kv.first;
    std::vector<int> Ks = kv.second;
    std::sort(Ks.begin(), Ks.end());
    for (int j = 0; j <= jmax; j++) {
      const int g = std::min(J, j) + 1;
      // the K blocks of this J must be 0..min(J, jmax) for the coupling to close
      Group grp;
      for (int K = 0; K < g; K++) {
        auto it = fcol0.find({J, K});
        if (it == fcol0.end()) return fail(QD_ERR_ARG, "J > 0 needs every block K = 0..min(J, jmax)");
        grp.cols.push_back(it->second + (j - K));
      }
      if (q->coriolis && g > 1) {
        std::vector<double> U;
        std::vector<int> l;
        qdt::coriolis_eigen(J, j, U, l);
        grp.u0 = (int)S.umat.size();
        S.umat.insert(S.umat.end(), U.begin(), U.end());
        for (int m = 0; m < g; m++) {
          S.fcol_lamval[grp.cols[m]] = (long long)l[m] * (l[m] + 1);
          S.fcol_j[grp.cols[m]] = j;
        }
      } else {
        // J = 0, or the centrifugal-sudden limit: (J - j)^2 -> J(J+1) + j(j+1) - 2K^2
        for (int K = 0; K < g; K++) {
          S.fcol_lamval[grp.cols[K]] = (long long)J * (J + 1) + (long long)j * (j + 1) - 2LL * K * K;
          S.fcol_j[grp.cols[K]] = j;
        }
        grp.u0 = -1;
      }
      if (g > 1 && q->coriolis) {
        S.groups.push_back(grp);
      } else {
        for (int K = 0; K < g; K++) {
          Group one;
          one.cols.push_back(grp.cols[K]);
          S.groups.push_back(one);
        }
      }
    }
  }
  return QD_OK;
}

// Drops the captured flux block: every change of a launch choice after the
// capture (qd_debug_select) must call it, or replays would keep running the
// kernels captured before the change.
void drop_graph(qd_handle *h) {
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
}

void release(qd_handle *h) {
  if (!h) return;
  void *bufs[] = {h->psi64,    h->lpsi64, h->psi32,     h->lpsi32,   h->V,      h->expV32,   h->dcol_vidx,
                  h->dcol_w,   h->fcol_lam, h->fcol_j,  h->d_blocks, h->d_blocks_k, h->d1,   h->d2,       h->d1f,
                  h->d2f,      h->dR,     h->dr,        h->kinR,     h->kinr,   h->twR,      h->twr,
                  h->ltwR,     h->ltwr,   h->lkinR,     h->lkinr,    h->lphR,   h->lphr,
                  h->Wf,       h->Wb,     h->bundle_u0, h->units,    h->unit_cols, h->umat,  h->bpsi, h->ftau,
                  h->bdpsi,    h->fai,    h->dfai,      h->energies, h->eta_sq, h->crp,      h->partial,
                  h->bundle_l0, h->line_off, h->line_lam, h->d_acc_steps, h->mxs[0].head, h->mxs[0].off,
                  h->mxs[0].lam, h->mxs[0].unit, h->mxs[1].head, h->mxs[1].off, h->mxs[1].lam, h->mxs[1].unit,
                  h->mx_umat,  h->d_tfu_items, h->stage, h->diag_tab, h->crp_part, h->sp_tmp,
                  h->Wpar};
  drop_graph(h);
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (auto &e : h->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : h->marks)
    if (e) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
}

// Captures tblock steps and the flux accumulation that closes the block.  The
// host-side block state advanced by the capture is put back: replays advance
// it (propagate_impl) and the device counter (k_counter_add).
int capture_block(qd_handle *h) {
  const int slot0 = h->slot;
  const long long acc0 = h->acc_steps;
  const bool ext0 = h->ext_pending;  // false: a block starts at slot 0, after a fold
  QD_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = QD_OK;
  // the block's last step leaves psi valid (a CUDA-order fused handle)
  for (int k = 0; k < h->tblock && rc == QD_OK; k++)
    rc = h->prec == QD_PREC_FP64 ? step_impl<double>(h, false, k + 1 == h->tblock)
                                 : step_impl<float>(h, false, k + 1 == h->tblock);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(h->stream, &g);
  h->slot = slot0;
  h->acc_steps = acc0;
  h->ext_pending = ext0;
  if (rc != QD_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  QD_HIP(e);
  h->graph = g;
  QD_HIP(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
  return QD_OK;
}

int propagate_impl(qd_handle *h, int32_t nsteps) {
  QD_HIP(hipSetDevice(h->dev));
  if (nsteps < 0) return fail(QD_ERR_ARG, "n_steps < 0");
  if (nsteps == 0) return QD_OK;
  const int32_t steps0 = h->steps;
  for (int k = 0; k < TK_N; k++) {
    h->t_ms[k] = 0;
    h->t_n[k] = 0;
  }
  if (h->pending_prephase) {
    QD_RC(materialize_psi(h));
    QD_RC(apply_prephase(h));
    h->pending_prephase = false;
  }
  QD_RC(enter_fbr_state(h));
  if (h->fuse_theta) h->dvr_valid = false;
  hipEvent_t e0 = h->ev[0], e1 = h->ev[1];
  h->nmarks = 0;
  QD_HIP(hipEventRecord(e0, h->stream));
  const bool graphs = h->use_graph && !h->profile;
  for (int s = 0; s < nsteps;) {
    // a whole flux block as one graph replay when it starts a block and no
    // P(E) / callback point falls inside it; otherwise one step eagerly
    const int B = h->tblock;
    bool replay = graphs && h->slot == 0 && nsteps - s >= B;
    if (replay && h->crp_every > 0 && (h->cb || h->do_crp)) {
      const long long next = (h->steps / h->crp_every + 1) * h->crp_every;
      replay = next >= h->steps + B;
    }
    // a diagnostics step runs eagerly (it reads its sums back at its end)
    const bool diag = h->diag_every > 0 && (h->steps + 1) % h->diag_every == 0;
    if (replay && h->diag_every > 0) {
      const long long next = (h->steps / h->diag_every + 1) * h->diag_every;
      replay = next > h->steps + B;
    }
    if (replay) {
      if (!h->gexec) QD_RC(capture_block(h));
      QD_HIP(hipGraphLaunch(h->gexec, h->stream));
      h->steps += B;
      s += B;
      if (h->do_crp) h->acc_steps += B;
      if (h->fuse_theta && h->variant == QD_VARIANT_CUDA) h->dvr_valid = true;  // its last step wrote psi
    } else {
      // psi is read after this step: the call ends, a diagnostics step or a
      // callback follows
      const bool psi_out = s + 1 == nsteps ||
                           (h->diag_every > 0 && (h->steps + 2) % h->diag_every == 0) ||
                           (h->crp_every > 0 && (h->cb || h->do_crp) && (h->steps + 1) % h->crp_every == 0);
      if (h->prec == QD_PREC_FP64) QD_RC(step_impl<double>(h, diag, psi_out));
      else QD_RC(step_impl<float>(h, diag, psi_out));
      h->steps++;
      s++;
    }
    const bool every = h->crp_every > 0 && (h->steps % h->crp_every == 0);
    if (every && (h->cb || h->do_crp)) {
      std::vector<double> tot;
      const double *c = nullptr;
      if (h->do_crp) {
        QD_RC(compute_crp(h));
        tot.assign(h->ne, 0.0);
        for (int b = 0; b < h->nb; b++)
          for (int e = 0; e < h->ne; e++) tot[e] += h->host_crp[(size_t)b * h->ne + e];
        c = tot.data();
      }
      if (h->cb && h->cb(h->steps, c, h->cb_user)) break;
    }
  }
  QD_RC(resolve_ext(h));  // the last step's extraction: psi may be read next
  if (h->profile) {
    QD_RC(mark(h, -1));
    QD_RC(drain_marks(h));
    h->nmarks = 0;
  }
  QD_HIP(hipEventRecord(e1, h->stream));
  QD_HIP(hipEventSynchronize(e1));
  float ms = 0;
  QD_HIP(hipEventElapsedTime(&ms, e0, e1));
  // a callback may stop the loop early: average over the steps actually taken
  const int32_t taken = h->steps - steps0;
  h->last_step_ms = taken > 0 ? ms / taken : 0.0;
  return QD_OK;
}

int get_psi_impl(qd_handle *h, double *out) {
  QD_HIP(hipSetDevice(h->dev));
  QD_RC(materialize_psi(h));
  return get_complex(h, h->prec == QD_PREC_FP64 ? (const void *)h->psi64 : (const void *)h->psi32, 0, out, h->npt);
}

int norm_impl(qd_handle *h, double *out) {
  QD_HIP(hipSetDevice(h->dev));
  QD_RC(materialize_psi(h));
  if (h->prec == QD_PREC_FP64)
    hipLaunchKernelGGL(k_norm_partial<double>, dim3(h->npartial), dim3(256), 0, h->stream, h->psi64, h->dcol_w,
                       h->n12, h->npt, h->partial);
  else
    hipLaunchKernelGGL(k_norm_partial<float>, dim3(h->npartial), dim3(256), 0, h->stream, h->psi32, h->dcol_w,
                       h->n12, h->npt, h->partial);
  QD_HIP(hipGetLastError());
  std::vector<double> part(h->npartial);
  QD_HIP(hipMemcpyAsync(part.data(), h->partial, sizeof(double) * h->npartial, hipMemcpyDeviceToHost, h->stream));
  QD_HIP(hipStreamSynchronize(h->stream));
  double s = 0;
  for (double v : part) s += v;
  *out = s * h->dr1 * h->dr2;
  return QD_OK;
}

// ---- autotuner ----
// Times each candidate configuration of a pass on the handle's own arrays
// (theta forward reads psi and writes the FBR scratch array; the R and r
// passes work in place on the FBR array, which the next step overwrites) and
// keeps the fastest.  psi is not modified.
template <typename F>
int time_launches(qd_handle *h, F launch, int reps, float *ms) {
  QD_RC(launch());  // warm-up (and first-use setup)
  QD_HIP(hipEventRecord(h->ev[0], h->stream));
  for (int k = 0; k < reps; k++) QD_RC(launch());
  QD_HIP(hipEventRecord(h->ev[1], h->stream));
  QD_HIP(hipEventSynchronize(h->ev[1]));
  QD_HIP(hipEventElapsedTime(ms, h->ev[0], h->ev[1]));
  *ms /= reps;
  return QD_OK;
}

int autotune_handle(qd_handle *h) {
  const int reps = 3;
  const int full = FFT_PHASE | FFT_FWD | FFT_KIN | FFT_INV;
  const bool f64 = h->prec == QD_PREC_FP64;
  void *lpsi = f64 ? (void *)h->lpsi64 : (void *)h->lpsi32;
  // theta: 16-row tiles per work group of the matrix-stationary kernel
  if (h->tmx_ks) {
    const long long ntiles = (h->n12 + 15) / 16;
    const int base = h->tmx_tpw;
    std::vector<int> cands = {base, std::max(1, base / 2), base * 3 / 2, base * 2, base * 4};
    float best = 1e30f;
    int pick = base;
    for (int c : cands) {
      if (c < 1 || c > ntiles) continue;
      h->tmx_tpw = c;
      float ms = 0;
      QD_RC(time_launches(h, [&] { return launch_theta(h, true, 1, 0); }, reps, &ms));
      if (ms < best) {
        best = ms;
        pick = c;
      }
    }
    h->tmx_tpw = pick;
  }
  // fused theta kernel: tile-blocks per work item (items of blocks sharing V
  // are at least one tile per wave, so smaller values matter too)
  if (h->fuse_theta) {
    // kernel, at the starting tile count: the dense transforms (512 threads),
    // the parity halves (512 or 768 threads for half bases of at most 32, 256
    // above), or the parity halves with wave pairs (512 threads, half bases of
    // 33..64)
    if (h->par_ok) {
      struct Cand {
        int kind, nth;  // kind 0 dense, 1 parity halves, 2 wave pairs
      };
      std::vector<Cand> cands;
      if (h->dense_ok) cands.push_back({0, TMX_THREADS});
      if (h->par_ksh == 1) {
        cands.push_back({1, 512});
        cands.push_back({1, 768});
      } else {
        cands.push_back({1, 256});
      }
      if (h->pair_ok) cands.push_back({2, 512});
      if (cands.size() > 1) {
        // three interleaved rounds, the fastest round of each candidate: the
        // kernels can be within a few per cent, and one round flips on noise
        std::vector<float> tmin(cands.size(), 1e30f);
        for (int round = 0; round < 3; round++)
          for (size_t c = 0; c < cands.size(); c++) {
            h->tfu_par = cands[c].kind >= 1;
            h->tfu_pair = cands[c].kind == 2;
            h->tfu_nth = cands[c].nth;
            QD_RC(build_tfu_items(h));
            float ms = 0;
            QD_RC(time_launches(h, [&] { return launch_theta_fused(h, false); }, reps, &ms));
            tmin[c] = std::min(tmin[c], ms);
          }
        const size_t pick = (size_t)(std::min_element(tmin.begin(), tmin.end()) - tmin.begin());
        h->tfu_par = cands[pick].kind >= 1;
        h->tfu_pair = cands[pick].kind == 2;
        h->tfu_nth = cands[pick].nth;
      }
    }
    const long long ntiles = (h->n12 + 15) / 16;
    const int base = h->tfu_tpw;
    std::vector<int> cands = {base,     std::max(1, base / 2), std::max(1, base / 4), std::max(1, base / 8),
                              base * 3 / 2, base * 2,          base * 4};
    float best = 1e30f;
    int pick = base;
    for (int c : cands) {
      if (c < 1 || c > ntiles) continue;
      h->tfu_tpw = c;
      QD_RC(build_tfu_items(h));
      float ms = 0;
      QD_RC(time_launches(h, [&] { return launch_theta_fused(h, false); }, reps, &ms));
      if (ms < best) {
        best = ms;
        pick = c;
      }
    }
    h->tfu_tpw = pick;
    QD_RC(build_tfu_items(h));
  }
  // R and r passes: register FFT or Stockham kernels
  auto pick_pass = [&](bool *flag, bool alongR) -> int {
    if (!*flag) return QD_OK;
    float t_rf = 0, t_st = 0;
    auto run = [&] {
      if (alongR)
        return f64 ? launch_rpass<double>(h, full, (double2 *)lpsi) : launch_rpass<float>(h, full, (float2 *)lpsi);
      return f64 ? launch_rline<double>(h, full, (double2 *)lpsi) : launch_rline<float>(h, full, (float2 *)lpsi);
    };
    QD_RC(time_launches(h, run, reps, &t_rf));
    *flag = false;
    QD_RC(time_launches(h, run, reps, &t_st));
    *flag = t_rf <= t_st;
    return QD_OK;
  };
  // R pass: the prefetching mixing kernel (one or two work groups per CU),
  // the register FFT with LDS-staged mixing, the Stockham kernels or the split
  // pass, whichever of them this handle has
  {
    const bool has_mx = h->rp_mx, has_mx2 = h->rp_mx && h->mxs[0].slots2 > 0, has_rf = h->rp_rf,
               has_st = h->rp_legacy, has_sp = h->sp_ok;
    auto runR = [&] {
      return f64 ? launch_rpass<double>(h, full, (double2 *)lpsi) : launch_rpass<float>(h, full, (float2 *)lpsi);
    };
    const bool cand[5] = {has_mx, has_rf, has_st, has_mx2, has_sp};
    int ncand = 0;
    for (bool c : cand) ncand += c;
    // the split pass mixes into psi, which holds the state here: its trial
    // runs take a scratch array of their own
    const bool own_tmp = has_sp && ncand > 1 && !h->sp_tmp;
    if (own_tmp) QD_HIP(hipMalloc(&h->sp_tmp, (f64 ? sizeof(double2) : sizeof(float2)) * h->nlp));
    if (ncand > 1) {
      float best = 1e30f;
      int pick = -1;
      for (int c = 0; c < 5; c++) {
        if (!cand[c]) continue;
        h->rp_mx = c == 0 || c == 3;
        h->mx_two = c == 3;
        h->rp_rf = c == 1;
        h->rp_split = c == 4;
        float ms = 0;
        QD_RC(time_launches(h, runR, reps, &ms));
        if (ms < best) {
          best = ms;
          pick = c;
        }
      }
      h->rp_mx = pick == 0 || pick == 3;
      h->mx_two = pick == 3;
      h->rp_rf = pick == 1;
      h->rp_split = pick == 4;
    }
    if (own_tmp) {
      QD_HIP(hipStreamSynchronize(h->stream));
      QD_HIP(hipFree(h->sp_tmp));
      h->sp_tmp = nullptr;
    }
  }
  QD_RC(pick_pass(&h->rl_rf, false));
  // register-FFT r pass: 256 or 512 threads per work group
  if (h->rl_rf && h->rl_slots2 > 0) {
    float t256 = 0, t512 = 0;
    auto run = [&] {
      return f64 ? launch_rline<double>(h, full, (double2 *)lpsi) : launch_rline<float>(h, full, (float2 *)lpsi);
    };
    t256 = t512 = 1e30f;
    for (int round = 0; round < 3; round++) {  // interleaved, the fastest round of each
      float ms = 0;
      h->rl_nth = 256;
      QD_RC(time_launches(h, run, reps, &ms));
      t256 = std::min(t256, ms);
      h->rl_nth = 512;
      QD_RC(time_launches(h, run, reps, &ms));
      t512 = std::min(t512, ms);
    }
    h->rl_nth = t512 < t256 ? 512 : 256;
  }
  h->tuned = true;
  return QD_OK;
}

int create_from_spec(const BatchSpec &S, const qd_options *options, qd_handle **out) {
  qd_handle *h = new (std::nothrow) qd_handle;
  if (!h) return fail(QD_ERR_NOMEM, "out of host memory");
  int rc = set_options(h, options);
  if (rc == QD_OK) rc = create_batch(h, S);
  if (rc == QD_OK && options && options->autotune) rc = autotune_handle(h);
  if (rc != QD_OK) {
    release(h);
    delete h;
    return rc;
  }
  *out = h;
  return QD_OK;
}

}  // namespace

extern "C" {

const char *qd_last_error(void) { return g_last_error.c_str(); }

int qd_create(const qd_problem *problem, const qd_options *options, qd_handle **out) {
  if (!out) return fail(QD_ERR_NULL, "out is NULL");
  *out = nullptr;
  BatchSpec S;
  QD_RC(spec_from_problem(problem, S));
  return create_from_spec(S, options, out);
}

int qd_create_jk(const qd_jk_problem *problem, const qd_options *options, qd_handle **out) {
  if (!out) return fail(QD_ERR_NULL, "out is NULL");
  *out = nullptr;
  BatchSpec S;
  QD_RC(spec_from_jk(problem, S));
  return create_from_spec(S, options, out);
}

int qd_get_tuning(qd_handle *h, int32_t *vals) {
  if (!h || !vals) return fail(QD_ERR_NULL, "NULL argument");
  vals[0] = h->tuned ? 1 : 0;
  vals[1] = h->tmx_ks;
  vals[2] = h->tmx_tpw;
  vals[3] = h->rp_rf ? 1 : 0;
  vals[4] = h->rl_rf ? 1 : 0;
  vals[5] = h->rp_lw;
  vals[6] = h->rp_ml;
  vals[7] = h->rl_rf ? h->rl_nth / rf_n1(h->n2) : h->rl_lpb;
  vals[8] = h->rp_mx ? std::max(h->mxs[0].nl, h->mxs[1].nl) : 0;
  vals[9] = h->rp_mx ? std::max(h->mxs[0].gb, h->mxs[1].gb) : 0;
  vals[10] = h->fuse_theta ? (h->tfu_pair ? 3 : h->tfu_par ? 2 : 1) : 0;
  vals[11] = h->fuse_theta ? h->tfu_tpw : 0;
  vals[13] = h->fuse_theta ? h->tfu_nth : 0;
  vals[12] = h->rp_mx && h->mx_two ? 1 : 0;
  vals[14] = h->rp_split ? 1 : 0;
  return QD_OK;
}

int qd_jk_theta_grid(int32_t K, int32_t jmax, double *x, double *w) {
  if (!x || !w) return fail(QD_ERR_NULL, "NULL argument");
  if (K < 0 || jmax < K) return fail(QD_ERR_ARG, "need 0 <= K <= jmax");
  std::vector<double> xv, wv, Pb;
  qdt::assoc_legendre_dvr(K, jmax, xv, wv, Pb);
  memcpy(x, xv.data(), sizeof(double) * xv.size());
  memcpy(w, wv.data(), sizeof(double) * wv.size());
  return QD_OK;
}

int qd_make_radial(int32_t n, double lo, double hi, double *r, double *dr) {
  if (!r || !dr) return fail(QD_ERR_NULL, "NULL argument");
  if (n < 2 || !(hi > lo)) return fail(QD_ERR_ARG, "need n >= 2 and hi > lo");
  const double step = (hi - lo) / (double)(n - 1);  // linspace: lo + i step, the last point hi exactly
  for (int i = 0; i < n; i++) r[i] = i == n - 1 ? hi : lo + i * step;
  *dr = r[1] - r[0];
  return QD_OK;
}

int qd_make_theta(int32_t n, int32_t m, double *x, double *w, double *legendre) {
  if (!x || !w) return fail(QD_ERR_NULL, "NULL argument");
  if (n < 1 || m < 0) return fail(QD_ERR_ARG, "need n >= 1 and m >= 0");
  qds::gauss_legendre(n, x, w);
  if (legendre) qds::legendre_p2(m, x, n, legendre);
  return QD_OK;
}

int qd_dividing_surface(double rd, double rmin, double dr, int32_t *index) {
  if (!index) return fail(QD_ERR_NULL, "NULL argument");
  if (!(dr > 0)) return fail(QD_ERR_ARG, "dr must be positive");
  *index = qds::dividing_surface(rd, rmin, dr);
  return QD_OK;
}

int qd_init_packet(const qd_radial *r1, const qd_radial *r2, const qd_angle *theta, const qd_packet *pk,
                   int32_t n_energies, const double *energies, double *psi, double *e_diatomic, double *eta_sq) {
  if (!r1 || !r2 || !theta || !pk || !e_diatomic) return fail(QD_ERR_NULL, "NULL argument");
  if (!r1->r || !r2->r || !pk->vdiat || (psi && !theta->x)) return fail(QD_ERR_NULL, "a grid or vdiat is NULL");
  if (n_energies > 0 && (!energies || !eta_sq)) return fail(QD_ERR_NULL, "energies or eta_sq is NULL");
  if (r1->n < 1 || r2->n < 1 || (psi && theta->n < 1)) return fail(QD_ERR_ARG, "grid sizes must be positive");
  if (!(r1->mass > 0) || !(r2->mass > 0) || !(r2->dr > 0) || !(pk->delta > 0))
    return fail(QD_ERR_ARG, "masses, dr and delta must be positive");
  if (pk->v0 < 0 || pk->v0 >= r2->n || pk->j0 < 0) return fail(QD_ERR_ARG, "need 0 <= v0 < r2.n and j0 >= 0");
  if (pk->J < 0 || pk->K < 0 || pk->K > pk->J || pk->K > pk->j0)
    return fail(QD_ERR_ARG, "need 0 <= K <= min(J, j0) for the packet's block");
  if (pk->J > 0 && pk->j0 != 0 && n_energies > 0)
    return fail(QD_ERR_ARG, "the partial-wave normalisation for J > 0 needs j0 = 0 (then l = J)");
  const int n1 = r1->n, n2 = r2->n;
  std::vector<double> phi(n2);
  double ev = 0;
  qds::diatomic_state(r2->r, n2, r2->dr, r2->mass, pk->vdiat, pk->j0, pk->v0, &ev, phi.data());
  *e_diatomic = ev;
  if (psi) {
    const int nt = theta->n;
    std::vector<double> th(nt);
    if (pk->J == 0) {  // sqrt(j0 + 1/2) P_j0(x) by LegendreP.m's recursion (InitWavePacket.m:21-22)
      std::vector<double> P((size_t)(pk->j0 + 1) * nt);
      qds::legendre_p2(pk->j0, theta->x, nt, P.data());
      for (int k = 0; k < nt; k++) th[k] = std::sqrt(pk->j0 + 0.5) * P[(size_t)(pk->j0 + 1) * k + pk->j0];
    } else {
      for (int k = 0; k < nt; k++) th[k] = qds::assoc_legendre_norm(pk->j0, pk->K, theta->x[k]);
    }
    std::vector<std::complex<double>> g(n1);
    for (int i = 0; i < n1; i++) g[i] = qds::packet_g(r1->r[i], pk->R0, pk->delta, pk->k0);
    for (int k = 0; k < nt; k++)
      for (int j = 0; j < n2; j++) {
        double *dst = psi + 2 * ((size_t)n1 * (j + (size_t)n2 * k));
        const double f = phi[j] * th[k];
        for (int i = 0; i < n1; i++) {
          dst[2 * i] = g[i].real() * f;
          dst[2 * i + 1] = g[i].imag() * f;
        }
      }
  }
  for (int e = 0; e < n_energies; e++) {
    const double Et = energies[e] - ev;
    eta_sq[e] = pk->J == 0 ? qds::eta_sq_plane(pk->delta, pk->k0, r1->mass, Et)
                           : qds::eta_sq_partial_wave(pk->R0, pk->delta, pk->k0, r1->mass, pk->J, Et);
  }
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
  for (int e = 0; e < h->ne; e++) {
    double s = 0;
    for (int b = 0; b < h->nb; b++) s += h->host_crp[(size_t)b * h->ne + e];
    crp_out[e] = s;
  }
  return QD_OK;
}

int qd_get_crp_blocks(qd_handle *h, double *crp_out) {
  if (!h || !crp_out) return fail(QD_ERR_NULL, "NULL argument");
  if (!h->do_crp) return fail(QD_ERR_ARG, "calculate_CRP was 0");
  QD_HIP(hipSetDevice(h->dev));
  QD_RC(compute_crp(h));
  memcpy(crp_out, h->host_crp.data(), sizeof(double) * h->host_crp.size());
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

int qd_get_diagnostics(qd_handle *h, double *out) {
  if (!h || !out) return fail(QD_ERR_NULL, "NULL argument");
  if (!h->diag_have) return fail(QD_ERR_ARG, "no step has evaluated the diagnostics (options.diagnostics_every)");
  for (int k = 0; k < 5; k++) out[k] = h->diag_last[k];
  return QD_OK;
}

int qd_get_timing(qd_handle *h, int32_t which, double *ms_out) {
  if (!h || !ms_out) return fail(QD_ERR_NULL, "NULL argument");
  if (which < 0 || which >= TK_N) return fail(QD_ERR_ARG, "bad timing selector");
  if (which == TK_STEP) *ms_out = h->last_step_ms;
  else *ms_out = h->t_n[which] ? h->t_ms[which] / (double)std::max<long long>(1, h->t_n[which]) : 0.0;
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
  if (rc == QD_OK && h->do_crp && problem->time.total_steps > 0) {
    if (!problem->crp.CRP) rc = fail(QD_ERR_NULL, "CRP.CRP is NULL");
    else rc = qd_get_crp(h, problem->crp.CRP);
  }
  if (rc == QD_OK) problem->time.steps = h->steps;
  qd_destroy(h);
  return rc;
}

}  // extern "C"

// ---- state access and test hooks ----
extern "C" {

int qd_set_psi(qd_handle *h, const double *psi_in) {
  if (!h || !psi_in) return fail(QD_ERR_NULL, "NULL argument");
  QD_HIP(hipSetDevice(h->dev));
  QD_RC(put_complex(h, h->prec == QD_PREC_FP64 ? (void *)h->psi64 : (void *)h->psi32, 0, psi_in, h->npt));
  h->state_fbr = false;
  h->dvr_valid = true;
  return QD_OK;
}

}  // extern "C"

// ---- checkpoints ----
namespace {
struct CkptHead {
  char magic[8];                                   // "QDCKPT1"
  int64_t npt, nlp, nx, ne, tblock, prec, nb;      // shape of the handle it came from
  int64_t steps, acc_steps, slot, state_fbr, pending_prephase, bytes;
};
// A CUDA-order fused handle keeps psi beside its FBR state (the state holds the
// dumped psi): its checkpoints carry both.
bool ckpt_both(const qd_handle *h) { return h->fuse_theta && h->variant == QD_VARIANT_CUDA; }
int64_t ckpt_bytes(const qd_handle *h) {
  const int64_t cs = h->prec == QD_PREC_FP64 ? 16 : 8;
  int64_t b = sizeof(CkptHead) + cs * (int64_t)std::max(h->npt, h->nlp);
  if (ckpt_both(h)) b += cs * (int64_t)h->npt;
  if (h->do_crp) b += 2 * 16 * (int64_t)h->nx * h->ne + 2 * 16 * (int64_t)h->nx * h->tblock;
  return b;
}
void ckpt_head(const qd_handle *h, CkptHead &c) {
  memset(&c, 0, sizeof c);
  memcpy(c.magic, "QDCKPT1", 8);
  c.npt = (int64_t)h->npt;
  c.nlp = (int64_t)h->nlp;
  c.nx = h->do_crp ? (int64_t)h->nx : 0;
  c.ne = h->ne;
  c.tblock = h->tblock;
  c.prec = h->prec;
  c.nb = h->nb;
  c.bytes = ckpt_bytes(h);
}
}  // namespace

extern "C" {

int qd_checkpoint_size(qd_handle *h, int64_t *bytes) {
  if (!h || !bytes) return fail(QD_ERR_NULL, "NULL argument");
  *bytes = ckpt_bytes(h);
  return QD_OK;
}

int qd_checkpoint_save(qd_handle *h, void *buf, int64_t bytes) {
  if (!h || !buf) return fail(QD_ERR_NULL, "NULL argument");
  if (bytes < ckpt_bytes(h)) return fail(QD_ERR_ARG, "checkpoint buffer smaller than qd_checkpoint_size");
  QD_HIP(hipSetDevice(h->dev));
  CkptHead c;
  ckpt_head(h, c);
  c.steps = h->steps;
  c.acc_steps = h->acc_steps;
  c.slot = h->slot;
  // 2: the FBR state and psi (a CUDA-order fused handle between calls)
  c.state_fbr = h->state_fbr ? (ckpt_both(h) && h->dvr_valid ? 2 : 1) : 0;
  c.pending_prephase = h->pending_prephase ? 1 : 0;
  char *p = static_cast<char *>(buf);
  memcpy(p, &c, sizeof c);
  p += sizeof c;
  const bool f64 = h->prec == QD_PREC_FP64;
  const size_t cs = f64 ? 16 : 8;
  // the wavefunction in the form the handle propagates it, raw (bit-exact)
  const void *src = h->state_fbr ? (f64 ? (const void *)h->lpsi64 : (const void *)h->lpsi32)
                                 : (f64 ? (const void *)h->psi64 : (const void *)h->psi32);
  const size_t n = h->state_fbr ? h->nlp : h->npt;
  QD_HIP(hipMemcpy(p, src, cs * n, hipMemcpyDeviceToHost));
  p += cs * std::max(h->npt, h->nlp);
  if (ckpt_both(h)) {
    if (c.state_fbr == 2)
      QD_HIP(hipMemcpy(p, f64 ? (const void *)h->psi64 : (const void *)h->psi32, cs * h->npt, hipMemcpyDeviceToHost));
    p += cs * h->npt;
  }
  if (h->do_crp) {
    const size_t acc = sizeof(double2) * h->nx * h->ne, ring = sizeof(double2) * h->nx * h->tblock;
    QD_HIP(hipMemcpy(p, h->fai, acc, hipMemcpyDeviceToHost));
    QD_HIP(hipMemcpy(p + acc, h->dfai, acc, hipMemcpyDeviceToHost));
    QD_HIP(hipMemcpy(p + 2 * acc, h->bpsi, ring, hipMemcpyDeviceToHost));
    QD_HIP(hipMemcpy(p + 2 * acc + ring, h->bdpsi, ring, hipMemcpyDeviceToHost));
  }
  return QD_OK;
}

int qd_checkpoint_load(qd_handle *h, const void *buf, int64_t bytes) {
  if (!h || !buf) return fail(QD_ERR_NULL, "NULL argument");
  if (bytes < (int64_t)sizeof(CkptHead)) return fail(QD_ERR_ARG, "checkpoint too short");
  CkptHead c, want;
  memcpy(&c, buf, sizeof c);
  ckpt_head(h, want);
  if (memcmp(c.magic, want.magic, 8) || c.npt != want.npt || c.nlp != want.nlp || c.nx != want.nx ||
      c.ne != want.ne || c.tblock != want.tblock || c.prec != want.prec || c.nb != want.nb || c.bytes != want.bytes)
    return fail(QD_ERR_ARG, "checkpoint does not belong to a handle of this problem and options");
  if (bytes < c.bytes) return fail(QD_ERR_ARG, "checkpoint buffer shorter than its header says");
  if (c.state_fbr && !h->fuse_theta) return fail(QD_ERR_ARG, "checkpoint holds the fused theta state");
  if (c.state_fbr < 0 || c.state_fbr > 2 || (c.state_fbr == 2 && !ckpt_both(h)))
    return fail(QD_ERR_ARG, "checkpoint state field out of range");
  QD_HIP(hipSetDevice(h->dev));
  QD_HIP(hipStreamSynchronize(h->stream));
  const char *p = static_cast<const char *>(buf) + sizeof c;
  const bool f64 = h->prec == QD_PREC_FP64;
  const size_t cs = f64 ? 16 : 8;
  void *dst = c.state_fbr ? (f64 ? (void *)h->lpsi64 : (void *)h->lpsi32) : (f64 ? (void *)h->psi64 : (void *)h->psi32);
  QD_HIP(hipMemcpy(dst, p, cs * (c.state_fbr ? h->nlp : h->npt), hipMemcpyHostToDevice));
  p += cs * std::max(h->npt, h->nlp);
  if (ckpt_both(h)) {
    if (c.state_fbr == 2)
      QD_HIP(hipMemcpy(f64 ? (void *)h->psi64 : (void *)h->psi32, p, cs * h->npt, hipMemcpyHostToDevice));
    p += cs * h->npt;
  }
  if (h->do_crp) {
    const size_t acc = sizeof(double2) * h->nx * h->ne, ring = sizeof(double2) * h->nx * h->tblock;
    QD_HIP(hipMemcpy(h->fai, p, acc, hipMemcpyHostToDevice));
    QD_HIP(hipMemcpy(h->dfai, p + acc, acc, hipMemcpyHostToDevice));
    QD_HIP(hipMemcpy(h->bpsi, p + 2 * acc, ring, hipMemcpyHostToDevice));
    QD_HIP(hipMemcpy(h->bdpsi, p + 2 * acc + ring, ring, hipMemcpyHostToDevice));
    const long long a = c.acc_steps;
    QD_HIP(hipMemcpy(h->d_acc_steps, &a, sizeof a, hipMemcpyHostToDevice));
  }
  h->steps = (int32_t)c.steps;
  h->acc_steps = c.acc_steps;
  h->part_valid = false;  // the loaded accumulators' P(E) partials are rebuilt by the next fold
  h->slot = (int)c.slot;
  h->state_fbr = c.state_fbr != 0;
  h->dvr_valid = c.state_fbr != 1;
  h->pending_prephase = c.pending_prephase != 0;
  return QD_OK;
}

}  // extern "C"

extern "C" {

int qd_debug_get_fbr(qd_handle *h, double *out) {
  if (!h || !out) return fail(QD_ERR_NULL, "NULL argument");
  QD_HIP(hipSetDevice(h->dev));
  return get_complex(h, h->prec == QD_PREC_FP64 ? (const void *)h->lpsi64 : (const void *)h->lpsi32, 0, out, h->nlp);
}

int qd_debug_set_fbr(qd_handle *h, const double *in) {
  if (!h || !in) return fail(QD_ERR_NULL, "NULL argument");
  QD_HIP(hipSetDevice(h->dev));
  return put_complex(h, h->prec == QD_PREC_FP64 ? (void *)h->lpsi64 : (void *)h->lpsi32, 0, in, h->nlp);
}

int qd_debug_select(qd_handle *h, int32_t what, int32_t value) {
  if (!h) return fail(QD_ERR_NULL, "handle is NULL");
  if (what == 0) {
    if (h->mxs[0].nbundle == 0 && h->mxs[1].nbundle == 0)
      return fail(QD_ERR_ARG, "this handle has no prefetching mixing R pass");
    if (value == 1 && h->mxs[0].slots2 == 0) return fail(QD_ERR_ARG, "two work groups per CU do not fit this handle");
    if (value != 0 && value != 1) return fail(QD_ERR_ARG, "value must be 0 or 1");
    drop_graph(h);
    h->rp_mx = true;
    h->rp_rf = false;
    h->rp_split = false;
    h->mx_two = value == 1;
    return QD_OK;
  }
  if (what == 2) {
    if (value != 0 && value != 1) return fail(QD_ERR_ARG, "value must be 0 or 1");
    if (value == 1 && !h->sp_ok) return fail(QD_ERR_ARG, "this handle has no split mixing R pass");
    if (value == 0 && !h->rp_legacy && !(h->mxs[0].nbundle || h->mxs[1].nbundle))
      return fail(QD_ERR_ARG, "the split R pass is the only one these mixing groups fit");
    drop_graph(h);
    h->rp_split = value == 1;
    if (value == 1) h->rp_mx = h->rp_rf = false;
    else if (!h->rp_legacy) h->rp_mx = true;
    return QD_OK;
  }
  if (what == 1) {
    if (value < 1 || value > 64) return fail(QD_ERR_ARG, "chunk must be 1..64");
    drop_graph(h);
    h->mx_chunk = value;
    return QD_OK;
  }
  if (what == 5) {
    if (value != 0 && value != 1) return fail(QD_ERR_ARG, "value must be 0 or 1");
    QD_RC(resolve_ext(h));
    drop_graph(h);
    h->ext_in_rline = value == 1;
    return QD_OK;
  }
  if (what == 6) {
    if (value != 0 && value != 1) return fail(QD_ERR_ARG, "value must be 0 or 1");
    if (!h->pair_ok) return fail(QD_ERR_ARG, "this handle has no wave-pair theta kernel");
    drop_graph(h);
    h->tfu_psync = value == 1;
    return QD_OK;
  }
  if (what == 4) {
    if (value != 256 && value != 512) return fail(QD_ERR_ARG, "value must be 256 or 512");
    if (!h->rl_rf) return fail(QD_ERR_ARG, "this handle has no register-FFT r pass");
    if (value == 512 && h->rl_slots2 == 0) return fail(QD_ERR_ARG, "the 512-thread r pass does not fit this handle");
    drop_graph(h);
    h->rl_nth = value;
    return QD_OK;
  }
  if (what == 3) {
    if (!h->fuse_theta) return fail(QD_ERR_ARG, "this handle has no fused theta step");
    if (value < 0 || value > 3) return fail(QD_ERR_ARG, "value must be 0, 1, 2 or 3");
    if (value > 0 && !h->par_ok) return fail(QD_ERR_ARG, "this handle's theta matrices have no parity halves");
    if (value == 0 && !h->dense_ok) return fail(QD_ERR_ARG, "the dense fused kernel does not fit this basis");
    if (value == 2 && h->par_ksh != 1) return fail(QD_ERR_ARG, "768 threads only for half bases of at most 32");
    if (value == 3 && !h->pair_ok) return fail(QD_ERR_ARG, "wave pairs only for half bases of 33..64");
    drop_graph(h);
    h->tfu_par = value > 0;
    h->tfu_pair = value == 3;
    h->tfu_nth = value == 2 ? 768 : value == 1 && h->par_ksh == 2 ? 256 : TMX_THREADS;
    return build_tfu_items(h);
  }

  return fail(QD_ERR_ARG, "unknown selection");
}

int qd_debug_stage(qd_handle *h, int32_t stage) {
  if (!h) return fail(QD_ERR_NULL, "handle is NULL");
  QD_HIP(hipSetDevice(h->dev));
  const int full = FFT_PHASE | FFT_FWD | FFT_KIN | FFT_INV;
  int rc = QD_OK;
  if (stage == 0) rc = launch_theta(h, true, 0, 0);
  else if (stage == 3) rc = launch_theta(h, false, 0, 0);
  else if (stage >= 1 && stage <= 7) {
    const int flags = stage == 1 || stage == 2 ? full : (stage == 4 || stage == 5 ? FFT_FWD : FFT_INV);
    const bool alongR = stage == 1 || stage == 4 || stage == 6;
    if (h->prec == QD_PREC_FP64)
      rc = alongR ? launch_rpass<double>(h, flags, h->lpsi64) : launch_rline<double>(h, flags, h->lpsi64);
    else
      rc = alongR ? launch_rpass<float>(h, flags, h->lpsi32) : launch_rline<float>(h, flags, h->lpsi32);
  } else {
    return fail(QD_ERR_ARG, "bad stage");
  }
  if (rc != QD_OK) return rc;
  QD_HIP(hipStreamSynchronize(h->stream));
  return QD_OK;
}

}  // extern "C"
