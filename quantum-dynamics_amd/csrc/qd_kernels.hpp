// qd_kernels.hpp -- CDNA4 (gfx950) kernels of the split-operator step.
//
// One step of the reference (common/timeEvol.C:619-647, evolution_dt) is
//   V phase -> theta forward (DVR->FBR) -> rot/2 -> 2-D FFT -> kinetic ->
//   2-D IFFT -> rot/2 -> theta backward (FBR->DVR) -> dump -> flux surface.
// Here it runs as four HBM sweeps over a batch of (J, K) blocks:
//   k_theta_*  (fwd)   V phase fused into the operand load
//   k_rpass            K->l mixing, D_R/2, FFT_R, exp(-i T1 dt)/n1, IFFT_R, D_R/2, l->K mixing
//   k_rline            D_r/2, FFT_r, exp(-i T2 dt)/n2, IFFT_r, D_r/2
//   k_theta_*  (bwd)   dump fused into the store
// plus the flux-surface extraction and, once per time block, the
// time->energy accumulation.
//
// Why the middle of the step splits this way.  With D_R = exp(-i dt/2 l(l+1)/(2 mu1 R^2))
// (l the orbital quantum number, = j for J = 0) and D_r = exp(-i dt/2 j(j+1)/(2 mu2 r^2)),
// the reference's  rot/2 . Kin . rot/2  equals
//   [D_r/2 Kin_r D_r/2] . U [D_R/2 Kin_R D_R/2] U^T
// exactly, because Kin_R, Kin_r, D_R, D_r act on different axes and U (the
// K <-> l transform of the body-fixed blocks, identity for J = 0) touches
// neither R nor r.  So each 1-D pass carries its own phases and the K mixing
// happens inside the R pass, on lines already in LDS.
//
// Layout.  Blocks sit back to back; block b owns DVR columns dcol0..dcol0+nd-1
// and FBR columns fcol0..fcol0+nf-1.  A column is one n1 x n2 slab,
// column-major with R fastest (common/mat.h:18-41), so psi[i + n1 j + n1 n2 c].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qd {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

template <typename T> struct cx_of;
template <> struct cx_of<double> { typedef double2 type; };
template <> struct cx_of<float> { typedef float2 type; };

template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename C, typename T> __device__ __forceinline__ C cscale(C a, T s) {
  C r;
  r.x = a.x * s;
  r.y = a.y * s;
  return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) {
  C r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}
template <typename C> __device__ __forceinline__ C csub(C a, C b) {
  C r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}
// i * s * a for real s
template <typename C, typename T> __device__ __forceinline__ C cmul_i(C a, T s) {
  C r;
  r.x = -a.y * s;
  r.y = a.x * s;
  return r;
}

// Fixed operator tables (FFT twiddles, kinetic and rotational phase factors).
// FP64 tables hold the value; FP32 tables hold hi + lo, hi = fl32(w) and
// lo = fl32(w - hi), applied with fused multiply-adds so that the FP32 product
// is the rounding of a*w rather than of a*fl32(w).  The same few twiddles and
// phase factors multiply every line of every step, so their FP32 rounding
// (|fl32(w)|^2 - 1 ~ 1e-8) would otherwise shrink or grow the norm coherently.
template <typename T> struct tab_of;
template <> struct tab_of<double> { typedef double2 type; };
template <> struct tab_of<float> { typedef float4 type; };

__device__ __forceinline__ double2 tmul(double2 a, double2 w, bool conj_w) {
  if (conj_w) w.y = -w.y;
  return cmul(a, w);
}
__device__ __forceinline__ float2 tmul(float2 a, float4 w, bool conj_w) {
  const float wy = conj_w ? -w.y : w.y, ly = conj_w ? -w.w : w.w;
  const float tx = a.x * w.z - a.y * ly;
  const float ty = a.x * ly + a.y * w.z;
  float2 r;
  r.x = fmaf(a.x, w.x, fmaf(-a.y, wy, tx));
  r.y = fmaf(a.x, wy, fmaf(a.y, w.x, ty));
  return r;
}
// Tables of the register-FFT passes (qd_rfft.hpp).  FP64: the complex values,
// as above.  FP32: every factor is a rotation applied as three shears
// (lifting steps, qdt::lift_consts), whose determinant is 1 for any FP32
// constants, so no table can scale the norm; an entry holds (t, s) =
// (tan(theta/2), sin theta) with |theta| <= pi/2:
//   tw   twiddle W_N^e for |e| <= N/4 (centred indices), float2 (t, s)
//   ph   D/2 phase, any angle theta in (-pi, pi], as two rotations by
//        theta/2: float4 (t, s, 2t, 0) of theta/2 (the middle shears merge,
//        five FMAs)
//   kin  kinetic factor sigma e^{i theta} / n as float4 (t, s, c_hi, c_lo),
//        sigma = +-1, c = sigma / n, in the order the threads read it
template <typename T> struct rf_tab;
template <> struct rf_tab<double> {
  typedef double2 tw;
  typedef double2 kin;
  typedef double2 ph;
};
template <> struct rf_tab<float> {
  typedef float2 tw;
  typedef float4 kin;
  typedef float4 ph;
};

// x * s for a real constant given as hi + lo (FP32) or a double
__device__ __forceinline__ float csmul(float x, float hi, float lo) { return fmaf(x, hi, x * lo); }
__device__ __forceinline__ double csmul(double x, double hi, double) { return x * hi; }

// ---------------------------------------------------------------------------
// Batch description
// ---------------------------------------------------------------------------
struct BlockDesc {
  int J, K;
  int nd, nf;        // theta DVR points, FBR functions
  int dcol0, fcol0;  // first DVR / FBR column
  int mat;           // index of the block's theta matrices
  int pad;
};

// ---------------------------------------------------------------------------
// Elementwise kernels
// ---------------------------------------------------------------------------
// V of point x (global DVR layout) is V[dcol_vidx[c] + (x mod n12)], c = x / n12.

// psi /= sqrt_ref(exp(-i V dt)), the first-step half phase of the CPU
// propagator (timeEvol.C:519-538).  sqrt_ref is the reference's pow(c, 0.5)
// with arg() = asin(Im/|c|) plus quadrant fixes (complex.h:60-68,148-159); it
// equals exp(+i V dt/2) only while |V dt| < pi, which is reproduced here.
template <typename C>
__global__ void k_prephase_cpu(C *psi, const double *V, const long long *dcol_vidx, size_t n12, double dt,
                               size_t n) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const double v = V[dcol_vidx[x / n12] + (long long)(x % n12)];
  double s, c;
  sincos(-dt * v, &s, &c);
  const double a = sqrt(c * c + s * s);
  double ang = asin(s / a);
  if (c < 0 && s > 0) ang = M_PI - ang;
  else if (c < 0 && s < 0) ang = -(M_PI + ang);
  const double th = 0.5 * ang;
  const double r = pow(c * c + s * s, 0.25);
  double sq_s, sq_c;
  sincos(th, &sq_s, &sq_c);
  const double2 q = make_double2(r * sq_c, r * sq_s);
  // c1 / abs2(c2) * conj(c2)
  const double q2 = q.x * q.x + q.y * q.y;
  double2 p = make_double2(psi[x].x, psi[x].y);  // FP64 arithmetic for either storage type
  p.x /= q2;
  p.y /= q2;
  p = cmul(p, make_double2(q.x, -q.y));
  psi[x].x = p.x;
  psi[x].y = p.y;
}

// psi *= exp(+i V dt/2): the CUDA propagator's first-step phase
// (evolutionCUDA2.cu:507, evolution_with_potential(-dt/2)).
template <typename C>
__global__ void k_prephase_cuda(C *psi, const double *V, const long long *dcol_vidx, size_t n12, double dt,
                                size_t n) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const double v = V[dcol_vidx[x / n12] + (long long)(x % n12)];
  double s, c;
  sincos(0.5 * dt * v, &s, &c);
  const double2 p = cmul(make_double2(psi[x].x, psi[x].y), make_double2(c, s));
  psi[x].x = p.x;
  psi[x].y = p.y;
}

// psi *= exp(-i V dt) (evolution_with_potential_dt, timeEvol.C:540-559); FP64
// V with the phase computed on the fly, or the FP32 phase table.
template <typename T>
__global__ void k_vphase(typename cx_of<T>::type *psi, const double *V, const typename cx_of<T>::type *expV,
                         const long long *dcol_vidx, size_t n12, double dt, size_t n) {
  typedef typename cx_of<T>::type C;
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const long long vi = dcol_vidx[x / n12] + (long long)(x % n12);
  C e;
  if (V) {
    double s, c;
    sincos(-dt * V[vi], &s, &c);
    e.x = (T)c;
    e.y = (T)s;
  } else {
    e = expV[vi];
  }
  psi[x] = cmul(psi[x], e);
}

// psi *= d1[i] d2[j] (dump_psi, timeEvol.C:713-735)
template <typename T>
__global__ void k_dump(typename cx_of<T>::type *psi, const double *d1, const double *d2, int n1, int n2,
                       size_t n) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const int i = (int)(x % n1);
  const int j = (int)((x / n1) % n2);
  const double d = d1[i] * d2[j];
  psi[x] = cscale(psi[x], (T)d);
}

// lpsi *= D_R/2[lam(c)][i] D_r/2[jr(c)][j]  (evolution_with_rotational_dt_2, timeEvol.C:561-582)
template <typename T>
__global__ void k_rot(typename cx_of<T>::type *lpsi, const typename tab_of<T>::type *dR,
                      const typename tab_of<T>::type *dr, const int *fcol_lam, const int *fcol_j, int n1, int n2,
                      size_t n) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const int i = (int)(x % n1);
  const size_t jc = x / n1;
  const int j = (int)(jc % n2);
  const int c = (int)(jc / n2);
  lpsi[x] = tmul(tmul(lpsi[x], dR[(size_t)fcol_lam[c] * n1 + i], false), dr[(size_t)fcol_j[c] * n2 + j], false);
}

// lpsi *= kinR[i] kinr[j]  (evolution_with_kinetic_dt, timeEvol.C:584-606; the
// tables carry 1/n1 and 1/n2)
template <typename T>
__global__ void k_kin2d(typename cx_of<T>::type *lpsi, const typename tab_of<T>::type *kinR,
                        const typename tab_of<T>::type *kinr, int n1, int n2, size_t n) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const int i = (int)(x % n1);
  const int j = (int)((x / n1) % n2);
  lpsi[x] = tmul(tmul(lpsi[x], kinR[i], false), kinr[j], false);
}

static __global__ void k_f64_to_f32(const double2 *a, float2 *b, size_t n) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) b[x] = make_float2((float)a[x].x, (float)a[x].y);
}
static __global__ void k_f32_to_f64(const float2 *a, double2 *b, size_t n) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) b[x] = make_double2((double)a[x].x, (double)a[x].y);
}

// ---------------------------------------------------------------------------
// Stockham FFT building blocks
// ---------------------------------------------------------------------------
// Pass with radix r and span Ns reads x[b + q n/r], twiddles by
// w^(q (b mod Ns) n/(Ns r)), and writes y[(b/Ns) Ns r + (b mod Ns) + q Ns]
// (autosort Stockham; any order of the radices gives the DFT).
constexpr int FFT_MAX_PASS = 12;
constexpr int FFT_BPT = 8;  // butterflies per thread per pass (max)

struct FftPlanArg {
  int n;
  int npass;
  int radix[FFT_MAX_PASS];
};

enum {
  FFT_FWD = 1,
  FFT_KIN = 2,       // multiply the spectrum by kin[p]
  FFT_INV = 4,
  FFT_PHASE = 8,     // D/2 before the forward and after the inverse transform
  FFT_MIX = 16       // K <-> l mixing of line groups (R pass)
};

template <typename C, typename T>
__device__ __forceinline__ void dft_small(C *v, int r, T sign) {
  if (r == 2) {
    const C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if (r == 4) {
    const C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const C t2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    const C wd = cmul_i(d, sign);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, wd);
    v[3] = csub(t1, wd);
  } else if (r == 3) {
    // sin(2 pi/3) as hi + lo in FP32
    const T s3h = (T)0.86602540378443864676372317075294;
    const T s3l = (T)(0.86602540378443864676372317075294 - (double)s3h);
    const C t = cadd(v[1], v[2]);
    const C d = csub(v[1], v[2]);
    const C a0 = v[0];
    v[0] = cadd(a0, t);
    const C mm = csub(a0, cscale(t, (T)0.5));
    C ds;
    ds.x = csmul(d.x, s3h, s3l);
    ds.y = csmul(d.y, s3h, s3l);
    const C jd = cmul_i(ds, sign);
    v[1] = cadd(mm, jd);
    v[2] = csub(mm, jd);
  } else if (r == 5) {
    const double C1 = 0.30901699437494742410229341718282, C2 = -0.80901699437494742410229341718282;
    const double S1 = 0.95105651629515357211643933337938, S2 = 0.58778525229247312916870595463907;
    const T c1 = (T)C1, c2 = (T)C2, s1 = (T)S1, s2 = (T)S2;
    const T c1l = (T)(C1 - (double)c1), c2l = (T)(C2 - (double)c2), s1l = (T)(S1 - (double)s1),
            s2l = (T)(S2 - (double)s2);
    auto sc = [](C z, T h, T l) {
      C r;
      r.x = csmul(z.x, h, l);
      r.y = csmul(z.y, h, l);
      return r;
    };
    const C a0 = v[0];
    const C b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
    const C d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
    v[0] = cadd(a0, cadd(b1, b2));
    const C m1 = cadd(a0, cadd(sc(b1, c1, c1l), sc(b2, c2, c2l)));
    const C m2 = cadd(a0, cadd(sc(b1, c2, c2l), sc(b2, c1, c1l)));
    const C n1 = cmul_i(cadd(sc(d1, s1, s1l), sc(d2, s2, s2l)), sign);
    const C n2 = cmul_i(csub(sc(d1, s2, s2l), sc(d2, s1, s1l)), sign);
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
  }
}

// One Stockham pass on the lines of a work group, in place in LDS.
// lidx(line, p) maps a line element to LDS; `slot`/`nslots` give this thread's
// lines and `bt`/`tpl` its butterflies.  Two barriers: all reads complete
// before any write.
template <typename T, typename LIDX, typename MUL>
__device__ __forceinline__ void stockham_pass(typename cx_of<T>::type *lds, LIDX lidx, int nlines, int slot,
                                              int nslots, int bt, int tpl, int n, int r, int Ns, bool fwd,
                                              const typename tab_of<T>::type *tw, MUL post) {
  typedef typename cx_of<T>::type C;
  const T sign = fwd ? (T)-1 : (T)1;
  const int nb = n / r;
  C v[FFT_BPT][5];
  const int line = slot;  // each slot owns at most one line (host guarantees nlines <= nslots)
  const bool act = line < nlines;
#pragma unroll
  for (int bb = 0; bb < FFT_BPT; bb++) {
    const int b = bt + bb * tpl;
    if (act && b < nb) {
#pragma unroll
      for (int q = 0; q < 5; q++) {
        if (q >= r) break;
        v[bb][q] = lds[lidx(line, b + q * nb)];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int bb = 0; bb < FFT_BPT; bb++) {
    const int b = bt + bb * tpl;
    if (act && b < nb) {
      const int k = b % Ns;
      if (Ns > 1) {
        const int estep = k * (n / (Ns * r));
#pragma unroll
        for (int q = 1; q < 5; q++) {
          if (q >= r) break;
          v[bb][q] = tmul(v[bb][q], tw[q * estep], !fwd);
        }
      }
      dft_small<C, T>(v[bb], r, sign);
      const int od = (b / Ns) * Ns * r + k;
#pragma unroll
      for (int q = 0; q < 5; q++) {
        if (q >= r) break;
        const int p = od + q * Ns;
        lds[lidx(line, p)] = post(line, p, v[bb][q]);
      }
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// R pass: lines along R (contiguous), grouped for the K <-> l mixing
// ---------------------------------------------------------------------------
// A work group takes one bundle of lines.  Lines come in units: a unit is the
// g FBR lines (J, K = 0..g-1, j) at one r index, mixed by the g x g matrix U
// of (J, j): c_m = sum_K U[K][m] c_K before the transform and c_K = sum_m
// U[K][m] c_m after it.  For J = 0 (and j = 0) units are single lines and the
// phase factors ride on the global load and store.  Twiddles and the kinetic
// factor sit in LDS; each line has `tpl` threads for its whole transform.
constexpr int RP_THREADS = 256;
constexpr int RP_GMAX = 32;  // largest unit (mixing group)

struct RUnit {
  int col0;   // index into unit_cols of the unit's first column
  int g;      // lines in the unit
  int u0;     // offset of U in umat (g*g, row K, column m), -1 = identity
  int r;      // r index of the lines
};

template <typename T>
struct RPassArgs {
  typename cx_of<T>::type *data;
  const typename tab_of<T>::type *tw;   // exp(-2 pi i e / n1)
  const typename tab_of<T>::type *kin;  // exp(-i T1 dt) / n1
  const typename tab_of<T>::type *dR;   // D_R/2 table [lam][n1]
  const int *fcol_lam;                 // lam index per FBR column (in the l basis)
  const int *bundle_u0;                // [nbundle + 1] range of units per bundle
  const RUnit *units;
  const int *unit_cols;                // FBR columns of the units
  const T *umat;
  int n1, n2;
  int tpl;
  int flags;
  FftPlanArg plan;
  int nfcol;                           // FBR columns (lines = nfcol * n2)
  // register-FFT mixing kernel: per-bundle line tables (built on the host)
  const int *bundle_l0;                // [nbundle + 1] first line of each bundle
  const long long *line_off;           // element offset of each line
  const int *line_lam;                 // D_R row of each line
  int nbundle;
  // the register-FFT kernels' tables (rf_tab; the FP64 ones are tw, kin, dR)
  const typename rf_tab<T>::tw *ltw;
  const typename rf_tab<T>::kin *lkin;
  const typename rf_tab<T>::ph *lph;
};

template <typename T>
__global__ __launch_bounds__(RP_THREADS) void k_rpass(RPassArgs<T> A) {
  typedef typename cx_of<T>::type C;
  typedef typename tab_of<T>::type TB;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = A.n1;
  TB *s_tw = reinterpret_cast<TB *>(smem_raw);
  TB *s_kin = s_tw + n;
  C *lds = reinterpret_cast<C *>(s_kin + n);
  __shared__ long long s_off[64];  // global offset of each line of the bundle
  __shared__ int s_lam[64];
  __shared__ int s_mixed[64];      // line belongs to a unit that is mixed
  __shared__ int s_unit_line0[65];
  const int t = threadIdx.x;
  const int ub = A.bundle_u0[blockIdx.x], ue = A.bundle_u0[blockIdx.x + 1];
  const int nu = ue - ub;
  const bool mix = A.flags & FFT_MIX;
  const bool phase = A.flags & FFT_PHASE;
  if (t == 0) {
    int L = 0;
    for (int u = 0; u < nu; u++) {
      s_unit_line0[u] = L;
      L += A.units[ub + u].g;
    }
    s_unit_line0[nu] = L;
  }
  for (int e = t; e < n; e += RP_THREADS) {
    s_tw[e] = A.tw[e];
    s_kin[e] = A.kin[e];
  }
  __syncthreads();
  const int nl = s_unit_line0[nu];
  if (t < nl) {
    int u = 0;
    while (s_unit_line0[u + 1] <= t) u++;
    const RUnit U = A.units[ub + u];
    const int col = A.unit_cols[U.col0 + (t - s_unit_line0[u])];
    s_off[t] = ((long long)col * A.n2 + U.r) * n;
    s_lam[t] = A.fcol_lam[col];
    s_mixed[t] = (mix && U.u0 >= 0) ? 1 : 0;
  }
  __syncthreads();
  // load; D_R/2 of unmixed lines rides on the load
  for (int e = t; e < nl * n; e += RP_THREADS) {
    const int L = e / n, p = e - L * n;
    C x = A.data[s_off[L] + p];
    if (phase && !s_mixed[L]) x = tmul(x, A.dR[(size_t)s_lam[L] * n + p], false);
    lds[L * n + p] = x;
  }
  __syncthreads();
  // K -> l mixing and D_R/2 of mixed units (one thread per unit position)
  if (mix) {
    for (int e = t; e < nu * n; e += RP_THREADS) {
      const int u = e / n, p = e - u * n;
      const RUnit U = A.units[ub + u];
      if (U.u0 < 0) continue;
      const int L0 = s_unit_line0[u];
      C in[RP_GMAX];
#pragma unroll
      for (int K = 0; K < RP_GMAX; K++)
        if (K < U.g) in[K] = lds[(L0 + K) * n + p];
#pragma unroll 1
      for (int m = 0; m < U.g; m++) {
        C acc;
        acc.x = 0;
        acc.y = 0;
#pragma unroll
        for (int K = 0; K < RP_GMAX; K++)
          if (K < U.g) acc = cadd(acc, cscale(in[K], A.umat[U.u0 + K * U.g + m]));
        if (phase) acc = tmul(acc, A.dR[(size_t)s_lam[L0 + m] * n + p], false);
        lds[(L0 + m) * n + p] = acc;
      }
    }
    __syncthreads();
  }
  // transforms: thread -> (slot = line, butterfly lane)
  const int tpl = A.tpl;
  const int nslots = RP_THREADS / tpl;
  const int slot = t / tpl, bt = t % tpl;
  auto lidx = [n](int line, int p) { return line * n + p; };
  auto keep = [](int, int, C x) { return x; };
  const bool fwd = A.flags & FFT_FWD, inv = A.flags & FFT_INV;
  const int np = A.plan.npass;
  if (fwd) {
    int Ns = 1;
    for (int pi = 0; pi < np; pi++) {
      const int r = A.plan.radix[pi];
      if (pi == np - 1 && (A.flags & FFT_KIN)) {
        auto kin = [&](int, int p, C x) { return tmul(x, s_kin[p], false); };
        stockham_pass<T>(lds, lidx, nl, slot, nslots, bt, tpl, n, r, Ns, true, s_tw, kin);
      } else {
        stockham_pass<T>(lds, lidx, nl, slot, nslots, bt, tpl, n, r, Ns, true, s_tw, keep);
      }
      Ns *= r;
    }
  }
  if (inv) {
    int Ns = 1;
    for (int pi = 0; pi < np; pi++) {
      const int r = A.plan.radix[pi];
      stockham_pass<T>(lds, lidx, nl, slot, nslots, bt, tpl, n, r, Ns, false, s_tw, keep);
      Ns *= r;
    }
  }
  // D_R/2 and l -> K mixing of mixed units
  if (mix) {
    for (int e = t; e < nu * n; e += RP_THREADS) {
      const int u = e / n, p = e - u * n;
      const RUnit U = A.units[ub + u];
      if (U.u0 < 0) continue;
      const int L0 = s_unit_line0[u];
      C in[RP_GMAX];
#pragma unroll
      for (int m = 0; m < RP_GMAX; m++)
        if (m < U.g) {
          C x = lds[(L0 + m) * n + p];
          if (phase) x = tmul(x, A.dR[(size_t)s_lam[L0 + m] * n + p], false);
          in[m] = x;
        }
#pragma unroll 1
      for (int K = 0; K < U.g; K++) {
        C acc;
        acc.x = 0;
        acc.y = 0;
#pragma unroll
        for (int m = 0; m < RP_GMAX; m++)
          if (m < U.g) acc = cadd(acc, cscale(in[m], A.umat[U.u0 + K * U.g + m]));
        lds[(L0 + K) * n + p] = acc;
      }
    }
    __syncthreads();
  }
  // store; D_R/2 of unmixed lines rides on the store
  for (int e = t; e < nl * n; e += RP_THREADS) {
    const int L = e / n, p = e - L * n;
    C x = lds[L * n + p];
    if (phase && !s_mixed[L]) x = tmul(x, A.dR[(size_t)s_lam[L] * n + p], false);
    A.data[s_off[L] + p] = x;
  }
}

// ---------------------------------------------------------------------------
// r pass: lines along r (stride n1); a work group takes LPB consecutive i of
// one FBR column so global accesses are coalesced across the group.  Twiddles,
// the kinetic factor and (when the group lies in one column) its D_r/2 row
// are staged in LDS.
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Flux surface: psi_s and d psi_s / dr on r = r_s (Gradient3.F:9-167,
// gradients2.cu:583-610) for every DVR column, into slot `slot` of the time block.
// ---------------------------------------------------------------------------
struct StencilArg {
  int h;
  double num[31];  // numerators c_j (integer form for 3..11 points) ...
  double den;      // ... over den: d psi = sum_j num_j (f(s+j) - f(s-j)) / (den dr)
};

// the extraction's arguments; nwg > 0: the first nwg work groups of a
// register-FFT r pass run it instead of transforming lines
template <typename T>
struct FluxExtArgs {
  const typename cx_of<T>::type *psi;
  double2 *bpsi, *bdpsi;
  int n1, n2, ncols, s, slot;
  double dr;
  StencilArg st;
  int nwg;
};

// surface point x (R index i, DVR column k) of the extraction
template <typename T>
__device__ __forceinline__ void flux_extract_at(const FluxExtArgs<T> &E, int x) {
  const int nx = E.n1 * E.ncols;
  if (x >= nx) return;
  const int i = x % E.n1, k = x / E.n1;
  const size_t base = (size_t)i + (size_t)E.n1 * E.n2 * k;
  auto F = [&](int j) {
    const typename cx_of<T>::type v = E.psi[base + (size_t)E.n1 * j];
    return make_double2((double)v.x, (double)v.y);
  };
  double2 acc = make_double2(0.0, 0.0);
  for (int j = E.st.h; j >= 1; j--) {
    const double2 a = F(E.s + j), b = F(E.s - j);
    acc.x += E.st.num[j - 1] * (a.x - b.x);
    acc.y += E.st.num[j - 1] * (a.y - b.y);
  }
  const double den = E.st.den * E.dr;
  E.bpsi[(size_t)E.slot * nx + x] = F(E.s);
  E.bdpsi[(size_t)E.slot * nx + x] = make_double2(acc.x / den, acc.y / den);
}

template <typename T>
__global__ void k_flux_extract(FluxExtArgs<T> E) {
  flux_extract_at<T>(E, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

template <typename T>
struct RLineArgs {
  typename cx_of<T>::type *data;
  const typename tab_of<T>::type *tw;   // exp(-2 pi i e / n2)
  const typename tab_of<T>::type *kin;  // exp(-i T2 dt) / n2
  const typename tab_of<T>::type *dr;   // D_r/2 table [j][n2]
  const int *fcol_j;
  int n1, n2;
  int lpb;
  int nlines;  // n1 * FBR columns
  int flags;
  FftPlanArg plan;
  // the register-FFT kernel's tables (rf_tab; the FP64 ones are tw, kin, dr)
  const typename rf_tab<T>::tw *ltw;
  const typename rf_tab<T>::kin *lkin;
  const typename rf_tab<T>::ph *lph;
  // the previous step's flux-surface extraction, run by the first ext.nwg
  // work groups of the register-FFT pass (nwg 0: none)
  FluxExtArgs<T> ext;
};

template <typename T>
__global__ __launch_bounds__(256) void k_rline(RLineArgs<T> A) {
  typedef typename cx_of<T>::type C;
  typedef typename tab_of<T>::type TB;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = A.n2;
  TB *s_tw = reinterpret_cast<TB *>(smem_raw);
  TB *s_kin = s_tw + n;
  TB *s_dr = s_kin + n;
  C *lds = reinterpret_cast<C *>(s_dr + n);
  const int lpb = A.lpb;
  const int tpl = 256 / lpb;
  const int t = threadIdx.x;
  const int line_local = t % lpb;
  const int bt = t / lpb;
  const int line = blockIdx.x * lpb + line_local;
  const bool active = line < A.nlines;
  const int lc = active ? line : 0;
  const int i = lc % A.n1;
  const int c = lc / A.n1;
  const size_t base = (size_t)c * A.n1 * A.n2 + i;
  const size_t stride = A.n1;
  // one column per group when lpb divides n1 (the host arranges it)
  const int c0 = (blockIdx.x * lpb) / A.n1;
  const int j0 = A.fcol_j[min(c0, A.nlines / A.n1 - 1)];
  const bool phase = A.flags & FFT_PHASE;
  for (int e = t; e < n; e += 256) {
    s_tw[e] = A.tw[e];
    s_kin[e] = A.kin[e];
    if (phase) s_dr[e] = A.dr[(size_t)j0 * n + e];
  }
  __syncthreads();
  auto lidx = [lpb, line_local](int, int p) { return p * lpb + line_local; };
  for (int p = bt; p < n; p += tpl) {
    C x;
    x.x = 0;
    x.y = 0;
    if (active) {
      x = A.data[base + (size_t)p * stride];
      if (phase) x = tmul(x, s_dr[p], false);
    }
    lds[p * lpb + line_local] = x;
  }
  __syncthreads();
  auto keep = [](int, int, C x) { return x; };
  const int np = A.plan.npass;
  if (A.flags & FFT_FWD) {
    int Ns = 1;
    for (int pi = 0; pi < np; pi++) {
      const int r = A.plan.radix[pi];
      if (pi == np - 1 && (A.flags & FFT_KIN)) {
        auto kin = [&](int, int p, C x) { return tmul(x, s_kin[p], false); };
        stockham_pass<T>(lds, lidx, lpb, line_local, lpb, bt, tpl, n, r, Ns, true, s_tw, kin);
      } else {
        stockham_pass<T>(lds, lidx, lpb, line_local, lpb, bt, tpl, n, r, Ns, true, s_tw, keep);
      }
      Ns *= r;
    }
  }
  if (A.flags & FFT_INV) {
    int Ns = 1;
    for (int pi = 0; pi < np; pi++) {
      const int r = A.plan.radix[pi];
      stockham_pass<T>(lds, lidx, lpb, line_local, lpb, bt, tpl, n, r, Ns, false, s_tw, keep);
      Ns *= r;
    }
  }
  for (int p = bt; p < n; p += tpl) {
    if (active) {
      C x = lds[p * lpb + line_local];
      if (phase) x = tmul(x, s_dr[p], false);
      A.data[base + (size_t)p * stride] = x;
    }
  }
}

// ---------------------------------------------------------------------------
// theta transform, FP64 on MFMA (v_mfma_f64_16x16x4f64)
// ---------------------------------------------------------------------------
// Per block: C[2R x N] = A[2R x K] * B[K x N] on psi viewed as a real matrix
// (W and P are real, LegTransform.F:6-59).  Work group: 64 complex rows
// (128 real rows) x 16 CB columns of one block (CB = 8, or 4 when no block
// has more than 64 columns, so that bases of 41..64 functions waste no half
// of the products), 4 waves of 32 real rows, K in chunks of 16.  B is packed
// per matrix as [Kpad][Npad] (row-major, Npad a multiple of 16 CB).
// Prologue PRO: 1 = multiply A by exp(-i V dt); 2 = by dump * exp(-i V dt).
// Epilogue EPI: 1 = multiply the result by the dump d1[i] d2[j].
constexpr int TG_ROWS = 64;    // complex rows per work group
constexpr int TG_KC64 = 16;    // K chunk for FP64
constexpr int TG_LDS64 = 144;  // padded row length (doubles) of the LDS tiles

struct ThetaArgs {
  const void *A;
  void *C;
  const void *B;               // packed matrices, mat_stride elements apart
  long long mat_stride;
  int Kpad, Npad;
  const BlockDesc *blocks;
  int fwd;                     // 1: DVR -> FBR, 0: FBR -> DVR
  int R;                       // n1 * n2
  const double *V;             // FP64 potential (FP64 path)
  const void *expV;            // FP32 phase table (mixed path)
  const long long *dcol_vidx;  // V column start per DVR column
  double dt;
  const double *d1, *d2;
  const float *d1f, *d2f;
  int n1;
  int tpw;                     // 16-row tiles per work group (matrix-stationary kernel)
};

template <int PRO, int EPI, int CB>
__global__ __launch_bounds__(256) void k_theta_f64(ThetaArgs P) {
  static_assert(CB == 4 || CB == 8, "16 or 8 B rows per thread and chunk");
  constexpr int COLS = 16 * CB;
  __shared__ __align__(16) double As[TG_KC64][TG_LDS64];
  __shared__ __align__(16) double Bs[TG_KC64][TG_LDS64];
  const BlockDesc bd = P.blocks[blockIdx.z];
  const int R = P.R;
  const int K = P.fwd ? bd.nd : bd.nf;
  const int N = P.fwd ? bd.nf : bd.nd;
  const size_t acol0 = P.fwd ? bd.dcol0 : bd.fcol0;
  const size_t ccol0 = P.fwd ? bd.fcol0 : bd.dcol0;
  const double2 *A = reinterpret_cast<const double2 *>(P.A) + acol0 * R;
  double2 *Cm = reinterpret_cast<double2 *>(P.C) + ccol0 * R;
  const double *Bp = reinterpret_cast<const double *>(P.B) + (size_t)bd.mat * P.mat_stride;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int row0 = blockIdx.x * TG_ROWS;
  const int col0 = blockIdx.y * COLS;
  if (col0 >= N) return;
  d4v acc[2][CB];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < CB; b++) acc[a][b] = d4v{0.0, 0.0, 0.0, 0.0};

  // The next K chunk is fetched into registers (A with the phase prologue
  // applied) while the current one is on the matrix cores.
  const int cr = t & 63;
  const int r = row0 + cr;
  double2 pa[4];
  double pb[CB];  // TG_KC64 rows x COLS columns over 256 threads
  auto fetch = [&](int k0) {
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int k = k0 + (t >> 6) + 4 * s;
      double2 v = make_double2(0.0, 0.0);
      if (r < R && k < K) {
        v = A[(size_t)r + (size_t)R * k];
        if (PRO) {
          double sn, cs;
          sincos(-P.dt * P.V[P.dcol_vidx[acol0 + k] + r], &sn, &cs);
          v = cmul(v, make_double2(cs, sn));
          if (PRO == 2) {
            const double d = P.d1[r % P.n1] * P.d2[r / P.n1];
            v.x *= d;
            v.y *= d;
          }
        }
      }
      pa[s] = v;
    }
#pragma unroll
    for (int s = 0; s < CB; s++) pb[s] = Bp[(size_t)(k0 + t / COLS + (256 / COLS) * s) * P.Npad + col0 + t % COLS];
  };
  // v_mfma_f64_16x16x4f64 returns D[(lane/16) + 4v][lane%16] in register v.
  // Feeding MFMA row i from tile row 4(i%4) + i/4 makes register v of lane l
  // hold tile row 4(l/16) + v, so each lane ends with two whole complex rows.
  const int arow = 4 * (lane & 3) + ((lane & 15) >> 2);
  fetch(0);
  for (int k0 = 0; k0 < K; k0 += TG_KC64) {
    __syncthreads();  // the previous chunk's operands have been read
#pragma unroll
    for (int s = 0; s < 4; s++) *reinterpret_cast<double2 *>(&As[(t >> 6) + 4 * s][2 * cr]) = pa[s];
#pragma unroll
    for (int s = 0; s < CB; s++) Bs[t / COLS + (256 / COLS) * s][t % COLS] = pb[s];
    __syncthreads();
    if (k0 + TG_KC64 < K) fetch(k0 + TG_KC64);
#pragma unroll
    for (int ks = 0; ks < TG_KC64 / 4; ks++) {
      const int kr = ks * 4 + (lane >> 4);
      double a[2];
#pragma unroll
      for (int rb = 0; rb < 2; rb++) a[rb] = As[kr][wave * 32 + rb * 16 + arow];
#pragma unroll
      for (int cb = 0; cb < CB; cb++) {
        const double b = Bs[kr][cb * 16 + (lane & 15)];
#pragma unroll
        for (int rb = 0; rb < 2; rb++) acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb], b, acc[rb][cb], 0, 0, 0);
      }
    }
  }
  // lane l holds tile rows 4(l/16)+0..3 (complex rows 2(l/16), 2(l/16)+1) of column l%16
#pragma unroll
  for (int rb = 0; rb < 2; rb++) {
    const int rr = wave * 32 + rb * 16 + 4 * (lane >> 4);
    const int r = row0 + rr / 2;
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
      const int col = col0 + cb * 16 + (lane & 15);
      if (col >= N) continue;
      double2 c0 = make_double2(acc[rb][cb][0], acc[rb][cb][1]);
      double2 c1 = make_double2(acc[rb][cb][2], acc[rb][cb][3]);
      if (EPI) {
        if (r < R) {
          const double d = P.d1[r % P.n1] * P.d2[r / P.n1];
          c0.x *= d;
          c0.y *= d;
        }
        if (r + 1 < R) {
          const double d = P.d1[(r + 1) % P.n1] * P.d2[(r + 1) / P.n1];
          c1.x *= d;
          c1.y *= d;
        }
      }
      const size_t x = (size_t)r + (size_t)R * col;
      if (r < R) Cm[x] = c0;
      if (r + 1 < R) Cm[x + 1] = c1;
    }
  }
}

// ---------------------------------------------------------------------------
// theta transform, mixed precision on MFMA (v_mfma_f32_16x16x32_bf16)
// ---------------------------------------------------------------------------
// Each FP32 operand is split into NS bf16 pieces a = a0 + a1 (+ a2); the
// products a_p b_q with p + q <= NS - 1 are accumulated in FP32 (NS = 2: the
// hi*hi + hi*lo + lo*hi scheme; NS = 3: six products, FP32-level accuracy).
// B is pre-split per matrix as [piece][Npad][Kpad] (k contiguous).
// NCB = 16-column tiles per work group (8: 128 columns, 4: 64 columns).
constexpr int TG_KC16 = 32;
constexpr int TG_LDS16 = 40;  // padded k length (bf16) of an LDS row

__device__ __forceinline__ __bf16 f2bf(float x) { return (__bf16)x; }

template <int NS, int NCB, int PRO, int EPI>
__global__ __launch_bounds__(256) void k_theta_bf16(ThetaArgs P) {
  __shared__ __align__(16) __bf16 As[NS][128][TG_LDS16];
  __shared__ __align__(16) __bf16 Bs[NS][NCB * 16][TG_LDS16];
  const BlockDesc bd = P.blocks[blockIdx.z];
  const int R = P.R;
  const int K = P.fwd ? bd.nd : bd.nf;
  const int N = P.fwd ? bd.nf : bd.nd;
  const size_t acol0 = P.fwd ? bd.dcol0 : bd.fcol0;
  const size_t ccol0 = P.fwd ? bd.fcol0 : bd.dcol0;
  const float2 *A = reinterpret_cast<const float2 *>(P.A) + acol0 * R;
  float2 *Cm = reinterpret_cast<float2 *>(P.C) + ccol0 * R;
  const __bf16 *Bsplit = reinterpret_cast<const __bf16 *>(P.B) + (size_t)bd.mat * P.mat_stride;
  const float2 *expV = reinterpret_cast<const float2 *>(P.expV);
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int row0 = blockIdx.x * TG_ROWS;
  const int col0 = blockIdx.y * NCB * 16;
  if (col0 >= N) return;
  f4v acc[2][NCB];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < NCB; b++) acc[a][b] = f4v{0.f, 0.f, 0.f, 0.f};

  // The next K chunk is fetched into registers while the current one is on
  // the matrix cores: A as 8 complex values (row cr, k = kg*8..kg*8+7) with
  // the phase prologue applied, B as bf16 vectors.
  const int cr = t & 63;
  const int kg = t >> 6;  // 0..3
  const int r = row0 + cr;
  constexpr int BV = (NS * NCB * 16 * 2 + 255) / 256;  // B vectors (2 x bf8v) per thread
  float2 pa[8];
  bf8v pb[BV][2];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int k = k0 + kg * 8 + e;
      float2 v = make_float2(0.f, 0.f);
      if (r < R && k < K) {
        v = A[(size_t)r + (size_t)R * k];
        if (PRO) {
          v = cmul(v, expV[P.dcol_vidx[acol0 + k] + r]);
          if (PRO == 2) {
            const float d = P.d1f[r % P.n1] * P.d2f[r / P.n1];
            v.x *= d;
            v.y *= d;
          }
        }
      }
      pa[e] = v;
    }
#pragma unroll
    for (int i = 0; i < BV; i++) {
      const int u = t + i * 256;
      if (u < NS * NCB * 32) {
        const int p = u / (NCB * 32);
        const int rem = u % (NCB * 32);
        const __bf16 *src = Bsplit + ((size_t)p * P.Npad + col0 + (rem >> 1)) * P.Kpad + k0 + (rem & 1) * 16;
        pb[i][0] = *reinterpret_cast<const bf8v *>(src);
        pb[i][1] = *reinterpret_cast<const bf8v *>(src + 8);
      }
    }
  };
  auto stash = [&]() {
    float re[8], im[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
      re[e] = pa[e].x;
      im[e] = pa[e].y;
    }
#pragma unroll
    for (int p = 0; p < NS; p++) {
      bf8v hr, hi;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const __bf16 a = f2bf(re[e]);
        const __bf16 b = f2bf(im[e]);
        hr[e] = a;
        hi[e] = b;
        re[e] -= (float)a;
        im[e] -= (float)b;
      }
      *reinterpret_cast<bf8v *>(&As[p][2 * cr][kg * 8]) = hr;
      *reinterpret_cast<bf8v *>(&As[p][2 * cr + 1][kg * 8]) = hi;
    }
#pragma unroll
    for (int i = 0; i < BV; i++) {
      const int u = t + i * 256;
      if (u < NS * NCB * 32) {
        const int p = u / (NCB * 32);
        const int rem = u % (NCB * 32);
        *reinterpret_cast<bf8v *>(&Bs[p][rem >> 1][(rem & 1) * 16]) = pb[i][0];
        *reinterpret_cast<bf8v *>(&Bs[p][rem >> 1][(rem & 1) * 16 + 8]) = pb[i][1];
      }
    }
  };

  fetch(0);
  for (int k0 = 0; k0 < K; k0 += TG_KC16) {
    __syncthreads();  // the previous chunk's fragments have been read
    stash();
    __syncthreads();
    if (k0 + TG_KC16 < K) fetch(k0 + TG_KC16);
    bf8v a[NS][2];
#pragma unroll
    for (int p = 0; p < NS; p++)
#pragma unroll
      for (int rb = 0; rb < 2; rb++)
        a[p][rb] = *reinterpret_cast<const bf8v *>(&As[p][wave * 32 + rb * 16 + (lane & 15)][(lane >> 4) * 8]);
#pragma unroll
    for (int cb = 0; cb < NCB; cb++) {
      bf8v b[NS];
#pragma unroll
      for (int q = 0; q < NS; q++)
        b[q] = *reinterpret_cast<const bf8v *>(&Bs[q][cb * 16 + (lane & 15)][(lane >> 4) * 8]);
      // small products first
#pragma unroll
      for (int s = NS - 1; s >= 0; s--)
#pragma unroll
        for (int p = 0; p <= s; p++) {
          const int q = s - p;
#pragma unroll
          for (int rb = 0; rb < 2; rb++)
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p][rb], b[q], acc[rb][cb], 0, 0, 0);
        }
    }
  }
#pragma unroll
  for (int rb = 0; rb < 2; rb++) {
    const int rr = wave * 32 + rb * 16 + 4 * (lane >> 4);
    const int rw = row0 + rr / 2;
#pragma unroll
    for (int cb = 0; cb < NCB; cb++) {
      const int col = col0 + cb * 16 + (lane & 15);
      if (col >= N) continue;
      float2 c0 = make_float2(acc[rb][cb][0], acc[rb][cb][1]);
      float2 c1 = make_float2(acc[rb][cb][2], acc[rb][cb][3]);
      if (EPI) {
        if (rw < R) {
          const float d = P.d1f[rw % P.n1] * P.d2f[rw / P.n1];
          c0.x *= d;
          c0.y *= d;
        }
        if (rw + 1 < R) {
          const float d = P.d1f[(rw + 1) % P.n1] * P.d2f[(rw + 1) / P.n1];
          c1.x *= d;
          c1.y *= d;
        }
      }
      const size_t x = (size_t)rw + (size_t)R * col;
      if (rw < R) Cm[x] = c0;
      if (rw + 1 < R) Cm[x + 1] = c1;
    }
  }
}

// ---------------------------------------------------------------------------
// theta transform, mixed precision, matrix-stationary (n_theta <= 128)
// ---------------------------------------------------------------------------
// The product is taken transposed, C^T[n][row] = sum_k M[k][n] psi[row][k]:
// the theta matrix (three bf16 pieces, [piece][n][k]) is the MFMA A operand
// and stays in LDS for the whole work group; psi goes from HBM straight into
// registers, one k-step (32 columns) ahead of the matrix cores, and is split
// into bf16 pieces there.  Lane l owns grid row (l % 16) of a 16-row tile, so
// its output column n is a 128-byte coalesced store over the 16 lanes, and the
// phase / dump factors are per-lane scalars.
// Work: a work group takes tiles [blockIdx.x * tpw, +tpw) of block blockIdx.z;
// its waves take every W-th tile.
// Raw buffer access (gfx9 descriptor, stride 0).  An offset at or beyond the
// descriptor's byte count reads as zero and drops a store, which makes the
// edge tiles branch-free: masked elements get offset TM_OOB.
constexpr unsigned TM_OOB = 0x80000000u;
constexpr int TMX_THREADS = 512;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float2 buf_ld_f2(__amdgpu_buffer_rsrc_t r, unsigned off) {
  const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
  return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// 8- or 16-byte element load through a descriptor (float2, double2, float4)
template <typename V>
__device__ __forceinline__ V buf_ld(__amdgpu_buffer_rsrc_t r, unsigned off) {
  static_assert(sizeof(V) == 8 || sizeof(V) == 16, "8- or 16-byte elements");
  if constexpr (sizeof(V) == 8) {
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
  } else {
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
  }
}
__device__ __forceinline__ void buf_st_f2(float2 x, __amdgpu_buffer_rsrc_t r, unsigned off) {
  u32x2 v;
  v.x = __float_as_uint(x.x);
  v.y = __float_as_uint(x.y);
  __builtin_amdgcn_raw_buffer_store_b64(v, r, (int)off, 0, 0);
}

template <int KS, int NT, int PRO, int EPI>
__global__ __launch_bounds__(512) void k_theta_mx(ThetaArgs P) {
  extern __shared__ __align__(16) unsigned char tm_lds[];
  constexpr int KP = KS * 32, NP = NT * 16, LD = KP + 8;
  __bf16 *Ms = reinterpret_cast<__bf16 *>(tm_lds);  // [3][NP][LD]
  const BlockDesc bd = P.blocks[blockIdx.z];
  const int R = P.R;
  const int K = P.fwd ? bd.nd : bd.nf;
  const int N = P.fwd ? bd.nf : bd.nd;
  const size_t acol0 = P.fwd ? bd.dcol0 : bd.fcol0;
  const size_t ccol0 = P.fwd ? bd.fcol0 : bd.dcol0;
  // block-local descriptors: psi in, phase columns (contiguous per block), out
  const __amdgpu_buffer_rsrc_t rA = buf_rsrc(reinterpret_cast<const float2 *>(P.A) + acol0 * R, (unsigned)K * R * 8u);
  const __amdgpu_buffer_rsrc_t rV =
      buf_rsrc(PRO ? reinterpret_cast<const float2 *>(P.expV) + P.dcol_vidx[acol0] : nullptr, PRO ? (unsigned)K * R * 8u : 0u);
  const __amdgpu_buffer_rsrc_t rC = buf_rsrc(reinterpret_cast<float2 *>(P.C) + ccol0 * R, (unsigned)N * R * 8u);
  // EPI == 2 (FBR -> DVR): the output columns times conj(exp(-i V dt))
  const __amdgpu_buffer_rsrc_t rVo =
      buf_rsrc(EPI == 2 ? reinterpret_cast<const float2 *>(P.expV) + P.dcol_vidx[ccol0] : nullptr,
               EPI == 2 ? (unsigned)N * R * 8u : 0u);
  {
    const __bf16 *Mg = reinterpret_cast<const __bf16 *>(P.B) + (size_t)bd.mat * P.mat_stride;
    constexpr int QV = KP / 8;  // bf8v per matrix row
    for (int u = threadIdx.x; u < 3 * NP * QV; u += blockDim.x) {
      const int pn = u / QV, kq = u % QV;
      *reinterpret_cast<bf8v *>(&Ms[pn * LD + kq * 8]) = *reinterpret_cast<const bf8v *>(&Mg[(size_t)pn * KP + kq * 8]);
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, W = blockDim.x >> 6;
  const int rl = lane & 15, kg = lane >> 4;
  const int ntiles = (R + 15) >> 4;
  const int t1 = min(ntiles, (int)(blockIdx.x + 1) * P.tpw);
  float2 ra[8], rv[8];
  auto fetch = [&](int tile, int s) {
    const int row = tile * 16 + rl;
    const int k0 = s * 32 + kg * 8;
    // a row past the end starts beyond the descriptor (K * R * 8 < 2^31)
    const unsigned base = (row < R ? (unsigned)row * 8u : TM_OOB) + (unsigned)(R * k0) * 8u;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const unsigned off = k0 + e < K ? base + (unsigned)(R * e) * 8u : TM_OOB;
      ra[e] = buf_ld_f2(rA, off);
      if (PRO) rv[e] = buf_ld_f2(rV, off);
    }
  };
  int tile = blockIdx.x * P.tpw + wave;
  if (tile < t1) fetch(tile, 0);
  for (; tile < t1; tile += W) {
    const int row = tile * 16 + rl;
    float dfac = 1.f;
    if ((PRO == 2 || EPI == 1) && row < R) dfac = P.d1f[row % P.n1] * P.d2f[row / P.n1];
    f4v acc[2][NT];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < NT; b++) acc[a][b] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int s = 0; s < KS; s++) {
      bf8v bre[3], bim[3];
      {
        float re[8], im[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
          float2 v = ra[e];
          if (PRO) v = cmul(v, rv[e]);
          if (PRO == 2) {
            v.x *= dfac;
            v.y *= dfac;
          }
          re[e] = v.x;
          im[e] = v.y;
        }
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
          for (int e = 0; e < 8; e++) {
            const __bf16 a = f2bf(re[e]);
            const __bf16 b = f2bf(im[e]);
            bre[q][e] = a;
            bim[q][e] = b;
            re[e] -= (float)a;
            im[e] -= (float)b;
          }
      }
      {
        const bool last = s + 1 == KS;
        const int ft = last ? tile + W : tile;
        if (ft < t1) fetch(ft, last ? 0 : s + 1);
      }
#pragma unroll
      for (int nt = 0; nt < NT; nt++) {
        bf8v a[3];
#pragma unroll
        for (int p = 0; p < 3; p++) a[p] = *reinterpret_cast<const bf8v *>(&Ms[(p * NP + nt * 16 + rl) * LD + s * 32 + kg * 8]);
        // small products first
#pragma unroll
        for (int d = 2; d >= 0; d--)
#pragma unroll
          for (int p = 0; p <= d; p++) {
            acc[0][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], bre[d - p], acc[0][nt], 0, 0, 0);
            acc[1][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], bim[d - p], acc[1][nt], 0, 0, 0);
          }
      }
    }
#pragma unroll
    for (int nt = 0; nt < NT; nt++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int n = nt * 16 + 4 * kg + v;
        float2 c = make_float2(acc[0][nt][v], acc[1][nt][v]);
        const unsigned off = (row < R && n < N) ? (unsigned)(row + R * n) * 8u : TM_OOB;
        if (EPI == 1) {
          c.x *= dfac;
          c.y *= dfac;
        }
        if (EPI == 2) {
          const float2 w = buf_ld_f2(rVo, off);
          c = make_float2(c.x * w.x + c.y * w.y, c.y * w.x - c.x * w.y);
        }
        buf_st_f2(c, rC, off);
      }
  }
}

// Fused FBR -> DVR -> FBR theta kernel (mixed precision, three BF16 pieces,
// matrix-stationary like k_theta_mx) for the CPU-propagator step order: the
// theta-backward GEMM of step n, the dump, the DVR values of the flux-surface
// rows r_lo..r_hi (written to psi for k_flux_extract; no other DVR value is
// stored), the V phase and the theta-forward GEMM of step n + 1, in place on
// the FBR array.  Each wave owns 16-row tiles; lane (rl, kg) holds row rl.
// The backward GEMM leaves DVR column nt * 16 + 4 kg + v in lane (rl, kg); the
// forward GEMM wants columns 32 s + 8 kg + e.  The values pass through a
// per-wave LDS tile to get there.  (Feeding them straight from the registers
// with the forward matrix's k index permuted instead groups different columns
// into each MFMA k-block, and that grouping was measured to gain norm
// coherently, 6e-8 per step; the natural grouping of the unfused kernels
// does not.)
// Blocks that share the theta matrices and the potential (the J blocks of one
// K) form a group.  A work group takes one work item: a range of tiles of one
// group; for each tile a wave loads the V phase once and runs every block of
// the group through it, so V is read from HBM once per group, not per block.
struct ThetaFusedArgs {
  float2 *L;                   // FBR array, in and out
  float2 *psi;                 // DVR array: the flux rows are written
  const void *Bb, *Bf;         // packed backward and forward matrices, mat_stride apart
  long long mat_stride;
  const BlockDesc *blocks;
  int R;                       // n1 * n2
  const float2 *expV;          // exp(-i V dt) per V column
  const long long *dcol_vidx;
  const float *d1f, *d2f;      // dump factors (nullptr: no dump)
  int n1;
  int r_lo, r_hi;              // psi rows written (r_lo > r_hi: none)
  int pre_dump;                // 1: the rows as before the dump (CUDA propagator order), 0: after it
  const int4 *items;           // per work group: first block of its group, blocks, tile begin, tile end
};

template <int KS, int NTH>
__global__ __launch_bounds__(NTH) void k_theta_fused(ThetaFusedArgs P) {
  extern __shared__ __align__(16) unsigned char tm_lds[];
  constexpr int KP = KS * 32, NT = 2 * KS, NP = NT * 16, LD = KP + 8, XS = KP + 1;
  __bf16 *Mb = reinterpret_cast<__bf16 *>(tm_lds);  // [3][NP][LD], row n = DVR column, k = FBR column
  __bf16 *Mf = Mb + 3 * NP * LD;                     // [3][NP][LD], row n = FBR column, k = DVR column
  // per wave, the tile's DVR values [16][XS] in natural column order
  float2 *Xw = reinterpret_cast<float2 *>(Mf + 3 * NP * LD) + (threadIdx.x >> 6) * 16 * XS;
  const int4 item = P.items[blockIdx.x];
  const int gb0 = item.x, gnb = item.y, t0 = item.z, t1 = item.w;
  const BlockDesc bd = P.blocks[gb0];  // the group shares mat, nd, nf and the V columns
  const int R = P.R, nd = bd.nd, nf = bd.nf;
  auto rL_of = [&](int b) { return buf_rsrc(P.L + (size_t)P.blocks[gb0 + b].fcol0 * R, (unsigned)nf * R * 8u); };
  auto rP_of = [&](int b) { return buf_rsrc(P.psi + (size_t)P.blocks[gb0 + b].dcol0 * R, (unsigned)nd * R * 8u); };
  const __amdgpu_buffer_rsrc_t rV = buf_rsrc(P.expV + P.dcol_vidx[bd.dcol0], (unsigned)nd * R * 8u);
  {
    const __bf16 *Gb = reinterpret_cast<const __bf16 *>(P.Bb) + (size_t)bd.mat * P.mat_stride;
    const __bf16 *Gf = reinterpret_cast<const __bf16 *>(P.Bf) + (size_t)bd.mat * P.mat_stride;
    constexpr int QV = KP / 8;
    for (int u = threadIdx.x; u < 3 * NP * QV; u += blockDim.x) {
      const int pn = u / QV, kq = u % QV;
      *reinterpret_cast<bf8v *>(&Mb[pn * LD + kq * 8]) = *reinterpret_cast<const bf8v *>(&Gb[(size_t)pn * KP + kq * 8]);
      *reinterpret_cast<bf8v *>(&Mf[pn * LD + kq * 8]) = *reinterpret_cast<const bf8v *>(&Gf[(size_t)pn * KP + kq * 8]);
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, W = blockDim.x >> 6;
  const int rl = lane & 15, kg = lane >> 4;
  float2 ra[8];
  auto fetch = [&](const __amdgpu_buffer_rsrc_t &rd, int tile, int s) {
    const int row = tile * 16 + rl;
    const int k0 = s * 32 + kg * 8;
    const unsigned base = (row < R ? (unsigned)row * 8u : TM_OOB) + (unsigned)(R * k0) * 8u;
#pragma unroll
    for (int e = 0; e < 8; e++) ra[e] = buf_ld_f2(rd, k0 + e < nf ? base + (unsigned)(R * e) * 8u : TM_OOB);
  };
  // three BF16 pieces of 8 FP32 values
  auto split3 = [](const float *x, bf8v *out) {
    float r[8];
#pragma unroll
    for (int e = 0; e < 8; e++) r[e] = x[e];
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const __bf16 b = f2bf(r[e]);
        out[q][e] = b;
        r[e] -= (float)b;
      }
  };
  // this wave's (tile, block) sequence: tiles t0 + wave, + W, ...; every
  // block of the group for each tile
  int tile = t0 + wave, bi = 0;
  __amdgpu_buffer_rsrc_t rL = rL_of(0);
  if (tile < t1) fetch(rL, tile, 0);
  float2 wv[NT][4];
  while (tile < t1) {
    const int row = tile * 16 + rl;
    const bool rin = row < R;
    const bool lastb = bi + 1 == gnb;
    const int ntile = lastb ? tile + W : tile, nbi = lastb ? 0 : bi + 1;
    const __amdgpu_buffer_rsrc_t rLn = rL_of(nbi);
    // V phase of this tile's DVR columns, loaded once per tile ahead of the
    // first block's backward GEMM
    if (bi == 0) {
#pragma unroll
      for (int nt = 0; nt < NT; nt++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int n = nt * 16 + 4 * kg + v;
          wv[nt][v] = buf_ld_f2(rV, (rin && n < nd) ? (unsigned)(row + R * n) * 8u : TM_OOB);
        }
    }
    f4v acc[2][NT];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < NT; b++) acc[a][b] = f4v{0.f, 0.f, 0.f, 0.f};
    // backward GEMM: DVR[n] = sum_k Pb[n][k] FBR[k]
#pragma unroll 1
    for (int s = 0; s < KS; s++) {
      bf8v bre[3], bim[3];
      {
        float re[8], im[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
          re[e] = ra[e].x;
          im[e] = ra[e].y;
        }
        split3(re, bre);
        split3(im, bim);
      }
      if (s + 1 < KS) fetch(rL, tile, s + 1);
      else if (ntile < t1) fetch(rLn, ntile, 0);
#pragma unroll
      for (int nt = 0; nt < NT; nt++) {
        bf8v a[3];
#pragma unroll
        for (int p = 0; p < 3; p++) a[p] = *reinterpret_cast<const bf8v *>(&Mb[(p * NP + nt * 16 + rl) * LD + s * 32 + kg * 8]);
#pragma unroll
        for (int d = 2; d >= 0; d--)
#pragma unroll
          for (int p = 0; p <= d; p++) {
            acc[0][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], bre[d - p], acc[0][nt], 0, 0, 0);
            acc[1][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], bim[d - p], acc[1][nt], 0, 0, 0);
          }
      }
    }
    // DVR domain: dump, psi rows (before or after the dump), V phase; then the
    // pieces of the forward B operand
    float dfac = 1.f;
    if (P.d1f && rin) dfac = P.d1f[row % P.n1] * P.d2f[row / P.n1];
    const int rr = row / P.n1;
    const bool frow = rin && rr >= P.r_lo && rr <= P.r_hi;
    const __amdgpu_buffer_rsrc_t rP = rP_of(bi);
#pragma unroll
    for (int nt = 0; nt < NT; nt++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int n = nt * 16 + 4 * kg + v;
        const float2 c = make_float2(acc[0][nt][v] * dfac, acc[1][nt][v] * dfac);
        if (frow)
          buf_st_f2(P.pre_dump ? make_float2(acc[0][nt][v], acc[1][nt][v]) : c, rP,
                    n < nd ? (unsigned)(row + R * n) * 8u : TM_OOB);
        const float2 w = wv[nt][v];
        Xw[rl * XS + n] = make_float2(c.x * w.x - c.y * w.y, c.x * w.y + c.y * w.x);
      }
    // the wave's own LDS rows: its writes land before its reads, which take
    // the columns in the forward GEMM's order
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // forward GEMM: FBR[m] = sum_n Wf[n][m] DVR[n]; each k-step's B operand
    // is split from the LDS tile just before its products
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < NT; b++) acc[a][b] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s2 = 0; s2 < KS; s2++) {
      bf8v fre[3], fim[3];
      {
        float re[8], im[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const float2 x = Xw[rl * XS + s2 * 32 + kg * 8 + e];
          re[e] = x.x;
          im[e] = x.y;
        }
        split3(re, fre);
        split3(im, fim);
      }
      if (s2 + 1 == KS) {
        // reads done before the next tile's writes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
#pragma unroll
      for (int nt = 0; nt < NT; nt++) {
        bf8v a[3];
#pragma unroll
        for (int p = 0; p < 3; p++) a[p] = *reinterpret_cast<const bf8v *>(&Mf[(p * NP + nt * 16 + rl) * LD + s2 * 32 + kg * 8]);
#pragma unroll
        for (int d = 2; d >= 0; d--)
#pragma unroll
          for (int p = 0; p <= d; p++) {
            acc[0][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], fre[d - p], acc[0][nt], 0, 0, 0);
            acc[1][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], fim[d - p], acc[1][nt], 0, 0, 0);
          }
      }
    }
#pragma unroll
    for (int nt = 0; nt < NT; nt++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int m = nt * 16 + 4 * kg + v;
        buf_st_f2(make_float2(acc[0][nt][v], acc[1][nt][v]), rL, (rin && m < nf) ? (unsigned)(row + R * m) * 8u : TM_OOB);
      }
    tile = ntile;
    bi = nbi;
    rL = rLn;
  }
}
template <int KS, int NTH>
constexpr size_t theta_fused_lds() {
  return (size_t)2 * 3 * (2 * KS * 16) * (KS * 32 + 8) * 2 + (size_t)(NTH / 64) * 16 * (KS * 32 + 1) * 8;
}

// The fused theta step on the parity halves of the two transforms.  Every
// theta grid is symmetric, x_{nd-1-i} = -x_i, and FBR function c has parity
// (-1)^c (P_l(-x) = (-1)^l P_l(x); (-1)^(j-K) for the associated functions of a
// K block), so with i' = nd - 1 - i and i < NH = ceil(nd / 2):
//   backward  E_i = sum_{c even} Pb[c][i] F_c,   O_i = sum_{c odd} Pb[c][i] F_c,
//             DVR_i = E_i + O_i,   DVR_i' = E_i - O_i;
//   forward   F_c = sum_{i < NH} Wf[i][c] (DVR_i + DVR_i')   (c even),
//             F_c = sum_{i < NH} Wf[i][c] (DVR_i - DVR_i')   (c odd).
// An odd grid's middle node is its own mirror: its odd backward row is zero
// and its even forward row is halved on the host (exact), so the same formulas
// hold.  Each half product is half as long and half as wide as the full one:
// half the MFMA work of k_theta_fused and a quarter of its matrix LDS per
// half.  The host packs the four half matrices of a set as [4][3][NP][KP]
// (backward even, backward odd, forward even, forward odd; three BF16 pieces;
// k contiguous) after checking the symmetry numerically.  Everything else --
// work items, dump, flux rows, V phase, both propagator orders -- is as in
// k_theta_fused; the per-wave LDS tile holds the sums (columns 0 .. KP - 1)
// and differences (KP .. 2 KP - 1) for the forward products -- or, for half
// bases of 33..64 (KSH = 2, whose four half matrices take 110.6 KB), the 32
// columns of one forward k-step at a time (four rounds), so that the kernel
// fits a CU's LDS with eight waves.
template <int KSH, int NTH>
__global__ __launch_bounds__(NTH) void k_theta_fused_par(ThetaFusedArgs P) {
  extern __shared__ __align__(16) unsigned char tm_lds[];
  constexpr int KP = KSH * 32, NT = 2 * KSH, NP = NT * 16, LD = KP + 8, XS = (KSH > 1 ? 32 : 2 * KP) + 1;
  __bf16 *M = reinterpret_cast<__bf16 *>(tm_lds);  // [4][3][NP][LD]
  float2 *Xw = reinterpret_cast<float2 *>(M + 4 * 3 * NP * LD) + (threadIdx.x >> 6) * 16 * XS;
  const int4 item = P.items[blockIdx.x];
  const int gb0 = item.x, gnb = item.y, t0 = item.z, t1 = item.w;
  const BlockDesc bd = P.blocks[gb0];  // the group shares mat, nd, nf and the V columns
  const int R = P.R, nd = bd.nd, nf = bd.nf;
  const int NH = (nd + 1) >> 1, CE = (nf + 1) >> 1, CO = nf >> 1;
  auto rL_of = [&](int b) { return buf_rsrc(P.L + (size_t)P.blocks[gb0 + b].fcol0 * R, (unsigned)nf * R * 8u); };
  auto rP_of = [&](int b) { return buf_rsrc(P.psi + (size_t)P.blocks[gb0 + b].dcol0 * R, (unsigned)nd * R * 8u); };
  const __amdgpu_buffer_rsrc_t rV = buf_rsrc(P.expV + P.dcol_vidx[bd.dcol0], (unsigned)nd * R * 8u);
  {
    const __bf16 *G = reinterpret_cast<const __bf16 *>(P.Bb) + (size_t)bd.mat * P.mat_stride;
    constexpr int QV = KP / 8;
    for (int u = threadIdx.x; u < 4 * 3 * NP * QV; u += blockDim.x) {
      const int pn = u / QV, kq = u % QV;
      *reinterpret_cast<bf8v *>(&M[pn * LD + kq * 8]) = *reinterpret_cast<const bf8v *>(&G[(size_t)pn * KP + kq * 8]);
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, W = blockDim.x >> 6;
  const int rl = lane & 15, kg = lane >> 4;
  float2 ra[2][8];  // [parity][e]: FBR column 2 (32 s + 8 kg + e) + parity
  auto fetch = [&](const __amdgpu_buffer_rsrc_t &rd, int tile, int s) {
    const int row = tile * 16 + rl;
    const int k0 = s * 32 + kg * 8;
    const unsigned base = row < R ? (unsigned)row * 8u : TM_OOB;
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int c = 2 * (k0 + e) + q;
        ra[q][e] = buf_ld_f2(rd, c < nf ? base + (unsigned)(R * c) * 8u : TM_OOB);
      }
  };
  auto split3 = [](const float *x, bf8v *out) {
    float r[8];
#pragma unroll
    for (int e = 0; e < 8; e++) r[e] = x[e];
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const __bf16 b = f2bf(r[e]);
        out[q][e] = b;
        r[e] -= (float)b;
      }
  };
  int tile = t0 + wave, bi = 0;
  __amdgpu_buffer_rsrc_t rL = rL_of(0);
  if (tile < t1) fetch(rL, tile, 0);
  float2 wv[2][NT][4];  // V phase at columns i and i'
  while (tile < t1) {
    const int row = tile * 16 + rl;
    const bool rin = row < R;
    const bool lastb = bi + 1 == gnb;
    const int ntile = lastb ? tile + W : tile, nbi = lastb ? 0 : bi + 1;
    const __amdgpu_buffer_rsrc_t rLn = rL_of(nbi);
    if (bi == 0) {
#pragma unroll
      for (int nt = 0; nt < NT; nt++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int i = nt * 16 + 4 * kg + v;
          const bool ok = rin && i < NH;
          wv[0][nt][v] = buf_ld_f2(rV, ok ? (unsigned)(row + R * i) * 8u : TM_OOB);
          wv[1][nt][v] = buf_ld_f2(rV, ok ? (unsigned)(row + R * (nd - 1 - i)) * 8u : TM_OOB);
        }
    }
    f4v acc[2][2][NT];  // [parity][re, im][n tile]
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < NT; b++) acc[q][a][b] = f4v{0.f, 0.f, 0.f, 0.f};
    // backward: E and O over the even and odd FBR columns
#pragma unroll 1
    for (int s = 0; s < KSH; s++) {
      bf8v bre[2][3], bim[2][3];
#pragma unroll
      for (int q = 0; q < 2; q++) {
        float re[8], im[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
          re[e] = ra[q][e].x;
          im[e] = ra[q][e].y;
        }
        split3(re, bre[q]);
        split3(im, bim[q]);
      }
      if (s + 1 < KSH) fetch(rL, tile, s + 1);
      else if (ntile < t1) fetch(rLn, ntile, 0);
#pragma unroll
      for (int q = 0; q < 2; q++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
          bf8v a[3];
#pragma unroll
          for (int p = 0; p < 3; p++)
            a[p] = *reinterpret_cast<const bf8v *>(&M[((q * 3 + p) * NP + nt * 16 + rl) * LD + s * 32 + kg * 8]);
#pragma unroll
          for (int d = 2; d >= 0; d--)
#pragma unroll
            for (int p = 0; p <= d; p++) {
              acc[q][0][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], bre[q][d - p], acc[q][0][nt], 0, 0, 0);
              acc[q][1][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], bim[q][d - p], acc[q][1][nt], 0, 0, 0);
            }
        }
    }
    // DVR domain: the pair (i, i'), dump, psi rows, V phase; sums and
    // differences into the wave's tile
    float dfac = 1.f;
    if (P.d1f && rin) dfac = P.d1f[row % P.n1] * P.d2f[row / P.n1];
    const int rr = row / P.n1;
    const bool frow = rin && rr >= P.r_lo && rr <= P.r_hi;
    const __amdgpu_buffer_rsrc_t rP = rP_of(bi);
    float2 pp[2][NT][4];  // the pair's values after dump and V phase
#pragma unroll
    for (int nt = 0; nt < NT; nt++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = nt * 16 + 4 * kg + v, im = nd - 1 - i;
        const float2 d1 = make_float2(acc[0][0][nt][v] + acc[1][0][nt][v], acc[0][1][nt][v] + acc[1][1][nt][v]);
        const float2 d2 = make_float2(acc[0][0][nt][v] - acc[1][0][nt][v], acc[0][1][nt][v] - acc[1][1][nt][v]);
        const float2 c1 = make_float2(d1.x * dfac, d1.y * dfac), c2 = make_float2(d2.x * dfac, d2.y * dfac);
        if (frow) {
          buf_st_f2(P.pre_dump ? d1 : c1, rP, i < NH ? (unsigned)(row + R * i) * 8u : TM_OOB);
          buf_st_f2(P.pre_dump ? d2 : c2, rP, i < nd / 2 ? (unsigned)(row + R * im) * 8u : TM_OOB);
        }
        const float2 w1 = wv[0][nt][v], w2 = wv[1][nt][v];
        pp[0][nt][v] = make_float2(c1.x * w1.x - c1.y * w1.y, c1.x * w1.y + c1.y * w1.x);
        pp[1][nt][v] = make_float2(c2.x * w2.x - c2.y * w2.y, c2.x * w2.y + c2.y * w2.x);
      }
    auto wave_sync = [] {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    auto sum_or_diff = [&](int q, int nt, int v) {
      const float2 p1 = pp[0][nt][v], p2 = pp[1][nt][v];
      return q == 0 ? make_float2(p1.x + p2.x, p1.y + p2.y) : make_float2(p1.x - p2.x, p1.y - p2.y);
    };
    // forward: even FBR columns from the sums (q = 0), odd from the
    // differences (q = 1); k-step s2's B operand is tile columns col0 + 8 kg + e
    auto tile_split = [&](int col0, bf8v *fre, bf8v *fim) {
      float re[8], im[8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float2 x = Xw[rl * XS + col0 + kg * 8 + e];
        re[e] = x.x;
        im[e] = x.y;
      }
      split3(re, fre);
      split3(im, fim);
    };
    auto forward_products = [&](int q, int s2, const bf8v *fre, const bf8v *fim) {
#pragma unroll
      for (int nt = 0; nt < NT; nt++) {
        bf8v a[3];
#pragma unroll
        for (int p = 0; p < 3; p++)
          a[p] = *reinterpret_cast<const bf8v *>(&M[(((2 + q) * 3 + p) * NP + nt * 16 + rl) * LD + s2 * 32 + kg * 8]);
#pragma unroll
        for (int d = 2; d >= 0; d--)
#pragma unroll
          for (int p = 0; p <= d; p++) {
            acc[q][0][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], fre[d - p], acc[q][0][nt], 0, 0, 0);
            acc[q][1][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], fim[d - p], acc[q][1][nt], 0, 0, 0);
          }
      }
    };
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < NT; b++) acc[q][a][b] = f4v{0.f, 0.f, 0.f, 0.f};
    if constexpr (KSH == 1) {
      // both halves into the tile, then the two products
#pragma unroll
      for (int q = 0; q < 2; q++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++)
#pragma unroll
          for (int v = 0; v < 4; v++) Xw[rl * XS + q * KP + nt * 16 + 4 * kg + v] = sum_or_diff(q, nt, v);
      wave_sync();
      bf8v fre[3], fim[3];
      tile_split(0, fre, fim);
      forward_products(0, 0, fre, fim);
      tile_split(KP, fre, fim);
      wave_sync();  // reads done before the next tile's writes
      forward_products(1, 0, fre, fim);
    } else {
      // one k-step's 32 columns at a time: n tiles 2 s2 and 2 s2 + 1
#pragma unroll
      for (int q = 0; q < 2; q++)
#pragma unroll
        for (int s2 = 0; s2 < KSH; s2++) {
#pragma unroll
          for (int t = 0; t < 2; t++)
#pragma unroll
            for (int v = 0; v < 4; v++) Xw[rl * XS + t * 16 + 4 * kg + v] = sum_or_diff(q, 2 * s2 + t, v);
          wave_sync();
          bf8v fre[3], fim[3];
          tile_split(0, fre, fim);
          wave_sync();  // reads done before the next round's writes
          forward_products(q, s2, fre, fim);
        }
    }
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int nt = 0; nt < NT; nt++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int cq = nt * 16 + 4 * kg + v, c = 2 * cq + q;
          const bool ok = rin && cq < (q ? CO : CE);
          buf_st_f2(make_float2(acc[q][0][nt][v], acc[q][1][nt][v]), rL, ok ? (unsigned)(row + R * c) * 8u : TM_OOB);
        }
    tile = ntile;
    bi = nbi;
    rL = rLn;
  }
}
template <int KSH, int NTH>
constexpr size_t theta_fused_par_lds() {
  return (size_t)4 * 3 * (2 * KSH * 16) * (KSH * 32 + 8) * 2 + (size_t)(NTH / 64) * 16 * ((KSH > 1 ? 32 : 64) + 1) * 8;
}

__device__ __forceinline__ float buf_ld_f(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ void buf_st_f(float x, __amdgpu_buffer_rsrc_t r, unsigned off) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x), r, (int)off, 0, 0);
}

// The fused theta step on the parity halves for half bases of 33..64 (two
// k-steps per half) with each 16-point tile shared by a pair of waves: wave
// 2 p + h computes component h (0: real, 1: imaginary) of every product of
// pair p's tiles.  The transforms are real, so the two components' products
// are independent; only the V phase couples them, and it is applied where the
// pair meets anyway: each wave writes its component of the DVR pair values
// (after the dump) into the pair's LDS tile, and both read complex values back
// in the layout of the forward B operand (DVR rows 8 kg + e of a k-step at
// point rl), one k-step (32 DVR rows, with their mirrors) per round.  Per wave
// this halves the accumulators and B operands of k_theta_fused_par<2, *>:
// eight waves (two per SIMD) fit in 256 registers, where one wave per tile
// needs about 400 and leaves one wave per SIMD.  Loads and stores move one
// float component (4 bytes at an 8-byte stride; the pair's two halves meet
// in the caches).  The pair meets twice per round, so both its waves run the
// same number of (tile, block) iterations (every wave of the group does);
// tiles past the item's end are masked (loads read zeros, stores are dropped).  Formulas, middle node,
// dump, flux rows and both propagator orders as in k_theta_fused_par.
// PSYNC: the two waves of a pair meet through a pair of LDS counters instead
// of work-group barriers, so the pairs of a work group need not move in step:
// each wave publishes its count of meetings after its LDS writes (release)
// and waits until its partner's count has caught up (acquire).  Both waves
// of a pair make the same meetings, and the wait gives up after about 2^22
// polls (tenths of a second; a meeting takes microseconds), so it ends even
// if that did not hold.
template <int NTH, bool PSYNC>
__global__ __launch_bounds__(NTH) void k_theta_fused_pair(ThetaFusedArgs P) {
  extern __shared__ __align__(16) unsigned char tm_lds[];
  constexpr int KSH = 2, KP = 64, NT = 4, NP = 64, LD = KP + 8, XS = 2 * 32 + 1, NPR = NTH / 128;
  __bf16 *M = reinterpret_cast<__bf16 *>(tm_lds);  // [4][3][NP][LD]
  // wave-uniform values in scalar registers (buffer descriptors built from
  // them must be, or every access becomes a loop over lanes)
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), pr = wave >> 1,
            h = wave & 1;
  // the pair's tile: [16 points][XS]; columns 0..31 DVR rows 32 s2 + j, 32..63 their mirrors
  float2 *Xp = reinterpret_cast<float2 *>(M + 4 * 3 * NP * LD) + pr * 16 * XS;
  float *Xf = reinterpret_cast<float *>(Xp);
  // the pairs' meeting counters, after the tiles: [pair][wave of the pair]
  int *meet = reinterpret_cast<int *>(reinterpret_cast<float2 *>(M + 4 * 3 * NP * LD) + NPR * 16 * XS);
  int nmeet = 0;
  auto pair_meet = [&]() {
    if constexpr (PSYNC) {
      nmeet++;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __hip_atomic_store(&meet[2 * pr + h], nmeet, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      for (int poll = 0; poll < (1 << 22); poll++) {
        if (__hip_atomic_load(&meet[2 * pr + (h ^ 1)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= nmeet) break;
        __builtin_amdgcn_s_sleep(1);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
      __syncthreads();
    }
  };
  const int4 item = P.items[blockIdx.x];
  const int gb0 = item.x, gnb = item.y, t0 = item.z, t1 = item.w;
  const BlockDesc bd = P.blocks[gb0];  // the group shares mat, nd, nf and the V columns
  const int R = P.R, nd = bd.nd, nf = bd.nf;
  const int NH = (nd + 1) >> 1, CE = (nf + 1) >> 1, CO = nf >> 1;
  auto rL_of = [&](int b) {
    const int f0 = __builtin_amdgcn_readfirstlane(P.blocks[gb0 + b].fcol0);
    return buf_rsrc(P.L + (size_t)f0 * R, (unsigned)nf * R * 8u);
  };
  auto rP_of = [&](int b) {
    const int d0 = __builtin_amdgcn_readfirstlane(P.blocks[gb0 + b].dcol0);
    return buf_rsrc(P.psi + (size_t)d0 * R, (unsigned)nd * R * 8u);
  };
  const __amdgpu_buffer_rsrc_t rV = buf_rsrc(P.expV + P.dcol_vidx[bd.dcol0], (unsigned)nd * R * 8u);
  const int rl = lane & 15, kg = lane >> 4;
  const unsigned hoff = 4u * (unsigned)h;
  float ra[2][8];  // [parity][e]: component h of FBR column 2 (32 s + 8 kg + e) + parity
  auto fetch = [&](const __amdgpu_buffer_rsrc_t &rd, int row, bool ok, int s, int Rx) {
    const int k0 = s * 32 + kg * 8;
    const unsigned base = ok ? (unsigned)row * 8u + hoff : TM_OOB;
    asm volatile("" : "+s"(Rx));
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int c = 2 * (k0 + e) + q;
        ra[q][e] = buf_ld_f(rd, c < nf ? base + (unsigned)(Rx * c) * 8u : TM_OOB);
      }
  };
  auto split3 = [](const float *x, bf8v *out) {
    float r[8];
#pragma unroll
    for (int e = 0; e < 8; e++) r[e] = x[e];
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const __bf16 b = f2bf(r[e]);
        out[q][e] = b;
        r[e] -= (float)b;
      }
  };
  // (tile, block) iterations: the same count for every wave of the group
  const int nit = (t1 - t0 + NPR - 1) / NPR * gnb;
  int tile = t0 + pr, bi = 0;
  __amdgpu_buffer_rsrc_t rL = rL_of(0);
  // the first tile's data are on their way while the matrices are copied
  fetch(rL, tile * 16 + rl, tile < t1 && tile * 16 + rl < R, 0, R);
  {
    const __bf16 *G = reinterpret_cast<const __bf16 *>(P.Bb) + (size_t)bd.mat * P.mat_stride;
    constexpr int QV = KP / 8;
    for (int u = threadIdx.x; u < 4 * 3 * NP * QV; u += blockDim.x) {
      const int pn = u / QV, kq = u % QV;
      *reinterpret_cast<bf8v *>(&M[pn * LD + kq * 8]) = *reinterpret_cast<const bf8v *>(&G[(size_t)pn * KP + kq * 8]);
    }
  }
  if (threadIdx.x < 2 * NPR) meet[threadIdx.x] = 0;
  __syncthreads();
  for (int it = 0; it < nit; it++) {
    // R as an opaque per-iteration value: the offsets R c 8 are formed where
    // they are used instead of being hoisted out of the loop (about a hundred
    // of them would not fit in scalar registers)
    int Ri = R;
    asm volatile("" : "+s"(Ri));
    const int row = tile * 16 + rl;
    const bool rin = tile < t1 && row < R;
    const bool lastb = bi + 1 == gnb;
    const int ntile = lastb ? tile + NPR : tile, nbi = lastb ? 0 : bi + 1;
    const int nrow = ntile * 16 + rl;
    const bool nin = ntile < t1 && nrow < R;
    const __amdgpu_buffer_rsrc_t rLn = rL_of(nbi);
    f4v acc[2][NT];  // [parity][n tile]: component h of E and O
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int b = 0; b < NT; b++) acc[q][b] = f4v{0.f, 0.f, 0.f, 0.f};
    // backward: E and O over the even and odd FBR columns
#pragma unroll 1
    for (int s = 0; s < KSH; s++) {
      bf8v bx[2][3];
#pragma unroll
      for (int q = 0; q < 2; q++) split3(ra[q], bx[q]);
      // the next k-step, or the next (tile, block)'s first: the block is
      // selected as an index, so that the descriptor stays scalar
      const bool more = s + 1 < KSH;
      fetch(rL_of(more ? bi : nbi), more ? row : nrow, more ? rin : nin, more ? s + 1 : 0, Ri);
#pragma unroll
      for (int q = 0; q < 2; q++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
          bf8v a[3];
#pragma unroll
          for (int p = 0; p < 3; p++)
            a[p] = *reinterpret_cast<const bf8v *>(&M[((q * 3 + p) * NP + nt * 16 + rl) * LD + s * 32 + kg * 8]);
#pragma unroll
          for (int d = 2; d >= 0; d--)
#pragma unroll
            for (int p = 0; p <= d; p++)
              acc[q][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], bx[q][d - p], acc[q][nt], 0, 0, 0);
        }
    }
    float dfac = 1.f;
    if (P.d1f && rin) dfac = P.d1f[row % P.n1] * P.d2f[row / P.n1];
    const int rr = row / P.n1;
    const bool frow = rin && rr >= P.r_lo && rr <= P.r_hi;
    const __amdgpu_buffer_rsrc_t rP = rP_of(bi);
    f4v accf[2][NT];  // [parity][n tile]: component h of the forward products
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int b = 0; b < NT; b++) accf[q][b] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s2 = 0; s2 < KSH; s2++) {
      // V phase at the forward B-operand rows i = 32 s2 + 8 kg + e and their
      // mirrors, read for each block and round (registers are short here)
      float2 wv[8][2];
      int Rv = Ri;
      asm volatile("" : "+s"(Rv));
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int i = 32 * s2 + 8 * kg + e;
        const bool ok = rin && i < NH;
        wv[e][0] = buf_ld_f2(rV, ok ? (unsigned)(row + Rv * i) * 8u : TM_OOB);
        wv[e][1] = buf_ld_f2(rV, ok ? (unsigned)(row + Rv * (nd - 1 - i)) * 8u : TM_OOB);
      }
      // DVR rows i = 32 s2 .. 32 s2 + 31 (n tiles 2 s2, 2 s2 + 1): the pair
      // (i, i') after the dump, psi rows, then component h into the tile
      int Rf = Ri;
      asm volatile("" : "+s"(Rf));
#pragma unroll
      for (int t = 0; t < 2; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int nt = 2 * s2 + t, j = t * 16 + 4 * kg + v, i = 32 * s2 + j, im = nd - 1 - i;
          const float e1 = acc[0][nt][v] + acc[1][nt][v], e2 = acc[0][nt][v] - acc[1][nt][v];
          const float c1 = e1 * dfac, c2 = e2 * dfac;
          if (frow) {
            buf_st_f(P.pre_dump ? e1 : c1, rP, i < NH ? (unsigned)(row + Rf * i) * 8u + hoff : TM_OOB);
            buf_st_f(P.pre_dump ? e2 : c2, rP, i < nd / 2 ? (unsigned)(row + Rf * im) * 8u + hoff : TM_OOB);
          }
          Xf[(rl * XS + j) * 2 + h] = c1;
          Xf[(rl * XS + 32 + j) * 2 + h] = c2;
        }
      pair_meet();
      // V phase at the B-operand positions, component h; sums and differences
      float sm[8], df[8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float2 x1 = Xp[rl * XS + 8 * kg + e], x2 = Xp[rl * XS + 32 + 8 * kg + e];
        const float2 w1 = wv[e][0], w2 = wv[e][1];
        const float p1 = h ? x1.x * w1.y + x1.y * w1.x : x1.x * w1.x - x1.y * w1.y;
        const float p2 = h ? x2.x * w2.y + x2.y * w2.x : x2.x * w2.x - x2.y * w2.y;
        sm[e] = p1 + p2;
        df[e] = p1 - p2;
      }
      pair_meet();  // reads done before the next round's writes
      // forward k-step s2: even FBR columns from the sums, odd from the differences
#pragma unroll
      for (int q = 0; q < 2; q++) {
        bf8v fb[3];
        split3(q ? df : sm, fb);
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
          bf8v a[3];
#pragma unroll
          for (int p = 0; p < 3; p++)
            a[p] = *reinterpret_cast<const bf8v *>(&M[(((2 + q) * 3 + p) * NP + nt * 16 + rl) * LD + s2 * 32 + kg * 8]);
#pragma unroll
          for (int d = 2; d >= 0; d--)
#pragma unroll
            for (int p = 0; p <= d; p++)
              accf[q][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], fb[d - p], accf[q][nt], 0, 0, 0);
        }
      }
    }
    int Ro = Ri;
    asm volatile("" : "+s"(Ro));
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int nt = 0; nt < NT; nt++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int cq = nt * 16 + 4 * kg + v, c = 2 * cq + q;
          const bool ok = rin && cq < (q ? CO : CE);
          buf_st_f(accf[q][nt][v], rL, ok ? (unsigned)(row + Ro * c) * 8u + hoff : TM_OOB);
        }
    tile = ntile;
    bi = nbi;
    rL = rLn;
  }
}
template <int NTH>
constexpr size_t theta_fused_pair_lds() {
  return (size_t)4 * 3 * 64 * 72 * 2 + (size_t)(NTH / 128) * 16 * 65 * 8 + (size_t)(NTH / 128) * 2 * sizeof(int);
}

template <int KS, int NT>
constexpr size_t theta_mx_lds() {
  return (size_t)3 * NT * 16 * (KS * 32 + 8) * 2;
}

// (The flux-surface extraction, k_flux_extract, is defined with the r pass's
// arguments above: the register-FFT r pass can run it in its first work groups.)

// ---------------------------------------------------------------------------
// Time -> energy accumulation over one time block (CRPAux.F:4-31 summed over
// the block):  fai[x, e] += sum_t tau[t, e] psi_s[t, x],  same for d fai,
// tau[t, e] = dt exp(i E_e (n0 + t + 1) dt)  (the reference's recursion
// timeEvol.C:799-817 evaluated directly, as evolutionCUDA2.cu:143-145 does).
// k_flux_tau writes the block's tau table once; k_flux_accumulate stages it
// and the block's surface values through LDS.  Work group: 128 x-values x 16
// energies; each thread holds 2 x-values (xl, xl + 64) x 4 energies, so every
// tau value read from LDS feeds two complex products per array; each complex
// multiply-add is four FMAs.  Per (x, e) the sum runs over t in increasing
// order.  The work groups of one x range (its energy blocks) get linear ids
// 8 apart, so they run on one XCD and read the surface values from its L2.
// (8 energies per thread: 194 registers, two waves per SIMD, and a 20-step
// fold of the batched workload took 12.6 ms; 4 per thread: 9.6 ms.)
// ---------------------------------------------------------------------------
constexpr int FX_X = 128;
constexpr int FX_E = 16;
constexpr int FX_EPT = 4;
constexpr int FX_TC = 8;

// counter += add (one thread): advances device-side step counts between
// kernels, so a captured graph can be replayed without new arguments
static __global__ void k_counter_add(long long *counter, int add) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter += add;
}

// tau[t * ne + e] for the block's nt steps, after n0 steps accumulated before it
static __global__ __launch_bounds__(256) void k_flux_tau(double2 *__restrict__ tau,
                                                  const double *__restrict__ energies, int ne, int nt,
                                                  const long long *__restrict__ n0p, double dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt * ne) return;
  const int t = i / ne, e = i % ne;
  const double time = (double)(*n0p + t + 1) * dt;
  double s, c;
  sincos(energies[e] * time, &s, &c);
  tau[i] = make_double2(c * dt, s * dt);
}

// With part != nullptr it also leaves, for its 128 x-values and each of its
// energies, the partial sum  sum_x w_x Im(conj(fai) dfai)  of the updated
// accumulators in part[xb][e] (w_x = dcol_w[x / n1]), so that P(E) after the
// fold needs these partials only (k_crp_parts) instead of another pass over
// fai and dfai.
static __global__ __launch_bounds__(256) void k_flux_accumulate(const double2 *__restrict__ bpsi,
                                                         const double2 *__restrict__ bdpsi,
                                                         double2 *__restrict__ fai, double2 *__restrict__ dfai,
                                                         const double2 *__restrict__ tau, int nx, int ne, int nt,
                                                         double *__restrict__ part, const double *__restrict__ dcol_w,
                                                         int n1) {
  __shared__ double2 sb[FX_TC][FX_X];
  __shared__ double2 sdb[FX_TC][FX_X];
  __shared__ double2 stau[FX_TC][FX_E];
  // linear id = (xb / 8) * 8 * ney + eb * 8 + xb % 8
  const int ney = (ne + FX_E - 1) / FX_E, lid = blockIdx.x;
  const int xb = (lid / (8 * ney)) * 8 + lid % 8, eb = (lid / 8) % ney;
  if (xb * FX_X >= nx) return;  // padding of the last group of eight x ranges (whole work group)
  const int t = threadIdx.x;
  const int xl = t & 63;
  const int eg = t >> 6;
  const int e0 = eb * FX_E;
  // the accumulators start from the stored fai, dfai: their loads are issued
  // first and arrive while the time block is staged and multiplied
  double2 af[2][FX_EPT], adf[2][FX_EPT];
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int q = 0; q < FX_EPT; q++) {
      const int x = xb * FX_X + xl + 64 * h, e = e0 + eg * FX_EPT + q;
      const bool ok = x < nx && e < ne;
      const size_t o = (size_t)e * nx + x;
      af[h][q] = ok ? fai[o] : make_double2(0.0, 0.0);
      adf[h][q] = ok ? dfai[o] : make_double2(0.0, 0.0);
    }
  for (int tc = 0; tc < nt; tc += FX_TC) {
    for (int u = t; u < FX_TC * FX_X; u += 256) {
      const int tt = u / FX_X, xx = u % FX_X;
      const int gx = xb * FX_X + xx;
      double2 b = make_double2(0.0, 0.0), db = b;
      if (tc + tt < nt && gx < nx) {
        b = bpsi[(size_t)(tc + tt) * nx + gx];
        db = bdpsi[(size_t)(tc + tt) * nx + gx];
      }
      sb[tt][xx] = b;
      sdb[tt][xx] = db;
    }
    for (int u = t; u < FX_TC * FX_E; u += 256) {
      const int tt = u / FX_E, ee = u % FX_E;
      stau[tt][ee] = (tc + tt < nt && e0 + ee < ne) ? tau[(size_t)(tc + tt) * ne + e0 + ee] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int tmax = min(FX_TC, nt - tc);
    for (int tt = 0; tt < tmax; tt++) {
      const double2 b0 = sb[tt][xl], db0 = sdb[tt][xl], b1 = sb[tt][xl + 64], db1 = sdb[tt][xl + 64];
#pragma unroll
      for (int q = 0; q < FX_EPT; q++) {
        const double2 tw = stau[tt][eg * FX_EPT + q];
        auto mac = [&](double2 &acc, const double2 v) {  // acc += tw v
          acc.x = fma(tw.x, v.x, acc.x);
          acc.x = fma(-tw.y, v.y, acc.x);
          acc.y = fma(tw.x, v.y, acc.y);
          acc.y = fma(tw.y, v.x, acc.y);
        };
        mac(af[0][q], b0);
        mac(adf[0][q], db0);
        mac(af[1][q], b1);
        mac(adf[1][q], db1);
      }
    }
    __syncthreads();
  }
  double pim[FX_EPT];
#pragma unroll
  for (int q = 0; q < FX_EPT; q++) pim[q] = 0.0;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int x = xb * FX_X + xl + 64 * h;
    if (x < nx) {
      const double w = part ? dcol_w[x / n1] : 0.0;
#pragma unroll
      for (int q = 0; q < FX_EPT; q++) {
        const int e = e0 + eg * FX_EPT + q;
        if (e < ne) {
          const size_t o = (size_t)e * nx + x;
          const double2 a = af[h][q], b = adf[h][q];
          fai[o] = a;
          dfai[o] = b;
          pim[q] = fma(w, a.x * b.y - a.y * b.x, pim[q]);  // w Im(conj(a) b)
        }
      }
    }
  }
  if (part) {
    // one wave per energy group: sum its 64 lanes (fixed order: a tree over lanes)
#pragma unroll
    for (int q = 0; q < FX_EPT; q++) {
      double v = pim[q];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      const int e = e0 + eg * FX_EPT + q;
      if (xl == 0 && e < ne) part[(size_t)xb * ne + e] = v;
    }
  }
}

// P(E) of block b from the partials of k_flux_accumulate: the sum over the
// block's 128-value x ranges (blocks start on such a range), times the same
// factors as k_crp.  Grid (ne, nblocks); the sum runs in a fixed order.
static __global__ __launch_bounds__(64) void k_crp_parts(const double *__restrict__ part,
                                                  const double *__restrict__ eta_sq,
                                                  const BlockDesc *__restrict__ blocks, double *__restrict__ crp,
                                                  int n1, int ne, double dr1_mu2, double sign) {
  const int e = blockIdx.x;
  const BlockDesc bd = blocks[blockIdx.y];
  const int t0 = n1 * bd.dcol0 / FX_X, t1 = (n1 * (bd.dcol0 + bd.nd) + FX_X - 1) / FX_X;
  double s = 0.0;
  for (int t = t0 + (int)threadIdx.x; t < t1; t += 64) s += part[(size_t)t * ne + e];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (threadIdx.x == 0) crp[(size_t)blockIdx.y * ne + e] = sign * s / eta_sq[(size_t)blockIdx.y * ne + e] * dr1_mu2;
}

// P(E) of block b: Im(sum_k w_k sum_i conj(fai) dfai) / eta^2 * dr1/mu2
// (CRPAux.F:34-66); sign = -1 reproduces the CUDA propagator's
// Zdotc(dfai, fai) (evolutionCUDA2.cu:972-977).  Grid (ne, nblocks).
static __global__ __launch_bounds__(256) void k_crp(const double2 *__restrict__ fai, const double2 *__restrict__ dfai,
                                             const double *__restrict__ dcol_w, const double *__restrict__ eta_sq,
                                             const BlockDesc *__restrict__ blocks, double *__restrict__ crp,
                                             int n1, int nx, int ne, double dr1_mu2, double sign) {
  __shared__ double red[256];
  const int e = blockIdx.x;
  const BlockDesc bd = blocks[blockIdx.y];
  const int x0 = n1 * bd.dcol0, x1 = n1 * (bd.dcol0 + bd.nd);
  double s = 0.0;
  for (int x = x0 + threadIdx.x; x < x1; x += 256) {
    const double2 a = fai[(size_t)e * nx + x], b = dfai[(size_t)e * nx + x];
    s += dcol_w[x / n1] * (a.x * b.y - a.y * b.x);  // Im(conj(a) b)
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) crp[(size_t)blockIdx.y * ne + e] = sign * red[0] / eta_sq[(size_t)blockIdx.y * ne + e] * dr1_mu2;
}

// per-group partial sums of w_k |psi|^2 (module_for_psi, timeEvol.C:220-243)
template <typename T>
__global__ __launch_bounds__(256) void k_norm_partial(const typename cx_of<T>::type *__restrict__ psi,
                                                      const double *__restrict__ dcol_w, size_t n12, size_t n,
                                                      double *__restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (size_t x = (size_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (size_t)gridDim.x * 256) {
    const typename cx_of<T>::type v = psi[x];
    s += dcol_w[x / n12] * ((double)v.x * v.x + (double)v.y * v.y);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// One partial sum per work group of  wc[c] * f * |a[x]|^2  over the n points of
// a complex array with columns of n12 = n1 * n2 points (c = x / n12, i = x mod
// n1, j = (x / n1) mod n2), where f = fi[i] + fj[j] when fi is given, times
// V[dcol_vidx[c] + x mod n12] when V is given.  This is the common form of the
// reference's per-step diagnostics: module_for_psi (timeEvol.C:220-243, wc = w_k),
// potential_energy (:256-283, wc = w_k, V), kinetic_energy_for_legendre_psi
// (:390-436, wc = 1/(l+1/2), f = p1^2/2mu1 + p2^2/2mu2) and rotational_energy
// (:323-360, wc = l(l+1)/(l+1/2), f = 1/2mu1R^2 + 1/2mu2r^2).
template <typename T>
__global__ __launch_bounds__(256) void k_diag_sum(const typename cx_of<T>::type *__restrict__ a, size_t n,
                                                  size_t n12, int n1, const double *__restrict__ wc,
                                                  const double *__restrict__ fi, const double *__restrict__ fj,
                                                  const double *__restrict__ V,
                                                  const long long *__restrict__ dcol_vidx,
                                                  double *__restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (size_t x = (size_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (size_t)gridDim.x * 256) {
    const size_t c = x / n12, r = x - c * n12;
    double f = wc[c];
    if (fi) f *= fi[r % n1] + fj[r / n1];
    if (V) f *= V[dcol_vidx[c] + (long long)r];
    const typename cx_of<T>::type v = a[x];
    s += f * ((double)v.x * v.x + (double)v.y * v.y);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

}  // namespace qd
