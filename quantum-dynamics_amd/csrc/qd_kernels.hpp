// qd_kernels.hpp -- CDNA4 (gfx950) kernels of the split-operator step.
//
// One step of the reference (common/timeEvol.C:619-647, evolution_dt) is
//   V phase -> theta forward (DVR->FBR) -> rot/2 -> 2-D FFT -> kinetic ->
//   2-D IFFT -> rot/2 -> theta backward (FBR->DVR) -> dump -> flux surface.
// Here it runs as four HBM sweeps:
//   theta_gemm<fwd>   V phase fused into the operand load (prologue)
//   fft_lines<R>      rot/2, FFT along R, exp(-i T1 dt)/n1, IFFT along R
//   fft_lines<r>      FFT along r, exp(-i T2 dt)/n2, IFFT along r, rot/2
//   theta_gemm<bwd>   dump fused into the store (epilogue)
// plus a small flux-surface extraction kernel and, once per time block, the
// time->energy accumulation.  The kinetic propagator factorises,
// exp(-i(T1+T2)dt) = exp(-i T1 dt) exp(-i T2 dt) (timeEvol.C:514), so the 2-D
// FFT splits into independent 1-D passes that each fit in LDS.
//
// Layout (column-major, R fastest, common/mat.h:18-41): psi[i + n1 (j + n2 k)],
// complex interleaved.  The theta transforms treat psi as a real matrix of
// 2 n1 n2 rows and ntheta columns (W and P are real, LegTransform.F:6-59).
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
// i * s * a for s = +-1
template <typename C, typename T> __device__ __forceinline__ C cmul_i(C a, T s) {
  C r;
  r.x = -a.y * s;
  r.y = a.x * s;
  return r;
}

// ---------------------------------------------------------------------------
// Elementwise kernels (first-step phase, plain-pipeline operators, conversion)
// ---------------------------------------------------------------------------

// psi /= sqrt_ref(exp(-i V dt)), the first-step half phase of the CPU
// propagator (timeEvol.C:519-538).  sqrt_ref is the reference's pow(c, 0.5)
// with arg() = asin(Im/|c|) plus quadrant fixes (complex.h:60-68,148-159); it
// equals exp(+i V dt/2) only while |V dt| < pi, which is reproduced here.
__global__ void k_prephase_cpu(double2 *psi, const double *V, double dt, size_t n) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  double s, c;
  sincos(-dt * V[x], &s, &c);
  // arg of (c, s), |c| = sqrt(c^2 + s^2)
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
  double2 p = psi[x];
  p.x /= q2;
  p.y /= q2;
  psi[x] = cmul(p, make_double2(q.x, -q.y));
}

// psi *= exp(+i V dt/2): the CUDA propagator's first-step phase
// (evolutionCUDA2.cu:507, evolution_with_potential(-dt/2)).
__global__ void k_prephase_cuda(double2 *psi, const double *V, double dt, size_t n) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  double s, c;
  sincos(0.5 * dt * V[x], &s, &c);
  psi[x] = cmul(psi[x], make_double2(c, s));
}

// psi *= exp(-i V dt) with V in FP64 (evolution_with_potential_dt, timeEvol.C:540-559)
template <typename T>
__global__ void k_vphase(typename cx_of<T>::type *psi, const double *V, const typename cx_of<T>::type *expV,
                         double dt, size_t n) {
  typedef typename cx_of<T>::type C;
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  C e;
  if (V) {
    double s, c;
    sincos(-dt * V[x], &s, &c);
    e.x = (T)c;
    e.y = (T)s;
  } else {
    e = expV[x];
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

// lpsi *= rotR[l][i] rotr[l][j]  (evolution_with_rotational_dt_2, timeEvol.C:561-582)
template <typename T>
__global__ void k_rot(typename cx_of<T>::type *lpsi, const typename cx_of<T>::type *rotR,
                      const typename cx_of<T>::type *rotr, int n1, int n2, size_t n) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const int i = (int)(x % n1);
  const size_t jl = x / n1;
  const int j = (int)(jl % n2);
  const int l = (int)(jl / n2);
  lpsi[x] = cmul(lpsi[x], cmul(rotR[(size_t)l * n1 + i], rotr[(size_t)l * n2 + j]));
}

// lpsi *= kinR[i] kinr[j]  (evolution_with_kinetic_dt, timeEvol.C:584-606; the
// tables carry 1/n1 and 1/n2)
template <typename T>
__global__ void k_kin2d(typename cx_of<T>::type *lpsi, const typename cx_of<T>::type *kinR,
                        const typename cx_of<T>::type *kinr, int n1, int n2, size_t n) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const int i = (int)(x % n1);
  const int j = (int)((x / n1) % n2);
  lpsi[x] = cmul(lpsi[x], cmul(kinR[i], kinr[j]));
}

__global__ void k_f64_to_f32(const double2 *a, float2 *b, size_t n) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) b[x] = make_float2((float)a[x].x, (float)a[x].y);
}
__global__ void k_f32_to_f64(const float2 *a, double2 *b, size_t n) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) b[x] = make_double2((double)a[x].x, (double)a[x].y);
}

// ---------------------------------------------------------------------------
// LDS-staged Stockham FFT along one axis with fused phases
// ---------------------------------------------------------------------------
// A block transforms LPB lines of length n.  Lines along R are contiguous
// (line id j + n2 l, element stride 1); lines along r are strided by n1
// (line id i + n1 l, element stride n1) and a block takes LPB consecutive i so
// every global access is coalesced across the block.  Passes follow the
// autosort Stockham scheme: pass with radix r and span Ns reads
// x[b + q n/r], twiddles by w^(q (b mod Ns) n/(Ns r)), and writes
// y[(b/Ns) Ns r + (b mod Ns) + q Ns].  The first pass reads global memory,
// the last pass writes it; the passes in between go through LDS.

enum {
  FFT_PRE_ROT = 1,   // multiply by rotR[l][i] rotr[l][j] before the forward FFT
  FFT_FWD = 2,
  FFT_KIN = 4,       // multiply the spectrum by kin[p]
  FFT_INV = 8,
  FFT_POST_ROT = 16  // multiply by the rot tables after the inverse FFT
};

constexpr int FFT_MAX_PASS = 12;
constexpr int FFT_BPT = 4;  // butterflies per thread per pass (max)

struct FftPlanArg {
  int n;
  int npass;
  int radix[FFT_MAX_PASS];
};

template <typename T>
struct FftArgs {
  typename cx_of<T>::type *data;
  const typename cx_of<T>::type *tw;    // tw[e] = exp(-2 pi i e / n)
  const typename cx_of<T>::type *kin;   // [n] spectrum multiplier or null
  const typename cx_of<T>::type *rotR;  // [m][n1]
  const typename cx_of<T>::type *rotr;  // [m][n2]
  int n1, n2, m;
  int lpb;      // lines per block
  int nlines;   // total lines
  int flags;
  FftPlanArg plan;
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
    const T s3 = sign * (T)0.86602540378443864676372317075294;
    const C t = cadd(v[1], v[2]);
    const C d = csub(v[1], v[2]);
    const C a0 = v[0];
    v[0] = cadd(a0, t);
    const C mm = csub(a0, cscale(t, (T)0.5));
    const C jd = cmul_i(d, s3);
    v[1] = cadd(mm, jd);
    v[2] = csub(mm, jd);
  } else if (r == 5) {
    const T c1 = (T)0.30901699437494742410229341718282, c2 = (T)-0.80901699437494742410229341718282;
    const T s1 = (T)0.95105651629515357211643933337938, s2 = (T)0.58778525229247312916870595463907;
    const C a0 = v[0];
    const C b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
    const C d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
    v[0] = cadd(a0, cadd(b1, b2));
    const C m1 = cadd(a0, cadd(cscale(b1, c1), cscale(b2, c2)));
    const C m2 = cadd(a0, cadd(cscale(b1, c2), cscale(b2, c1)));
    const C n1 = cmul_i(cadd(cscale(d1, s1), cscale(d2, s2)), sign);
    const C n2 = cmul_i(csub(cscale(d1, s2), cscale(d2, s1)), sign);
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
  }
}

template <typename T, int AXIS>
__global__ __launch_bounds__(256) void k_fft_lines(FftArgs<T> A) {
  typedef typename cx_of<T>::type C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *lds = reinterpret_cast<C *>(smem_raw);
  const int n = A.plan.n;
  const int lpb = A.lpb;
  const int tpl = 256 / lpb;
  const int t = threadIdx.x;
  const int line_local = AXIS == 0 ? t / tpl : t % lpb;
  const int bt = AXIS == 0 ? t % tpl : t / lpb;
  const int line = blockIdx.x * lpb + line_local;
  const bool active = line < A.nlines;
  const size_t n12 = (size_t)A.n1 * A.n2;
  // line coordinates
  int jl_fixed;  // R axis: j of the line; r axis: i of the line
  int l;
  size_t base;
  size_t stride;
  if (AXIS == 0) {
    const int lc = active ? line : 0;
    jl_fixed = lc % A.n2;
    l = lc / A.n2;
    base = (size_t)lc * A.n1;
    stride = 1;
  } else {
    const int lc = active ? line : 0;
    jl_fixed = lc % A.n1;
    l = lc / A.n1;
    base = (size_t)l * n12 + jl_fixed;
    stride = A.n1;
  }
  auto lidx = [&](int p) { return AXIS == 0 ? line_local * n + p : p * lpb + line_local; };

  const bool fwd = A.flags & FFT_FWD, inv = A.flags & FFT_INV;
  const int npass = A.plan.npass;
  const int ptot = (fwd ? npass : 0) + (inv ? npass : 0);

  C v[FFT_BPT][5];
  for (int pass = 0; pass < ptot; pass++) {
    const bool is_fwd_pass = fwd && pass < npass;
    const int pi = is_fwd_pass ? pass : (fwd ? pass - npass : pass);
    const int r = A.plan.radix[pi];
    int Ns = 1;
    for (int q = 0; q < pi; q++) Ns *= A.plan.radix[q];
    const T sign = is_fwd_pass ? (T)-1 : (T)1;
    const int nb = n / r;
    const int stride_q = nb;
    // ---- read ----
#pragma unroll
    for (int bb = 0; bb < FFT_BPT; bb++) {
      const int b = bt + bb * tpl;
      if (b < nb && active) {
#pragma unroll
        for (int q = 0; q < 5; q++) {
          if (q >= r) break;
          const int p = b + q * stride_q;
          C x;
          if (pass == 0) {
            x = A.data[base + (size_t)p * stride];
            if ((A.flags & FFT_PRE_ROT) && is_fwd_pass) {
              const int i = AXIS == 0 ? p : jl_fixed;
              const int j = AXIS == 0 ? jl_fixed : p;
              x = cmul(x, cmul(A.rotR[(size_t)l * A.n1 + i], A.rotr[(size_t)l * A.n2 + j]));
            }
          } else {
            x = lds[lidx(p)];
          }
          v[bb][q] = x;
        }
      }
    }
    if (pass > 0) __syncthreads();
    // ---- compute + write ----
    const bool last = pass == ptot - 1;
    const bool last_fwd = is_fwd_pass && pass == npass - 1;
#pragma unroll
    for (int bb = 0; bb < FFT_BPT; bb++) {
      const int b = bt + bb * tpl;
      if (b < nb && active) {
        const int k = b % Ns;
        if (Ns > 1) {
          const int estep = k * (n / (Ns * r));
#pragma unroll
          for (int q = 1; q < 5; q++) {
            if (q >= r) break;
            C w = A.tw[q * estep];
            if (!is_fwd_pass) w.y = -w.y;
            v[bb][q] = cmul(v[bb][q], w);
          }
        }
        dft_small<C, T>(v[bb], r, sign);
        const int od = (b / Ns) * Ns * r + k;
#pragma unroll
        for (int q = 0; q < 5; q++) {
          if (q >= r) break;
          const int p = od + q * Ns;
          C x = v[bb][q];
          if (last_fwd && (A.flags & FFT_KIN)) x = cmul(x, A.kin[p]);
          if (last) {
            if ((A.flags & FFT_POST_ROT) && !is_fwd_pass) {
              const int i = AXIS == 0 ? p : jl_fixed;
              const int j = AXIS == 0 ? jl_fixed : p;
              x = cmul(x, cmul(A.rotR[(size_t)l * A.n1 + i], A.rotr[(size_t)l * A.n2 + j]));
            }
            A.data[base + (size_t)p * stride] = x;
          } else {
            lds[lidx(p)] = x;
          }
        }
      }
    }
    if (!last) __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// theta transform, FP64 on MFMA (v_mfma_f64_16x16x4f64)
// ---------------------------------------------------------------------------
// C[2R x N] = A[2R x K] * B[K x N] on psi viewed as a real matrix.
// Block: 64 complex rows (128 real rows) x 128 columns, 4 waves of 32 real
// rows each, K in chunks of 16.  B is pre-packed row-major [Kpad][Npad].
// Prologue (PRO): 1 = multiply A by exp(-i V dt); 2 = by dump * exp(-i V dt).
// Epilogue (EPI): 1 = multiply the result by the dump d1[i] d2[j].
constexpr int TG_ROWS = 64;    // complex rows per block
constexpr int TG_COLS = 128;   // output columns per block
constexpr int TG_KC64 = 16;    // K chunk for FP64
constexpr int TG_LDS64 = 144;  // padded row length (doubles) of the LDS tiles

template <int PRO, int EPI>
__global__ __launch_bounds__(256) void k_theta_f64(const double2 *__restrict__ A, double2 *__restrict__ Cm,
                                                   const double *__restrict__ Bp, int R, int K, int N,
                                                   int Npad, const double *__restrict__ V, double dt,
                                                   const double *__restrict__ d1, const double *__restrict__ d2,
                                                   int n1) {
  __shared__ __align__(16) double As[TG_KC64][TG_LDS64];
  __shared__ __align__(16) double Bs[TG_KC64][TG_LDS64];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int row0 = blockIdx.x * TG_ROWS;
  const int col0 = blockIdx.y * TG_COLS;
  d4v acc[2][8];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 8; b++) acc[a][b] = d4v{0.0, 0.0, 0.0, 0.0};

  for (int k0 = 0; k0 < K; k0 += TG_KC64) {
    // ---- A chunk: 64 complex rows x 16 k, phase prologue ----
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int cr = t & 63;
      const int kk = (t >> 6) + 4 * s;
      const int r = row0 + cr;
      const int k = k0 + kk;
      double2 v = make_double2(0.0, 0.0);
      if (r < R && k < K) {
        const size_t x = (size_t)r + (size_t)R * k;
        v = A[x];
        if (PRO) {
          double sn, cs;
          sincos(-dt * V[x], &sn, &cs);
          v = cmul(v, make_double2(cs, sn));
          if (PRO == 2) {
            const double d = d1[r % n1] * d2[r / n1];
            v.x *= d;
            v.y *= d;
          }
        }
      }
      *reinterpret_cast<double2 *>(&As[kk][2 * cr]) = v;
    }
    // ---- B chunk: 16 k x 128 columns ----
#pragma unroll
    for (int s = 0; s < 8; s++) {
      const int col = t & 127;
      const int kk = (t >> 7) + 2 * s;
      Bs[kk][col] = Bp[(size_t)(k0 + kk) * Npad + col0 + col];
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < TG_KC64 / 4; ks++) {
      const int kr = ks * 4 + (lane >> 4);
      double a[2];
#pragma unroll
      for (int rb = 0; rb < 2; rb++) a[rb] = As[kr][wave * 32 + rb * 16 + (lane & 15)];
#pragma unroll
      for (int cb = 0; cb < 8; cb++) {
        const double b = Bs[kr][cb * 16 + (lane & 15)];
#pragma unroll
        for (int rb = 0; rb < 2; rb++) acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb], b, acc[rb][cb], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // ---- epilogue: lane holds rows 4(lane/16)+0..3 of column lane%16 ----
#pragma unroll
  for (int rb = 0; rb < 2; rb++) {
    const int rr = wave * 32 + rb * 16 + 4 * (lane >> 4);  // real row, multiple of 4
    const int r = row0 + rr / 2;                           // first of two complex rows
#pragma unroll
    for (int cb = 0; cb < 8; cb++) {
      const int col = col0 + cb * 16 + (lane & 15);
      if (col >= N) continue;
      double2 c0 = make_double2(acc[rb][cb][0], acc[rb][cb][1]);
      double2 c1 = make_double2(acc[rb][cb][2], acc[rb][cb][3]);
      if (EPI) {
        if (r < R) {
          const double d = d1[r % n1] * d2[r / n1];
          c0.x *= d;
          c0.y *= d;
        }
        if (r + 1 < R) {
          const double d = d1[(r + 1) % n1] * d2[(r + 1) / n1];
          c1.x *= d;
          c1.y *= d;
        }
      }
      const size_t x = (size_t)r + (size_t)R * col;
      if (r + 1 < R) {
        Cm[x] = c0;
        Cm[x + 1] = c1;
      } else if (r < R) {
        Cm[x] = c0;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// theta transform, mixed precision on MFMA (v_mfma_f32_16x16x32_bf16)
// ---------------------------------------------------------------------------
// Each FP32 operand is split into NS bf16 pieces a = a0 + a1 (+ a2); the
// products a_p b_q with p + q <= NS - 1 are accumulated in FP32 (NS = 2: the
// hi*hi + hi*lo + lo*hi scheme; NS = 3: six products, FP32-level accuracy).
// B is pre-split on the host as Bs[piece][Npad][Kpad] (k contiguous).
// Block: 64 complex rows x 128 columns; K in chunks of 32.
constexpr int TG_KC16 = 32;
constexpr int TG_LDS16 = 40;  // padded k length (bf16) of an LDS row

__device__ __forceinline__ __bf16 f2bf(float x) { return (__bf16)x; }

template <int NS, int PRO, int EPI>
__global__ __launch_bounds__(256) void k_theta_bf16(const float2 *__restrict__ A, float2 *__restrict__ Cm,
                                                    const __bf16 *__restrict__ Bsplit, int R, int K, int N,
                                                    int Npad, int Kpad, const float2 *__restrict__ expV,
                                                    const float *__restrict__ d1, const float *__restrict__ d2,
                                                    int n1) {
  __shared__ __align__(16) __bf16 As[NS][128][TG_LDS16];
  __shared__ __align__(16) __bf16 Bs[NS][TG_COLS][TG_LDS16];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int row0 = blockIdx.x * TG_ROWS;
  const int col0 = blockIdx.y * TG_COLS;
  f4v acc[2][8];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 8; b++) acc[a][b] = f4v{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < K; k0 += TG_KC16) {
    // ---- A chunk: thread owns complex row cr and 8 consecutive k ----
    {
      const int cr = t & 63;
      const int kg = t >> 6;  // 0..3
      const int r = row0 + cr;
      float re[8], im[8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int k = k0 + kg * 8 + e;
        float2 v = make_float2(0.f, 0.f);
        if (r < R && k < K) {
          const size_t x = (size_t)r + (size_t)R * k;
          v = A[x];
          if (PRO) {
            v = cmul(v, expV[x]);
            if (PRO == 2) {
              const float d = d1[r % n1] * d2[r / n1];
              v.x *= d;
              v.y *= d;
            }
          }
        }
        re[e] = v.x;
        im[e] = v.y;
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
    }
    // ---- B chunk: 128 columns x 32 k per piece ----
    {
      const int col = t >> 1;
      const int half = t & 1;
#pragma unroll
      for (int p = 0; p < NS; p++) {
        const __bf16 *src = Bsplit + ((size_t)p * Npad + col0 + col) * Kpad + k0 + half * 16;
        const bf8v b0 = *reinterpret_cast<const bf8v *>(src);
        const bf8v b1 = *reinterpret_cast<const bf8v *>(src + 8);
        *reinterpret_cast<bf8v *>(&Bs[p][col][half * 16]) = b0;
        *reinterpret_cast<bf8v *>(&Bs[p][col][half * 16 + 8]) = b1;
      }
    }
    __syncthreads();
    bf8v a[NS][2];
#pragma unroll
    for (int p = 0; p < NS; p++)
#pragma unroll
      for (int rb = 0; rb < 2; rb++)
        a[p][rb] = *reinterpret_cast<const bf8v *>(&As[p][wave * 32 + rb * 16 + (lane & 15)][(lane >> 4) * 8]);
#pragma unroll
    for (int cb = 0; cb < 8; cb++) {
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
    __syncthreads();
  }
#pragma unroll
  for (int rb = 0; rb < 2; rb++) {
    const int rr = wave * 32 + rb * 16 + 4 * (lane >> 4);
    const int r = row0 + rr / 2;
#pragma unroll
    for (int cb = 0; cb < 8; cb++) {
      const int col = col0 + cb * 16 + (lane & 15);
      if (col >= N) continue;
      float2 c0 = make_float2(acc[rb][cb][0], acc[rb][cb][1]);
      float2 c1 = make_float2(acc[rb][cb][2], acc[rb][cb][3]);
      if (EPI) {
        if (r < R) {
          const float d = d1[r % n1] * d2[r / n1];
          c0.x *= d;
          c0.y *= d;
        }
        if (r + 1 < R) {
          const float d = d1[(r + 1) % n1] * d2[(r + 1) / n1];
          c1.x *= d;
          c1.y *= d;
        }
      }
      const size_t x = (size_t)r + (size_t)R * col;
      if (r + 1 < R) {
        Cm[x] = c0;
        Cm[x + 1] = c1;
      } else if (r < R) {
        Cm[x] = c0;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Flux surface: psi_s and d psi_s / dr on r = r_s (Gradient3.F:9-167,
// gradients2.cu:583-610), written into slot `slot` of the time block.
// ---------------------------------------------------------------------------
struct StencilArg {
  int h;
  double num[31];  // numerators c_j (integer form for 3..11 points) ...
  double den;      // ... over den: d psi = sum_j num_j (f(s+j) - f(s-j)) / (den dr)
};

template <typename T>
__global__ void k_flux_extract(const typename cx_of<T>::type *__restrict__ psi, double2 *__restrict__ bpsi,
                               double2 *__restrict__ bdpsi, int n1, int n2, int ntheta, int s, double dr,
                               StencilArg st, int slot) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = n1 * ntheta;
  if (x >= nx) return;
  const int i = x % n1, k = x / n1;
  const size_t base = (size_t)i + (size_t)n1 * n2 * k;
  auto F = [&](int j) {
    const typename cx_of<T>::type v = psi[base + (size_t)n1 * j];
    return make_double2((double)v.x, (double)v.y);
  };
  double2 acc = make_double2(0.0, 0.0);
  for (int j = st.h; j >= 1; j--) {
    const double2 a = F(s + j), b = F(s - j);
    acc.x += st.num[j - 1] * (a.x - b.x);
    acc.y += st.num[j - 1] * (a.y - b.y);
  }
  const double den = st.den * dr;
  bpsi[(size_t)slot * nx + x] = F(s);
  bdpsi[(size_t)slot * nx + x] = make_double2(acc.x / den, acc.y / den);
}

// ---------------------------------------------------------------------------
// Time -> energy accumulation over one time block (CRPAux.F:4-31 summed over
// the block):  fai[x, e] += sum_t tau[t, e] psi_s[t, x],  same for d fai,
// tau[t, e] = dt exp(i E_e (n0 + t + 1) dt)  (the reference's recursion
// timeEvol.C:799-817 evaluated directly, as evolutionCUDA2.cu:143-145 does).
// Block: 64 x-values x 32 energies; each thread holds 8 energies.
// ---------------------------------------------------------------------------
constexpr int FX_X = 64;
constexpr int FX_E = 32;
constexpr int FX_EPT = 8;
constexpr int FX_TC = 16;

__global__ __launch_bounds__(256) void k_flux_accumulate(const double2 *__restrict__ bpsi,
                                                         const double2 *__restrict__ bdpsi,
                                                         double2 *__restrict__ fai, double2 *__restrict__ dfai,
                                                         const double *__restrict__ energies, int nx, int ne,
                                                         int nt, long long n0, double dt) {
  __shared__ double2 sb[FX_TC][FX_X];
  __shared__ double2 sdb[FX_TC][FX_X];
  __shared__ double2 stau[FX_TC][FX_E];
  const int t = threadIdx.x;
  const int xl = t & 63;
  const int eg = t >> 6;  // 0..3 -> energies eg*8 .. eg*8+7 of the tile
  const int x = blockIdx.x * FX_X + xl;
  const int e0 = blockIdx.y * FX_E;
  double2 af[FX_EPT], adf[FX_EPT];
#pragma unroll
  for (int q = 0; q < FX_EPT; q++) {
    af[q] = make_double2(0.0, 0.0);
    adf[q] = make_double2(0.0, 0.0);
  }
  for (int tc = 0; tc < nt; tc += FX_TC) {
    for (int u = t; u < FX_TC * FX_X; u += 256) {
      const int tt = u / FX_X, xx = u % FX_X;
      const int gx = blockIdx.x * FX_X + xx;
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
      double2 tau = make_double2(0.0, 0.0);
      if (tc + tt < nt && e0 + ee < ne) {
        const double time = (double)(n0 + tc + tt + 1) * dt;
        double s, c;
        sincos(energies[e0 + ee] * time, &s, &c);
        tau = make_double2(c * dt, s * dt);
      }
      stau[tt][ee] = tau;
    }
    __syncthreads();
    const int tmax = min(FX_TC, nt - tc);
    for (int tt = 0; tt < tmax; tt++) {
      const double2 b = sb[tt][xl], db = sdb[tt][xl];
#pragma unroll
      for (int q = 0; q < FX_EPT; q++) {
        const double2 tau = stau[tt][eg * FX_EPT + q];
        af[q] = cadd(af[q], cmul(tau, b));
        adf[q] = cadd(adf[q], cmul(tau, db));
      }
    }
    __syncthreads();
  }
  if (x < nx) {
#pragma unroll
    for (int q = 0; q < FX_EPT; q++) {
      const int e = e0 + eg * FX_EPT + q;
      if (e < ne) {
        const size_t o = (size_t)e * nx + x;
        fai[o] = cadd(fai[o], af[q]);
        dfai[o] = cadd(dfai[o], adf[q]);
      }
    }
  }
}

// P(E) = Im(sum_k w_k sum_i conj(fai) dfai) / eta^2 * dr1/mu2  (CRPAux.F:34-66);
// sign = -1 reproduces the CUDA propagator's Zdotc(dfai, fai) (evolutionCUDA2.cu:972-977).
__global__ __launch_bounds__(256) void k_crp(const double2 *__restrict__ fai, const double2 *__restrict__ dfai,
                                             const double *__restrict__ w, const double *__restrict__ eta_sq,
                                             double *__restrict__ crp, int n1, int ntheta, double dr1_mu2,
                                             double sign) {
  __shared__ double red[256];
  const int e = blockIdx.x;
  const int nx = n1 * ntheta;
  double s = 0.0;
  for (int x = threadIdx.x; x < nx; x += 256) {
    const double2 a = fai[(size_t)e * nx + x], b = dfai[(size_t)e * nx + x];
    s += w[x / n1] * (a.x * b.y - a.y * b.x);  // Im(conj(a) b)
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) crp[e] = sign * red[0] / eta_sq[e] * dr1_mu2;
}

// per-block partial sums of w_k |psi|^2 (module_for_psi, timeEvol.C:220-243)
template <typename T>
__global__ __launch_bounds__(256) void k_norm_partial(const typename cx_of<T>::type *__restrict__ psi,
                                                      const double *__restrict__ w, size_t n12, size_t n,
                                                      double *__restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (size_t x = (size_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (size_t)gridDim.x * 256) {
    const typename cx_of<T>::type v = psi[x];
    s += w[x / n12] * ((double)v.x * v.x + (double)v.y * v.y);
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
