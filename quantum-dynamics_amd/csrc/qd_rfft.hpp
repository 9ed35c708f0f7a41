// qd_rfft.hpp -- register-resident FFT passes (forward, kinetic, inverse).
//
// A line of N = N1 * N2 points is split x[a + N1 m], a < N1, m < N2:
//   stage A   thread a:  length-N2 DFT over m -> k2, times W_N^(a k2)
//   exchange  (LDS)
//   stage B   thread b:  length-N1 DFT over a -> k1, giving X[b + N2 k1];
//             times the kinetic factor; inverse length-N1 DFT -> c;
//             times W_N^(-b c)
//   exchange  (LDS)
//   stage B'  thread a:  inverse length-N2 DFT -> y[a + N1 c1]
// so the forward transform, the kinetic phase and the inverse transform of a
// line cost two LDS exchanges, and every thread ends with the elements it
// loaded (the same global addresses, which keeps the store coalesced).
// The small DFTs are unrolled at compile time with exact constants.  In FP32
// every twiddle and phase factor is a rotation applied as three shears, whose
// determinant is 1 whatever the rounding of its constants: a single-float
// complex factor would bias |W| and, applied coherently every step, the norm,
// and a hi + lo pair avoids that at 8 instead of 3 FMAs.  The radix-3 and -5
// constants, which are not rotations, stay hi + lo.
#pragma once

#include <type_traits>
#include <utility>

#include "qd_kernels.hpp"
#include "qd_records.hpp"

namespace qd {

// ---- compile-time loop and trigonometry ------------------------------------
template <int... Is, typename F>
__device__ __forceinline__ void sfor_impl(std::integer_sequence<int, Is...>, F &&f) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void sfor(F &&f) {
  sfor_impl(std::make_integer_sequence<int, N>{}, f);
}

constexpr double rf_pi = 3.14159265358979323846264338327950288;
constexpr double rf_sin_small(double x) {  // |x| <= pi/4, Taylor series
  double t = x, s = x;
  for (int n = 1; n < 14; n++) {
    t *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    s += t;
  }
  return s;
}
constexpr double rf_cos_small(double x) {
  double t = 1.0, s = 1.0;
  for (int n = 1; n < 14; n++) {
    t *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
    s += t;
  }
  return s;
}
// cos or sin of 2 pi e / S; the octant of the angle is found on the integers
// so the series only sees arguments in [0, pi/4].
constexpr double rf_trig(int e, int S, bool want_sin) {
  e %= S;
  if (e < 0) e += S;
  const long long f = 8LL * e;        // angle = (pi/4) f / S
  const int q = (int)(f / S);         // octant 0..7
  const long long rn = f - (long long)q * S;
  double c = 0, s = 0;
  int Q = 0;
  if (q % 2 == 0) {  // angle = (q/2)(pi/2) + r
    const double r = (rf_pi / 4) * (double)rn / (double)S;
    c = rf_cos_small(r);
    s = rf_sin_small(r);
    Q = q / 2;
  } else {           // angle = ((q+1)/2)(pi/2) - r'
    const double r = (rf_pi / 4) * (double)(S - rn) / (double)S;
    c = rf_cos_small(r);
    s = -rf_sin_small(r);
    Q = (q + 1) / 2;
  }
  double cc = c, ss = s;
  switch (Q & 3) {
    case 1: cc = -s; ss = c; break;
    case 2: cc = -c; ss = -s; break;
    case 3: cc = s; ss = -c; break;
    default: break;
  }
  return want_sin ? ss : cc;
}

// ---- rotations as three shears (FP32) ---------------------------------------
// x1 = x - t y, y' = y + s x1, x' = x1 - t y' rotates x + iy by theta for
// t = tan(theta/2), s = sin theta.  Its matrix has determinant 1 for any FP32
// t and s, so the rounding of the constants moves the angle but never scales
// the norm, and (-t, -s) undoes (t, s) exactly: three FMAs against the eight
// of a hi + lo complex product.  The constants are picked like
// qdt::lift_consts on the host: among the floats within 8 ulp of tan(theta/2),
// each with the three floats nearest (1 - cos theta) / t, the pair whose
// product is nearest 1 - cos theta (the composite's trace fixes its angle).
struct rf_lift_c {
  float t, s;
};
constexpr float rf_fstep(float f, int k) {  // f > 0 moved by k float spacings
  return __builtin_bit_cast(float, (unsigned)((int)__builtin_bit_cast(unsigned, f) + k));
}
constexpr double rf_fabs(double x) { return x < 0 ? -x : x; }
constexpr rf_lift_c rf_lift(double theta) {  // |theta| <= pi/2
  if (theta == 0) return {0.f, 0.f};
  const double th = theta < 0 ? -theta : theta;
  const double sh = rf_sin_small(th / 2), ch = rf_cos_small(th / 2);
  const double a = 2 * sh * sh;
  const float t0 = (float)(sh / ch);
  float tb = t0, sb = (float)(2 * sh * ch);
  double best = 1e300;
  for (int k = -8; k <= 8; k++) {
    const float t = rf_fstep(t0, k);
    if (!(t > 0)) continue;
    const float s0 = (float)(a / t);
    for (int q = -1; q <= 1; q++) {
      const float s = rf_fstep(s0, q);
      const double e = rf_fabs((double)t * (double)s - a);
      if (e < best) {
        best = e;
        tb = t;
        sb = s;
      }
    }
  }
  return theta < 0 ? rf_lift_c{-tb, -sb} : rf_lift_c{tb, sb};
}

__device__ __forceinline__ float2 lrot(float2 v, float t, float s) {
  const float x1 = fmaf(-t, v.y, v.x);
  const float y1 = fmaf(s, x1, v.y);
  float2 r;
  r.x = fmaf(-t, y1, x1);
  r.y = y1;
  return r;
}

// Table factors of the register passes (rf_tab in qd_kernels.hpp): twiddle
// (inv: its conjugate), D/2 phase, kinetic factor of a length-N line.
__device__ __forceinline__ double2 rf_tw(double2 v, double2 w, bool inv) { return tmul(v, w, inv); }
__device__ __forceinline__ float2 rf_tw(float2 v, float2 w, bool inv) {
  return inv ? lrot(v, -w.x, -w.y) : lrot(v, w.x, w.y);
}
__device__ __forceinline__ double2 rf_ph(double2 v, double2 w) { return tmul(v, w, false); }
// two rotations by theta/2 (t, s, 2t): |theta| up to pi with |t| <= 1
__device__ __forceinline__ float2 rf_ph(float2 v, float4 w) {
  const float x1 = fmaf(-w.x, v.y, v.x);
  const float y1 = fmaf(w.y, x1, v.y);
  const float x2 = fmaf(-w.z, y1, x1);
  const float y2 = fmaf(w.y, x2, y1);
  float2 r;
  r.x = fmaf(-w.x, y2, x2);
  r.y = y2;
  return r;
}
template <int N>
__device__ __forceinline__ double2 rf_kin(double2 v, double2 w) {
  return tmul(v, w, false);
}
template <int N>
__device__ __forceinline__ float2 rf_kin(float2 v, float4 w) {
  float2 u;
  if constexpr ((N & (N - 1)) == 0) {  // c = +-1/N, exact
    u.x = v.x * w.z;
    u.y = v.y * w.z;
  } else {                             // c as hi + lo
    u.x = fmaf(v.x, w.z, v.x * w.w);
    u.y = fmaf(v.y, w.z, v.y * w.w);
  }
  return lrot(u, w.x, w.y);
}

// x * W_S^E with W = exp(-2 pi i / S) (FWD) or its conjugate.  FP32: the
// nearest quarter turn exactly, the rest (|angle| <= pi/4) as three shears.
template <typename T, int S, int E, bool FWD>
__device__ __forceinline__ typename cx_of<T>::type rf_twc(typename cx_of<T>::type x) {
  typedef typename cx_of<T>::type C;
  constexpr int e = ((E % S) + S) % S;
  if constexpr (e == 0) {
    return x;
  } else if constexpr (2 * e == S) {
    C r;
    r.x = -x.x;
    r.y = -x.y;
    return r;
  } else if constexpr (4 * e == S || 4 * e == 3 * S) {
    // exp(-i pi/2) = -i for e = S/4 forward
    constexpr bool minus_i = (4 * e == S) == FWD;
    C r;
    if (minus_i) {
      r.x = x.y;
      r.y = -x.x;
    } else {
      r.x = -x.y;
      r.y = x.x;
    }
    return r;
  } else {
    constexpr double wr = rf_trig(e, S, false);
    constexpr double wi = FWD ? -rf_trig(e, S, true) : rf_trig(e, S, true);
    C r;
    if constexpr (sizeof(T) == 8) {
      r.x = x.x * wr - x.y * wi;
      r.y = x.x * wi + x.y * wr;
    } else {
      constexpr int q = (8 * e + S) / (2 * S);  // nearest quarter turn
      constexpr int rn = 4 * e - q * S;         // the rest, in units of 2 pi / (4 S)
      constexpr rf_lift_c L = rf_lift((FWD ? -2.0 : 2.0) * rf_pi * rn / (4.0 * S));
      const C y = rn == 0 ? x : lrot(x, L.t, L.s);
      constexpr int qi = FWD ? (4 - q % 4) % 4 : q % 4;  // (-i)^q forward, i^q inverse, as a power of i
      if constexpr (qi == 0) {
        r = y;
      } else if constexpr (qi == 1) {
        r.x = -y.y;
        r.y = y.x;
      } else if constexpr (qi == 2) {
        r.x = -y.x;
        r.y = -y.y;
      } else {
        r.x = y.y;
        r.y = -y.x;
      }
    }
    return r;
  }
}

// In-place DFT of S points held in registers (natural order in and out),
// decimation in time with the radix preference 4, 2, 3, 5.
template <typename T, int S, bool FWD>
__device__ __forceinline__ void rdft(typename cx_of<T>::type *x) {
  typedef typename cx_of<T>::type C;
  if constexpr (S == 2 || S == 3 || S == 4 || S == 5) {
    dft_small<C, T>(x, S, FWD ? (T)-1 : (T)1);
  } else if constexpr (S > 5) {
    constexpr int r = S % 4 == 0 ? 4 : S % 2 == 0 ? 2 : S % 3 == 0 ? 3 : 5;
    constexpr int m = S / r;
    static_assert(m * r == S, "length must factor into 2, 3, 5");
    C sub[r][m];
    sfor<r>([&](auto J) {
      constexpr int j = decltype(J)::value;
      sfor<m>([&](auto I) {
        constexpr int i = decltype(I)::value;
        sub[j][i] = x[j + r * i];
      });
      rdft<T, m, FWD>(sub[j]);
    });
    sfor<m>([&](auto Kc) {
      constexpr int k = decltype(Kc)::value;
      C y[r];
      sfor<r>([&](auto J) {
        constexpr int j = decltype(J)::value;
        y[j] = rf_twc<T, S, j * k, FWD>(sub[j][k]);
      });
      dft_small<C, T>(y, r, FWD ? (T)-1 : (T)1);
      sfor<r>([&](auto Q) {
        constexpr int q = decltype(Q)::value;
        x[k + m * q] = y[q];
      });
    });
  }
}

// Factorisation used for a line length: N1 threads per line (a divisor of 256),
// N2 <= N1 points per thread.
constexpr int rf_n1(int n) {
  return n == 512 ? 32 : n == 384 ? 32 : n == 256 ? 16 : n == 192 ? 16 : n == 128 ? 16 : n == 96 ? 16
       : n == 64 ? 8 : n == 48 ? 8 : n == 32 ? 8 : n == 24 ? 8 : 0;
}
constexpr bool rf_length(int n) { return rf_n1(n) != 0; }

// Stages A .. B' of one line (see the file comment).  v: the thread's N2
// points on entry and exit; xidx(n, k): LDS slot of exchange element (n, k)
// of this line; twa[a * (N2 + 1) + k] = W_N^(a' k'); kin[b + N2 k]: kinetic
// factor of the frequency stage B's thread b holds in u[k].
//
// FP32 runs the decomposition on centred indices, a' = a - N1 [a >= N1/2],
// k' = k - N2 [k >= (N2 + 1)/2]: with n = a' + N1 m and k = k' + N2 k1 every
// point and frequency is still met once (mod N), the length-N1 and length-N2
// DFTs are unchanged (their exponents only matter mod N1, N2), and the table
// twiddles W_N^(a' k') stay within a quarter turn, which the three-shear
// rotation needs.  So thread a holds point a' + N1 m in v[m] (rf_pos) and the
// kinetic table is stored in the threads' order.  FP64 keeps a' = a, k' = k.
// WAVE: the N1 threads of every line lie in one wave (consecutive threads,
// N1 dividing 64), so the two exchanges through the line's LDS rows need only
// the wave's own order, not a work-group barrier.
template <typename T, int N1, int N2, bool WAVE = false, typename XI>
__device__ __forceinline__ void rf_line(typename cx_of<T>::type *v, int a, bool act, typename cx_of<T>::type *xb,
                                        XI xidx, const typename rf_tab<T>::tw *twa,
                                        const typename rf_tab<T>::kin *kin) {
  typedef typename cx_of<T>::type C;
  constexpr int XS = N2 + 1;
  static_assert(!WAVE || 64 % N1 == 0, "a line's threads must stay within one wave");
  auto exchange_sync = [] {
    if constexpr (WAVE) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
      __syncthreads();
    }
  };
  if (act) {
    rdft<T, N2, true>(v);
#pragma unroll
    for (int k = 1; k < N2; k++) v[k] = rf_tw(v[k], twa[a * XS + k], false);
#pragma unroll
    for (int k = 0; k < N2; k++) xb[xidx(a, k)] = v[k];
  }
  exchange_sync();
  if (act && a < N2) {
    C u[N1];
#pragma unroll
    for (int n = 0; n < N1; n++) u[n] = xb[xidx(n, a)];
    rdft<T, N1, true>(u);
#pragma unroll
    for (int k = 0; k < N1; k++) u[k] = rf_kin<N1 * N2>(u[k], kin[a + N2 * k]);
    rdft<T, N1, false>(u);
#pragma unroll
    for (int c = 1; c < N1; c++) u[c] = rf_tw(u[c], twa[c * XS + a], true);
#pragma unroll
    for (int c = 0; c < N1; c++) xb[xidx(c, a)] = u[c];
  }
  exchange_sync();
  if (act) {
#pragma unroll
    for (int k = 0; k < N2; k++) v[k] = xb[xidx(a, k)];
    rdft<T, N2, false>(v);
  }
}

// Point of the line held in element m of thread a (see rf_line).
template <int N1, int N2, bool CEN>
struct rf_pos {
  int pb, p0;
  __device__ __forceinline__ explicit rf_pos(int a) {
    const bool up = CEN && a >= N1 / 2;
    pb = up ? a - N1 : a;
    p0 = up ? a - N1 + N1 * N2 : a;
  }
  __device__ __forceinline__ int operator()(int m) const { return m == 0 ? p0 : pb + N1 * m; }
};

template <typename T, int N1, int N2>
constexpr size_t rf_table_bytes() {
  return (size_t)N1 * (N2 + 1) * sizeof(typename rf_tab<T>::tw) + (size_t)N1 * N2 * sizeof(typename rf_tab<T>::kin);
}

// Twiddle and kinetic tables into LDS.  tw: FP64 W_N^e for e < N; FP32 the
// shear constants of W_N^e for -N/4 <= e <= N/4 at tw[e + N/4].
template <typename T, int N1, int N2>
__device__ __forceinline__ void rf_tables(typename rf_tab<T>::tw *twa, typename rf_tab<T>::kin *kin,
                                          const typename rf_tab<T>::tw *tw, const typename rf_tab<T>::kin *kg) {
  constexpr int N = N1 * N2, XS = N2 + 1;
  for (int e = threadIdx.x; e < N; e += blockDim.x) {
    const int aa = e / N2, kk = e % N2;
    if constexpr (sizeof(T) == 4) {
      const int ac = aa < N1 / 2 ? aa : aa - N1, kc = kk < (N2 + 1) / 2 ? kk : kk - N2;
      twa[aa * XS + kk] = tw[ac * kc + N / 4];
    } else {
      twa[aa * XS + kk] = tw[aa * kk];
    }
    kin[e] = kg[e];
  }
}

// out[m] = sum_K M[K][m] in[K] over the g (<= G) lines of a mixing unit at
// one point p, in place in the staged lines.  M is G x G row-major in LDS,
// zero-padded beyond g; rows are read as broadcast 4-wide vectors.
template <typename T, int G, int SP>
__device__ __forceinline__ void rf_mix_point(typename cx_of<T>::type *stg, int L0, int g, int p, const T *M) {
  typedef typename cx_of<T>::type C;
  typedef typename std::conditional<sizeof(T) == 8, double4, float4>::type V4;
  C acc[G];
#pragma unroll
  for (int m = 0; m < G; m++) {
    acc[m].x = 0;
    acc[m].y = 0;
  }
#pragma unroll
  for (int K = 0; K < G; K++) {
    if (K < g) {
      const C x = stg[(L0 + K) * SP + p];
#pragma unroll
      for (int m4 = 0; m4 < G / 4; m4++) {
        const V4 u = *reinterpret_cast<const V4 *>(M + K * G + 4 * m4);
        acc[4 * m4 + 0].x += x.x * u.x;
        acc[4 * m4 + 0].y += x.y * u.x;
        acc[4 * m4 + 1].x += x.x * u.y;
        acc[4 * m4 + 1].y += x.y * u.y;
        acc[4 * m4 + 2].x += x.x * u.z;
        acc[4 * m4 + 2].y += x.y * u.z;
        acc[4 * m4 + 3].x += x.x * u.w;
        acc[4 * m4 + 3].y += x.y * u.w;
      }
    }
  }
#pragma unroll
  for (int m = 0; m < G; m++)
    if (m < g) stg[(L0 + m) * SP + p] = acc[m];
}

// Mixing of every mixed unit of a bundle, one (unit, point) per thread.
template <typename T, int N, int SP, int NTH>
__device__ __forceinline__ void rf_mix_bundle(typename cx_of<T>::type *stg, int nu, const int *ug, const int *line0,
                                              const int *uoff, const T *mats) {
  for (int e = threadIdx.x; e < nu * N; e += NTH) {
    const int u = e / N, p = e % N;
    const int g = ug[u];
    if (g == 0) continue;
    const int L0 = line0[u];
    const T *M = mats + uoff[u];
    switch ((g + 3) >> 2) {
      case 1: rf_mix_point<T, 4, SP>(stg, L0, g, p, M); break;
      case 2: rf_mix_point<T, 8, SP>(stg, L0, g, p, M); break;
      case 3: rf_mix_point<T, 12, SP>(stg, L0, g, p, M); break;
      case 4: rf_mix_point<T, 16, SP>(stg, L0, g, p, M); break;
      case 5: rf_mix_point<T, 20, SP>(stg, L0, g, p, M); break;
      case 6: rf_mix_point<T, 24, SP>(stg, L0, g, p, M); break;
      case 7: rf_mix_point<T, 28, SP>(stg, L0, g, p, M); break;
      default: rf_mix_point<T, 32, SP>(stg, L0, g, p, M); break;
    }
  }
}

// R pass: lines (column, r) of N = n1 contiguous points; LW lines per work
// group of LW * N1 threads (LW >= the largest mixing group).  Without K mixing each thread loads its points straight from
// HBM; with mixing the lines of a bundle are staged in LDS, mixed into the l
// basis, transformed, and mixed back.
template <typename T, int N1, int N2, bool MIXT, int LW>
__global__ __launch_bounds__(LW * N1) void k_rpass_rf(RPassArgs<T> A) {
  typedef typename cx_of<T>::type C;
  typedef typename rf_tab<T>::tw TW;
  typedef typename rf_tab<T>::kin TK;
  typedef typename rf_tab<T>::ph TB;
  constexpr int N = N1 * N2, NTH = LW * N1, XS = N2 + 1, LX = N1 * XS;
  constexpr int SP = N + N / 32 + 1;  // staged line stride
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TW *s_twa = reinterpret_cast<TW *>(smem_raw);
  TK *s_kin = reinterpret_cast<TK *>(s_twa + N1 * XS);
  C *xch = reinterpret_cast<C *>(s_kin + N);  // [LW][LX]
  C *stg = xch + LW * LX;                       // [LW][SP] (mixing) or the D_R rows
  const int t = threadIdx.x;
  const int a = t % N1, L = t / N1;
  const rf_pos<N1, N2, sizeof(T) == 4> P(a);
  rf_tables<T, N1, N2>(s_twa, s_kin, A.ltw, A.lkin);
  C v[N2];
  C *xb = xch + L * LX;
  auto xidx = [](int n, int k) { return n * XS + k; };
  if constexpr (!MIXT) {
    // Persistent over batches of LW consecutive lines of one column (the host
    // requires n2 % LW == 0), so the batch has one D_R row.  The next batch's
    // points and D_R row are loaded while this one is transformed; the row
    // goes through a double-buffered LDS copy.
    constexpr int PHE = (N + NTH - 1) / NTH;
    TB *s_ph = reinterpret_cast<TB *>(stg);  // [2][N]
    const long long nlines = (long long)A.nfcol * A.n2;
    const long long nbatch = nlines / LW;
    C raw[N2];
    TB phr[PHE];
    // The D_R row index of the batch to prefetch is read at the top of the
    // previous iteration, when no other load is outstanding.
    // Loads go through descriptors over the batch and its row; with `has`
    // false the descriptors are empty, the loads return zeros and move no
    // memory (the loop body stays branch-free).
    auto fetch = [&](long long b, int lam, bool has) {
      const __amdgpu_buffer_rsrc_t rd = buf_rsrc(A.data + b * LW * N, has ? (unsigned)(LW * N * sizeof(C)) : 0u);
      const __amdgpu_buffer_rsrc_t rp = buf_rsrc(A.lph + (size_t)lam * N, has ? (unsigned)(N * sizeof(TB)) : 0u);
#pragma unroll
      for (int m = 0; m < N2; m++) raw[m] = buf_ld<C>(rd, (unsigned)((L * N + P(m)) * sizeof(C)));
#pragma unroll
      for (int q = 0; q < PHE; q++)
        if (N % NTH == 0 || q * NTH + t < N) phr[q] = buf_ld<TB>(rp, (unsigned)((q * NTH + t) * sizeof(TB)));
    };
    // one column per batch: the index is uniform (readfirstlane keeps the
    // descriptors in scalar registers)
    auto lam_of = [&](long long b) { return __builtin_amdgcn_readfirstlane(A.fcol_lam[(b * LW) / A.n2]); };
    long long b = blockIdx.x;
    if (b < nbatch) fetch(b, lam_of(b), true);
    int buf = 0;
    for (; b < nbatch; b += gridDim.x) {
      const bool has_n = b + gridDim.x < nbatch;
      const long long bn = has_n ? b + gridDim.x : b;
      const int lam_n = lam_of(bn);
      TB *ph = s_ph + buf * N;
#pragma unroll
      for (int q = 0; q < PHE; q++)
        if (N % NTH == 0 || q * NTH + t < N) ph[q * NTH + t] = phr[q];
      __syncthreads();  // row (and, first time round, the tables)
#pragma unroll
      for (int m = 0; m < N2; m++) v[m] = rf_ph(raw[m], ph[P(m)]);
      // unconditional (empty descriptors past the last batch) keeps the loop
      // body straight-line, so the waits stay exact across iterations
      fetch(bn, lam_n, has_n);
      // line L's threads are t = N1 L .. N1 L + N1 - 1: one wave
      rf_line<T, N1, N2, true>(v, a, true, xb, xidx, s_twa, s_kin);
      C *dst = A.data + (b * LW + L) * N;
#pragma unroll
      for (int m = 0; m < N2; m++) dst[P(m)] = rf_ph(v[m], ph[P(m)]);
      buf ^= 1;
    }
  } else {
    // Persistent over bundles of whole mixing units (the K blocks of one
    // (J, j, r)), staged in LDS:  load -> K-to-l mixing -> D_R/2, FFT,
    // kinetic, inverse FFT, D_R/2 (register stages; the phase rows are read
    // with the points) -> l-to-K mixing -> store.  Line offsets and D_R rows
    // come from host-built tables; the unit matrices are copied to LDS, so
    // the mixing loops touch no global memory.
    __shared__ long long s_off[LW];
    __shared__ int s_lam[LW];
    __shared__ int s_unit_line0[LW + 1];
    __shared__ int s_ug[LW], s_mg[LW], s_usrc[LW], s_uoff[LW + 1];
    // U (row K, column m) and its transpose per mixed unit, padded to G4 = 4 ceil(g/4)
    __shared__ __align__(32) T s_um[LW * 40], s_umt[LW * 40];
    constexpr int PER = (LW * N + NTH - 1) / NTH;
    const bool mix = (A.flags & FFT_MIX) != 0;
    for (int bdl = blockIdx.x; bdl < A.nbundle; bdl += gridDim.x) {
      const int ub = A.bundle_u0[bdl], nu = A.bundle_u0[bdl + 1] - ub;
      const int bl0 = A.bundle_l0[bdl], nl = A.bundle_l0[bdl + 1] - bl0;
      __syncthreads();  // the previous bundle is done with LDS (first time: tables)
      if (t < nu) {
        const RUnit U = A.units[ub + t];
        s_ug[t] = U.g;
        s_mg[t] = (mix && U.u0 >= 0) ? U.g : 0;
        s_usrc[t] = U.u0;
      }
      if (t < nl) {
        s_off[t] = A.line_off[bl0 + t];
        s_lam[t] = A.line_lam[bl0 + t];
      }
      __syncthreads();
      // all of a thread's loads in flight together, while the unit layout is built
      C x[PER];
#pragma unroll
      for (int i = 0; i < PER; i++) {
        const int e = t + i * NTH;
        if (e < nl * N) x[i] = A.data[s_off[e / N] + e % N];
      }
      if (t == 0) {
        int l0 = 0, no = 0;
        for (int u = 0; u < nu; u++) {
          s_unit_line0[u] = l0;
          l0 += s_ug[u];
          s_uoff[u] = no;
          const int g4 = (s_mg[u] + 3) & ~3;
          no += g4 * g4;
        }
        s_unit_line0[nu] = l0;
        s_uoff[nu] = no;
      }
      __syncthreads();
      for (int i = t; i < s_uoff[nu]; i += NTH) {
        int u = 0;
        while (s_uoff[u + 1] <= i) u++;
        const int g = s_mg[u], g4 = (g + 3) & ~3, ii = i - s_uoff[u];
        const int r = ii / g4, c = ii % g4;
        const T *src = A.umat + s_usrc[u];
        const bool in = r < g && c < g;
        s_um[i] = in ? src[r * g + c] : (T)0;
        s_umt[i] = in ? src[c * g + r] : (T)0;
      }
#pragma unroll
      for (int i = 0; i < PER; i++) {
        const int e = t + i * NTH;
        if (e < nl * N) stg[(e / N) * SP + e % N] = x[i];
      }
      __syncthreads();
      // K -> l: c_m = sum_K U[K, m] c_K
      if (mix) {
        rf_mix_bundle<T, N, SP, NTH>(stg, nu, s_mg, s_unit_line0, s_uoff, s_um);
        __syncthreads();
      }
      const bool act = L < nl;
      const TB *ph = A.lph + (size_t)(act ? s_lam[L] : 0) * N;
      if (act) {
#pragma unroll
        for (int m = 0; m < N2; m++) v[m] = rf_ph(stg[L * SP + P(m)], ph[P(m)]);
      }
      rf_line<T, N1, N2>(v, a, act, xb, xidx, s_twa, s_kin);
      if (act) {
#pragma unroll
        for (int m = 0; m < N2; m++) stg[L * SP + P(m)] = rf_ph(v[m], ph[P(m)]);
      }
      __syncthreads();
      // l -> K: c_K = sum_m U[K, m] c_m
      if (mix) {
        rf_mix_bundle<T, N, SP, NTH>(stg, nu, s_mg, s_unit_line0, s_uoff, s_umt);
        __syncthreads();
      }
#pragma unroll
      for (int i = 0; i < PER; i++) {
        const int e = t + i * NTH;
        if (e < nl * N) A.data[s_off[e / N] + e % N] = stg[(e / N) * SP + e % N];
      }
    }
  }
}

template <typename T, int N1, int N2, bool MIXT, int LW>
constexpr size_t rpass_rf_lds() {
  constexpr int N = N1 * N2;
  return rf_table_bytes<T, N1, N2>() + (size_t)LW * N1 * (N2 + 1) * sizeof(typename cx_of<T>::type) +
         (MIXT ? (size_t)LW * (N + N / 32 + 1) * sizeof(typename cx_of<T>::type)
               : (size_t)2 * N * sizeof(typename rf_tab<T>::ph));
}

// ---- mixing R pass with prefetch (FP32) ------------------------------------
// Bundles of whole mixing units (the K lines of one (J, j, r)), at most NL
// lines and two distinct unit matrices each, with host-built fixed-stride
// records.  A work group of NL * N1 threads stages one bundle in LDS, one line
// per row; the row doubles as that line's FFT exchange buffer.  Per bundle:
//   mix K -> l      one (unit, point) per thread, the unit's g outputs
//                   accumulated in registers and written in place (only this
//                   thread touches this point of the unit's lines); the
//                   unit matrix comes from HBM/L2 through scalar loads;
//   FFT phase       line L: D_R/2, forward FFT, kinetic, inverse FFT, D_R/2;
//   mix l -> K      read the g values of the point, store the K lines
//                   straight to HBM.
// The next bundle's descriptors are loaded during the K -> l mixing.  With one
// work group per CU its lines are loaded during the FFT and the l -> K mixing,
// so the HBM traffic of one bundle overlaps the arithmetic of the previous
// one; with two work groups per CU (TWO) the other work group covers it.
// (MxHead, the bundle header, is in qd_records.hpp with the host code that
// builds and checks the records.)

// A work group of the mixing R pass takes `chunk` consecutive bundles at a
// time (round-robin over the grid by chunks): consecutive bundles mostly share
// their unit matrices, which then stay in the CU's scalar cache.
constexpr int MX_CHUNK = 8;  // default (on the J = 0..20 batch: 1, 2, 4, 8, 16 -> 4.14, 3.92, 3.78, 3.71, 3.77 ms)

struct RMixArgs {
  float2 *data;
  const float2 *tw;          // W_n1^e, |e| <= n1/4, as shear constants (rf_tab<float>)
  const float4 *kin;         // exp(-i T1 dt) / n1, in the threads' order (rf_tab<float>)
  const float4 *dR;          // D_R/2 rows [lam][n1] (rf_tab<float>)
  const MxHead *head;        // [nbundle]
  const long long *line_off;  // [nbundle][NL] element offset of each line
  const int *line_lam;       // [nbundle][NL] D_R row of each line
  const int *unit;           // [nbundle][NL] line0 | g << 8 | slot << 16 (slot 0: no mixing)
  const float *umat;         // padded U and U^T of every unit matrix
  int nbundle;
  int chunk;                 // consecutive bundles per turn of a work group (>= 1)
};

template <int N1, int N2, int NL>
constexpr int mx_row_stride() {
  constexpr int N = N1 * N2, LX = N1 * (N2 + 1), SPA = N + N / 32 + 1;
  return LX > SPA ? LX : SPA;
}
template <int N1, int N2, int NL>
constexpr size_t rpass_mx_lds() {
  return rf_table_bytes<float, N1, N2>() + (size_t)NL * mx_row_stride<N1, N2, NL>() * sizeof(float2);
}

// Unit matrices are read through the constant address space: their addresses
// are uniform over a wave, so they become scalar loads into SGPRs that feed
// the FMAs directly, and the LDS bandwidth stays with the line values.
typedef float qd_f4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) float kfloat;
typedef const __attribute__((address_space(4))) qd_f4v kfloat4;

// acc_n = sum_k M[k][n] x_k for the outputs n < G over the matrix M (row
// stride GS = 4 ceil(G / 4), zero padded beyond the unit's g lines) for the
// lines line0 .. line0 + g - 1 of the staging rows at point p; out(n, acc_n)
// for n < g.  The k loop runs over
// the g lines (g is uniform over the wave); its body has no branches, so the
// compiler issues the loads ahead of the FMAs.
template <int G, int SP, int NL, typename OUT>
__device__ __forceinline__ void mx_apply(const float2 *stg, int line0, int g, int p, kfloat *M, OUT out) {
  constexpr int GS = (G + 3) & ~3;  // row stride (padded order)
  float2 acc[G];
#pragma unroll
  for (int n = 0; n < G; n++) acc[n] = make_float2(0.f, 0.f);
  // two k-steps per iteration up to 24 lines: enough loads in flight to cover
  // their latency within the scalar register budget
  constexpr int UK = G > 24 ? 1 : 2;
#pragma unroll UK
  for (int k = 0; k < g; k++) {
    const float2 x = stg[(line0 + k) * SP + p];
#pragma unroll
    for (int n4 = 0; n4 < GS / 4; n4++) {
      const qd_f4v w = *reinterpret_cast<kfloat4 *>(M + k * GS + 4 * n4);
#pragma unroll
      for (int c = 0; c < 4; c++)
        if (4 * n4 + c < G) {
          acc[4 * n4 + c].x = fmaf(x.x, w[c], acc[4 * n4 + c].x);
          acc[4 * n4 + c].y = fmaf(x.y, w[c], acc[4 * n4 + c].y);
        }
    }
  }
#pragma unroll
  for (int n = 0; n < G; n++)
    if (n < g) out(n, acc[n]);
}
// mx_apply with the unit's g outputs rounded up to even, 2 ceil(g / 2) (at
// most GB; rows of the matrix padded to a multiple of 4)
template <int GB, int SP, int NL, typename OUT>
__device__ __forceinline__ void mx_unit(const float2 *stg, int line0, int g, int p, kfloat *M, OUT out) {
#define QD_MXA(G_) mx_apply<G_, SP, NL>(stg, line0, g, p, M, out)
  switch ((g + 1) >> 1) {
    case 1: QD_MXA(2); break;
    case 2: QD_MXA(4); break;
    case 3: if constexpr (GB >= 8) QD_MXA(6); break;
    case 4: if constexpr (GB >= 8) QD_MXA(8); break;
    case 5: if constexpr (GB >= 12) QD_MXA(10); break;
    case 6: if constexpr (GB >= 12) QD_MXA(12); break;
    case 7: if constexpr (GB >= 16) QD_MXA(14); break;
    case 8: if constexpr (GB >= 16) QD_MXA(16); break;
    case 9: if constexpr (GB >= 20) QD_MXA(18); break;
    case 10: if constexpr (GB >= 20) QD_MXA(20); break;
    case 11: if constexpr (GB >= 24) QD_MXA(22); break;
    case 12: if constexpr (GB >= 24) QD_MXA(24); break;
    case 13: if constexpr (GB >= 28) QD_MXA(26); break;
    case 14: if constexpr (GB >= 28) QD_MXA(28); break;
    case 15: if constexpr (GB >= 32) QD_MXA(30); break;
    default: if constexpr (GB >= 32) QD_MXA(32); break;
  }
#undef QD_MXA
}

// Outputs o0 .. o0 + 15 of a unit (row stride gs4 = its padded order): for
// units of more than 32 lines, whose bundle holds that one unit, three threads
// share each point and each keeps 16 accumulators.  The matrix rows are read
// 16 columns at a time, past the row end for the last part (into the next row
// or the zero padding after the last matrix; those outputs are >= gs4 and not
// kept).
template <int SP>
__device__ __forceinline__ void mx_part16(const float2 *stg, int line0, int g, int gs4, int o0, int p, kfloat *M,
                                          float2 (&acc)[16]) {
#pragma unroll
  for (int n = 0; n < 16; n++) acc[n] = make_float2(0.f, 0.f);
#pragma unroll 2
  for (int k = 0; k < g; k++) {
    const float2 x = stg[(line0 + k) * SP + p];
#pragma unroll
    for (int n4 = 0; n4 < 4; n4++) {
      const qd_f4v w = *reinterpret_cast<kfloat4 *>(M + k * gs4 + o0 + 4 * n4);
#pragma unroll
      for (int c = 0; c < 4; c++) {
        acc[4 * n4 + c].x = fmaf(x.x, w[c], acc[4 * n4 + c].x);
        acc[4 * n4 + c].y = fmaf(x.y, w[c], acc[4 * n4 + c].y);
      }
    }
  }
}

// TWO: sized for two resident work groups per CU (at most 128 registers per
// thread, LDS within half the CU's): the lines are loaded at the top of their
// own iteration, and the other work group's bundle covers the latency.
template <int N1, int N2, int NL, int GB, bool TWO>
__global__ __launch_bounds__(NL * N1, TWO ? NL * N1 / 128 : 1) void k_rpass_mx(RMixArgs A) {
  typedef float2 C;
  typedef float4 TB;
  constexpr int N = N1 * N2, NT = NL * N1, XS = N2 + 1;
  constexpr int SP = mx_row_stride<N1, N2, NL>();
  static_assert(N % 64 == 0 && NT % N == 0, "the points of a unit must fill whole waves");
  static_assert(GB % 4 == 0 && GB <= NL && NL <= 64, "unit size bucket");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float2 *s_twa = reinterpret_cast<float2 *>(smem_raw);
  float4 *s_kin = reinterpret_cast<float4 *>(s_twa + N1 * XS);
  C *stg = reinterpret_cast<C *>(s_kin + N);          // [NL][SP]
  __shared__ long long s_off[NL];
  __shared__ int s_unit[NL];
  const int t = threadIdx.x, a = t % N1, L = t / N1;
  const rf_pos<N1, N2, true> P(a);
  rf_tables<float, N1, N2>(s_twa, s_kin, A.tw, A.kin);

  // descriptors of a bundle (registers): header, this thread's line, unit t
  int4 h0, h1;
  long long d_off = 0;
  int d_lam = 0, d_unit = 0;
  auto fetch_desc = [&](int bb, bool has) {
    if (has) {
      const int4 *hp = reinterpret_cast<const int4 *>(A.head + bb);
      h0 = hp[0];
      h1 = hp[1];
      d_off = A.line_off[(size_t)bb * NL + L];
      d_lam = A.line_lam[(size_t)bb * NL + L];
      if (t < NL) d_unit = A.unit[(size_t)bb * NL + t];
    } else {
      h0 = make_int4(0, 0, 0, 0);
      h1 = make_int4(0, 0, 0, 0);
    }
  };
  // lines of a bundle (registers), from its descriptors
  C xn[N2];
  auto fetch_data = [&]() {
    const int nl = h0.x;
    if (L < nl) {
      const C *src = A.data + d_off;
#pragma unroll
      for (int m = 0; m < N2; m++) xn[m] = src[a + N1 * m];
    }
  };

  // With 48-line bundles (768 threads, at most 168 registers per thread) the
  // lines are loaded at the top of their own iteration instead of one bundle
  // ahead, which keeps the 48 mixing accumulators in registers.
  constexpr bool PREF = NT <= 512 && !TWO;
  auto bundle_of = [&](int i) {  // the i-th bundle of this work group (increasing in i)
    return ((i / A.chunk) * (int)gridDim.x + (int)blockIdx.x) * A.chunk + i % A.chunk;
  };
  int b = bundle_of(0);
  if (b < A.nbundle) {
    fetch_desc(b, true);
    if (PREF) fetch_data();
  }
  for (int i = 0; b < A.nbundle; i++, b = bundle_of(i)) {
    if (!PREF) fetch_data();
    const int nl = __builtin_amdgcn_readfirstlane(h0.x), nu = __builtin_amdgcn_readfirstlane(h0.y);
    const int g40 = __builtin_amdgcn_readfirstlane(h1.x), g41 = __builtin_amdgcn_readfirstlane(h1.y);
    const int ms0 = __builtin_amdgcn_readfirstlane(h0.z), ms1 = __builtin_amdgcn_readfirstlane(h0.w);
    const int lam = d_lam;
    __syncthreads();  // the previous bundle is done with LDS (first time: tables)
    if (L < nl) {
#pragma unroll
      for (int m = 0; m < N2; m++) stg[L * SP + a + N1 * m] = xn[m];
      if (a == 0) s_off[L] = d_off;
    }
    if (t < NL) s_unit[t] = d_unit;
    const int bn = bundle_of(i + 1);
    fetch_desc(bn, bn < A.nbundle);
    __syncthreads();
    // K -> l: c_m = sum_K U[K][m] c_K, one (unit, point) per thread, in place
    // (only this thread touches this point of the unit's lines); units above
    // 32 lines: three threads per point (16 outputs each), all reads of a
    // point before any of its writes
    if constexpr (GB > 32) {
      static_assert(NT % (3 * N) == 0, "the three parts of a unit's points fill whole passes");
      for (int i0 = 0; i0 < nu * 3 * N; i0 += NT) {
        const int i = i0 + t;
        const bool act = i < nu * 3 * N;  // uniform per wave
        const int task = i / N, p = i % N, u = act ? task / 3 : 0, part = task % 3;
        const int ud = __builtin_amdgcn_readfirstlane(s_unit[u]);
        const int line0 = ud & 255, g = (ud >> 8) & 255, slot = ud >> 16;
        float2 res[16];
        if (act && slot) {
          kfloat *U = (kfloat *)(A.umat + (slot == 1 ? ms0 : ms1));
          mx_part16<SP>(stg, line0, g, slot == 1 ? g40 : g41, 16 * part, p, U, res);
        }
        __syncthreads();
        if (act && slot) {
#pragma unroll
          for (int n = 0; n < 16; n++)
            if (16 * part + n < g) stg[(line0 + 16 * part + n) * SP + p] = res[n];
        }
        __syncthreads();
      }
    } else
    for (int i0 = 0; i0 < nu * N; i0 += NT) {
      const int i = i0 + t;
      if (i < nu * N) {  // uniform per wave (N is a multiple of 64)
        const int u = i / N, p = i % N;
        const int ud = __builtin_amdgcn_readfirstlane(s_unit[u]);
        const int line0 = ud & 255, g = (ud >> 8) & 255, slot = ud >> 16;
        if (slot) {
          kfloat *U = (kfloat *)(A.umat + (slot == 1 ? ms0 : ms1));
          mx_unit<GB, SP, NL>(stg, line0, g, p, U, [&](int m, C c) { stg[(line0 + m) * SP + p] = c; });
        }
      }
    }
    __syncthreads();
    // the next bundle's lines stream in during the rest
    if (PREF) fetch_data();
    // FFT phase: line L in the l basis
    {
      const bool act = L < nl;
      C *row = stg + L * SP;
      C v[N2];
      TB phv[PREF ? N2 : 1];
      if (act) {
        const TB *ph = A.dR + (size_t)lam * N;
        if constexpr (PREF) {
#pragma unroll
          for (int m = 0; m < N2; m++) phv[m] = ph[P(m)];
#pragma unroll
          for (int m = 0; m < N2; m++) v[m] = rf_ph(row[P(m)], phv[m]);
        } else {
          // four rows at a time: the loads are not all hoisted ahead, which
          // keeps the pass within its register budget without spilling
#pragma unroll
          for (int m0 = 0; m0 < N2; m0 += 4) {
#pragma unroll
            for (int m = m0; m < m0 + 4; m++) v[m] = rf_ph(row[P(m)], ph[P(m)]);
            __builtin_amdgcn_sched_barrier(0);
          }
        }
      }
      // line L's threads are t = N1 L .. N1 L + N1 - 1: one wave
      rf_line<float, N1, N2, true>(v, a, act, row, [](int n, int k) { return n * XS + k; }, s_twa, s_kin);
      if (act) {
        const TB *ph = A.dR + (size_t)lam * N;
#pragma unroll
        for (int m = 0; m < N2; m++) row[P(m)] = rf_ph(v[m], PREF ? phv[PREF ? m : 0] : ph[P(m)]);
      }
    }
    __syncthreads();
    // l -> K and store: c_K = sum_m U[K][m] c_m (rows of U^T), straight to HBM
    if constexpr (GB > 32) {
      for (int i0 = 0; i0 < nu * 3 * N; i0 += NT) {
        const int i = i0 + t;
        if (i < nu * 3 * N) {
          const int task = i / N, p = i % N, u = task / 3, part = task % 3;
          const int ud = __builtin_amdgcn_readfirstlane(s_unit[u]);
          const int line0 = ud & 255, g = (ud >> 8) & 255, slot = ud >> 16;
          if (slot) {
            const int g4 = slot == 1 ? g40 : g41;
            kfloat *UT = (kfloat *)(A.umat + (slot == 1 ? ms0 : ms1) + g4 * g4);
            float2 res[16];
            mx_part16<SP>(stg, line0, g, g4, 16 * part, p, UT, res);
#pragma unroll
            for (int n = 0; n < 16; n++)
              if (16 * part + n < g) A.data[s_off[line0 + 16 * part + n] + p] = res[n];
          } else if (part == 0) {
            for (int K = 0; K < g; K++) A.data[s_off[line0 + K] + p] = stg[(line0 + K) * SP + p];
          }
        }
      }
    } else
    for (int i0 = 0; i0 < nu * N; i0 += NT) {
      const int i = i0 + t;
      if (i < nu * N) {
        const int u = i / N, p = i % N;
        const int ud = __builtin_amdgcn_readfirstlane(s_unit[u]);
        const int line0 = ud & 255, g = (ud >> 8) & 255, slot = ud >> 16;
        if (slot) {
          const int g4 = slot == 1 ? g40 : g41;
          kfloat *UT = (kfloat *)(A.umat + (slot == 1 ? ms0 : ms1) + g4 * g4);
          mx_unit<GB, SP, NL>(stg, line0, g, p, UT, [&](int K, C c) { A.data[s_off[line0 + K] + p] = c; });
        } else {
          for (int K = 0; K < g; K++) A.data[s_off[line0 + K] + p] = stg[(line0 + K) * SP + p];
        }
      }
    }
  }
}

// r pass: lines (column, i) of N = n2 points at stride n1; a batch is NTR / N1
// consecutive i of one column (n1 a multiple of NTR / N1), so the lanes of a
// load read consecutive addresses (128 B per point of the line with 256
// threads, 256 B with 512).  Persistent over batches, with the next batch
// loaded while the current one is transformed.
template <typename T, int N1, int N2, int NTR>
__global__ __launch_bounds__(NTR) void k_rline_rf(RLineArgs<T> A) {
  typedef typename cx_of<T>::type C;
  typedef typename rf_tab<T>::tw TW;
  typedef typename rf_tab<T>::kin TK;
  typedef typename rf_tab<T>::ph TB;
  constexpr int N = N1 * N2, LPB = NTR / N1, LS = LPB + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TW *s_twa = reinterpret_cast<TW *>(smem_raw);
  TK *s_kin = reinterpret_cast<TK *>(s_twa + N1 * (N2 + 1));
  C *xch = reinterpret_cast<C *>(s_kin + N);  // [N1][N2][LS]
  // the last ext.nwg work groups extract the previous step's flux surface
  // (it reads psi, which this pass does not touch) and are done; they are
  // dispatched as the persistent groups finish, into the pass's tail
  const int nblk = (int)gridDim.x - A.ext.nwg, bid = (int)blockIdx.x;
  if (bid >= nblk) {
    flux_extract_at<T>(A.ext, (bid - nblk) * NTR + (int)threadIdx.x);
    return;
  }
  const int t = threadIdx.x;
  const int L = t % LPB, a = t / LPB;
  const rf_pos<N1, N2, sizeof(T) == 4> P(a);
  rf_tables<T, N1, N2>(s_twa, s_kin, A.ltw, A.lkin);
  const size_t st = A.n1;
  // persistent: batch b = LPB consecutive i of one column (n1 % LPB == 0);
  // the next batch's points and D_r row are loaded while this one is
  // transformed, the row through a double-buffered LDS copy
  constexpr int PHE = (N + NTR - 1) / NTR;
  TB *s_ph = reinterpret_cast<TB *>(xch + N1 * N2 * LS);  // [2][N]
  const int nbatch = A.nlines / LPB;
  C raw[N2], v[N2];
  TB phr[PHE];
  // descriptors over the batch's column slab and its row; empty when `has`
  // is false (loads return zeros, no memory traffic)
  auto fetch = [&](int b, int jj, bool has) {
    const int l0 = b * LPB;
    const C *base = A.data + (size_t)(l0 / A.n1) * A.n1 * N + l0 % A.n1;
    const __amdgpu_buffer_rsrc_t rd =
        buf_rsrc(base, has ? (unsigned)(((size_t)(N - 1) * A.n1 + LPB) * sizeof(C)) : 0u);
    const __amdgpu_buffer_rsrc_t rp = buf_rsrc(A.lph + (size_t)jj * N, has ? (unsigned)(N * sizeof(TB)) : 0u);
#pragma unroll
    for (int m = 0; m < N2; m++) raw[m] = buf_ld<C>(rd, (unsigned)(((size_t)P(m) * st + L) * sizeof(C)));
#pragma unroll
    for (int q = 0; q < PHE; q++)
      if (N % NTR == 0 || q * NTR + t < N) phr[q] = buf_ld<TB>(rp, (unsigned)((q * NTR + t) * sizeof(TB)));
  };
  auto j_of = [&](int b) { return __builtin_amdgcn_readfirstlane(A.fcol_j[(b * LPB) / A.n1]); };
  int b = bid;
  if (b < nbatch) fetch(b, j_of(b), true);
  auto xidx = [L](int n, int k) { return (n * N2 + k) * LS + L; };
  int buf = 0;
  for (; b < nbatch; b += nblk) {
    const bool has_n = b + nblk < nbatch;
    const int bn = has_n ? b + nblk : b;
    const int j_n = j_of(bn);  // read while nothing else is outstanding
    TB *ph = s_ph + buf * N;
#pragma unroll
    for (int q = 0; q < PHE; q++)
      if (N % NTR == 0 || q * NTR + t < N) ph[q * NTR + t] = phr[q];
    __syncthreads();  // row (and, first time round, the tables)
    const int line = b * LPB + L;
#pragma unroll
    for (int m = 0; m < N2; m++) v[m] = rf_ph(raw[m], ph[P(m)]);
    fetch(bn, j_n, has_n);
    rf_line<T, N1, N2>(v, a, true, xch, xidx, s_twa, s_kin);
    C *dst = A.data + (size_t)(line / A.n1) * A.n1 * N + line % A.n1;
#pragma unroll
    for (int m = 0; m < N2; m++) dst[(size_t)P(m) * st] = rf_ph(v[m], ph[P(m)]);
    buf ^= 1;
  }
}

template <typename T, int N1, int N2, int NTR>
constexpr size_t rline_rf_lds() {
  return rf_table_bytes<T, N1, N2>() + (size_t)N1 * N2 * (NTR / N1 + 1) * sizeof(typename cx_of<T>::type) +
         (size_t)2 * N1 * N2 * sizeof(typename rf_tab<T>::ph);
}

}  // namespace qd
