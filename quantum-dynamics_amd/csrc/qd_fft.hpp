// qd_fft.hpp -- compile-time Stockham FFT passes for the grid lengths in use.
//
// Same algorithm as stockham_pass in qd_kernels.hpp (autosort Stockham,
// radices 4, 2, 3, 5 in that order of preference), but with the length N,
// the radix sequence and the threads per line known to the compiler: every
// index is a constant or a shift, every loop unrolls without guards.  The
// R pass and r pass kernels below are the instantiations used for
// N in {16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512}; other lengths run
// through the generic kernels.
#pragma once

#include "qd_kernels.hpp"

namespace qd {

constexpr int ct_radix_of(int m) { return m % 4 == 0 ? 4 : m % 2 == 0 ? 2 : m % 3 == 0 ? 3 : 5; }
constexpr int ct_npass(int n) {
  int c = 0;
  while (n > 1) {
    n /= ct_radix_of(n);
    c++;
  }
  return c;
}
constexpr int ct_radix(int n, int i) {
  for (int k = 0; k < i; k++) n /= ct_radix_of(n);
  return ct_radix_of(n);
}
constexpr int ct_span(int n, int i) {  // product of the radices before pass i
  int s = 1;
  for (int k = 0; k < i; k++) {
    const int r = ct_radix_of(n);
    s *= r;
    n /= r;
  }
  return s;
}
constexpr int ct_maxnb(int n) {
  int m = 0;
  for (int i = 0; i < ct_npass(n); i++) m = (n / ct_radix(n, i)) > m ? (n / ct_radix(n, i)) : m;
  return m;
}
constexpr bool ct_supported(int n) {
  int m = n;
  while (m > 1) {
    const int r = ct_radix_of(m);
    if (m % r) return false;
    m /= r;
  }
  return true;
}

// LDS index with one pad element every 32, against bank conflicts of the
// strided Stockham writes.
template <int N> struct LinePad {
  static constexpr int stride = N + N / 32 + 1;
  __device__ __forceinline__ static int at(int line, int p) { return line * stride + p + (p >> 5); }
};

// One pass with compile-time radix/span; TPL threads per line, slot = line.
template <typename T, int N, int PI, int TPL, bool FWD, typename LIDX, typename MUL>
__device__ __forceinline__ void ct_pass(typename cx_of<T>::type *lds, LIDX lidx, bool act, int line, int bt,
                                        const typename tab_of<T>::type *tw, MUL post) {
  typedef typename cx_of<T>::type C;
  constexpr int r = ct_radix(N, PI);
  constexpr int Ns = ct_span(N, PI);
  constexpr int nb = N / r;
  constexpr int BPT = (nb + TPL - 1) / TPL;
  constexpr T sign = FWD ? (T)-1 : (T)1;
  C v[BPT][r];
#pragma unroll
  for (int bb = 0; bb < BPT; bb++) {
    const int b = bt + bb * TPL;
    if (act && (nb % TPL == 0 || b < nb)) {
#pragma unroll
      for (int q = 0; q < r; q++) v[bb][q] = lds[lidx(line, b + q * nb)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int bb = 0; bb < BPT; bb++) {
    const int b = bt + bb * TPL;
    if (act && (nb % TPL == 0 || b < nb)) {
      const int k = Ns == 1 ? 0 : b % Ns;
      if (Ns > 1) {
        const int estep = k * (N / (Ns * r));
#pragma unroll
        for (int q = 1; q < r; q++) v[bb][q] = tmul(v[bb][q], tw[q * estep], !FWD);
      }
      dft_small<C, T>(v[bb], r, sign);
      const int od = (b / Ns) * Ns * r + k;
#pragma unroll
      for (int q = 0; q < r; q++) {
        const int p = od + q * Ns;
        lds[lidx(line, p)] = post(p, v[bb][q]);
      }
    }
  }
  __syncthreads();
}

template <typename T, int N, int PI, int TPL, bool FWD, bool KIN, typename LIDX>
__device__ __forceinline__ void ct_passes(typename cx_of<T>::type *lds, LIDX lidx, bool act, int line, int bt,
                                          const typename tab_of<T>::type *tw, const typename tab_of<T>::type *kin) {
  typedef typename cx_of<T>::type C;
  if constexpr (PI < ct_npass(N)) {
    if constexpr (KIN && PI == ct_npass(N) - 1) {
      ct_pass<T, N, PI, TPL, FWD>(lds, lidx, act, line, bt, tw, [&](int p, C x) { return tmul(x, kin[p], false); });
    } else {
      ct_pass<T, N, PI, TPL, FWD>(lds, lidx, act, line, bt, tw, [](int, C x) { return x; });
    }
    ct_passes<T, N, PI + 1, TPL, FWD, KIN>(lds, lidx, act, line, bt, tw, kin);
  }
}

// R pass with compile-time length; ML lines per work group of 256 threads.
template <typename T, int N, int ML, bool MIXT>
__global__ __launch_bounds__(256) void k_rpass_ct(RPassArgs<T> A) {
  typedef typename cx_of<T>::type C;
  typedef typename tab_of<T>::type TB;
  constexpr int TPL = 256 / ML;
  static_assert(TPL * FFT_BPT >= ct_maxnb(N), "too few threads per line");
  typedef LinePad<N> LP;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TB *s_tw = reinterpret_cast<TB *>(smem_raw);
  TB *s_kin = s_tw + N;
  C *lds = reinterpret_cast<C *>(s_kin + N);
  __shared__ long long s_off[64];
  __shared__ int s_lam[64];
  __shared__ int s_mixed[64];
  __shared__ int s_unit_line0[65];
  const int t = threadIdx.x;
  const int ub = A.bundle_u0[blockIdx.x], ue = A.bundle_u0[blockIdx.x + 1];
  const int nu = ue - ub;
  const bool mix = MIXT && (A.flags & FFT_MIX);
  const bool phase = A.flags & FFT_PHASE;
  if (t == 0) {
    int L = 0;
    for (int u = 0; u < nu; u++) {
      s_unit_line0[u] = L;
      L += A.units[ub + u].g;
    }
    s_unit_line0[nu] = L;
  }
  for (int e = t; e < N; e += 256) {
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
    s_off[t] = ((long long)col * A.n2 + U.r) * N;
    s_lam[t] = A.fcol_lam[col];
    s_mixed[t] = (mix && U.u0 >= 0) ? 1 : 0;
  }
  __syncthreads();
  for (int e = t; e < nl * N; e += 256) {
    const int L = e / N, p = e % N;
    C x = A.data[s_off[L] + p];
    if (phase && !s_mixed[L]) x = tmul(x, A.dR[(size_t)s_lam[L] * N + p], false);
    lds[LP::at(L, p)] = x;
  }
  __syncthreads();
  if (MIXT && mix) {
    for (int e = t; e < nu * N; e += 256) {
      const int u = e / N, p = e % N;
      const RUnit U = A.units[ub + u];
      if (U.u0 < 0) continue;
      const int L0 = s_unit_line0[u];
      C in[RP_GMAX];
#pragma unroll
      for (int K = 0; K < RP_GMAX; K++)
        if (K < U.g) in[K] = lds[LP::at(L0 + K, p)];
#pragma unroll 1
      for (int m = 0; m < U.g; m++) {
        C acc;
        acc.x = 0;
        acc.y = 0;
#pragma unroll
        for (int K = 0; K < RP_GMAX; K++)
          if (K < U.g) acc = cadd(acc, cscale(in[K], A.umat[U.u0 + K * U.g + m]));
        if (phase) acc = tmul(acc, A.dR[(size_t)s_lam[L0 + m] * N + p], false);
        lds[LP::at(L0 + m, p)] = acc;
      }
    }
    __syncthreads();
  }
  const int line = t / TPL, bt = t % TPL;
  const bool act = line < nl;
  auto lidx = [](int l, int p) { return LP::at(l, p); };
  if (A.flags & FFT_FWD) {
    if (A.flags & FFT_KIN) ct_passes<T, N, 0, TPL, true, true>(lds, lidx, act, line, bt, s_tw, s_kin);
    else ct_passes<T, N, 0, TPL, true, false>(lds, lidx, act, line, bt, s_tw, s_kin);
  }
  if (A.flags & FFT_INV) ct_passes<T, N, 0, TPL, false, false>(lds, lidx, act, line, bt, s_tw, s_kin);
  if (MIXT && mix) {
    for (int e = t; e < nu * N; e += 256) {
      const int u = e / N, p = e % N;
      const RUnit U = A.units[ub + u];
      if (U.u0 < 0) continue;
      const int L0 = s_unit_line0[u];
      C in[RP_GMAX];
#pragma unroll
      for (int m = 0; m < RP_GMAX; m++)
        if (m < U.g) {
          C x = lds[LP::at(L0 + m, p)];
          if (phase) x = tmul(x, A.dR[(size_t)s_lam[L0 + m] * N + p], false);
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
        lds[LP::at(L0 + K, p)] = acc;
      }
    }
    __syncthreads();
  }
  for (int e = t; e < nl * N; e += 256) {
    const int L = e / N, p = e % N;
    C x = lds[LP::at(L, p)];
    if (phase && !s_mixed[L]) x = tmul(x, A.dR[(size_t)s_lam[L] * N + p], false);
    A.data[s_off[L] + p] = x;
  }
}

// r pass with compile-time length; LPB consecutive i of one column per group.
template <typename T, int N, int LPB>
__global__ __launch_bounds__(256) void k_rline_ct(RLineArgs<T> A) {
  typedef typename cx_of<T>::type C;
  typedef typename tab_of<T>::type TB;
  constexpr int TPL = 256 / LPB;
  static_assert(TPL * FFT_BPT >= ct_maxnb(N), "too few threads per line");
  constexpr int LS = LPB + 1;  // padded row of the transposed tile
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TB *s_tw = reinterpret_cast<TB *>(smem_raw);
  TB *s_kin = s_tw + N;
  TB *s_dr = s_kin + N;
  C *lds = reinterpret_cast<C *>(s_dr + N);
  const int t = threadIdx.x;
  const int line_local = t % LPB;
  const int bt = t / LPB;
  const int line = blockIdx.x * LPB + line_local;
  const bool active = line < A.nlines;
  const int lc = active ? line : 0;
  const int i = lc % A.n1;
  const int c = lc / A.n1;
  const size_t base = (size_t)c * A.n1 * N + i;
  const size_t stride = A.n1;
  const int c0 = (blockIdx.x * LPB) / A.n1;
  const int j0 = A.fcol_j[min(c0, A.nlines / A.n1 - 1)];
  const bool phase = A.flags & FFT_PHASE;
  for (int e = t; e < N; e += 256) {
    s_tw[e] = A.tw[e];
    s_kin[e] = A.kin[e];
    if (phase) s_dr[e] = A.dr[(size_t)j0 * N + e];
  }
  __syncthreads();
  auto lidx = [line_local](int, int p) { return p * LS + line_local; };
#pragma unroll
  for (int p0 = 0; p0 < N; p0 += TPL) {
    const int p = p0 + bt;
    if (N % TPL == 0 || p < N) {
      C x;
      x.x = 0;
      x.y = 0;
      if (active) {
        x = A.data[base + (size_t)p * stride];
        if (phase) x = tmul(x, s_dr[p], false);
      }
      lds[p * LS + line_local] = x;
    }
  }
  __syncthreads();
  if (A.flags & FFT_FWD) {
    if (A.flags & FFT_KIN) ct_passes<T, N, 0, TPL, true, true>(lds, lidx, true, line_local, bt, s_tw, s_kin);
    else ct_passes<T, N, 0, TPL, true, false>(lds, lidx, true, line_local, bt, s_tw, s_kin);
  }
  if (A.flags & FFT_INV) ct_passes<T, N, 0, TPL, false, false>(lds, lidx, true, line_local, bt, s_tw, s_kin);
#pragma unroll
  for (int p0 = 0; p0 < N; p0 += TPL) {
    const int p = p0 + bt;
    if ((N % TPL == 0 || p < N) && active) {
      C x = lds[p * LS + line_local];
      if (phase) x = tmul(x, s_dr[p], false);
      A.data[base + (size_t)p * stride] = x;
    }
  }
}

}  // namespace qd
