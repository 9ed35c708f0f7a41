// qd_tables.hpp -- host-side construction of the angular tables.
//
// Gauss-Legendre grid and Legendre matrices of the J = 0 propagator
// (common/GaussLegendre.F:4-38, common/LegendreP2.m:4-23, timeEvol.C:737-760);
// for J > 0 (no reference counterpart) the associated-Legendre DVR of each K
// block and the eigen-decomposition of the Coriolis coupling.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

namespace qdt {

typedef long double ld;

// Gauss-Legendre nodes (ascending) and weights; Newton on P_n in long double.
inline void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  if (n == 1) {
    w[0] = 2.0;
    return;
  }
  const ld pi = 3.141592653589793238462643383279502884L;
  std::vector<ld> xs(n), ws(n);
  for (int k = 1; k <= n; k++) {
    ld z = cosl(pi * (k - 0.25L) / (n + 0.5L));
    ld dp = 1;
    for (int it = 0; it < 100; it++) {
      ld p0 = 1, p1 = z;
      for (int j = 2; j <= n; j++) {
        const ld p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1);
      const ld dz = p1 / dp;
      z -= dz;
      if (fabsl(dz) < 1e-19L) break;
    }
    ld p0 = 1, p1 = z;
    for (int j = 2; j <= n; j++) {
      const ld p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1);
    xs[k - 1] = z;
    ws[k - 1] = 2 / ((1 - z * z) * dp * dp);
  }
  // ascending, exactly symmetric
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return xs[a] < xs[b]; });
  std::vector<double> xd(n), wd(n);
  for (int i = 0; i < n; i++) {
    xd[i] = (double)xs[idx[i]];
    wd[i] = (double)ws[idx[i]];
  }
  for (int i = 0; i < n; i++) {
    x[i] = 0.5 * (xd[i] - xd[n - 1 - i]);
    w[i] = 0.5 * (wd[i] + wd[n - 1 - i]);
  }
  if (n % 2) x[n / 2] = 0.0;
}

// P_l(x_k), l = 0..m (LegendreP2.m:18-20 recursion): out[l][k]
inline std::vector<double> legendre_p2(int m, const std::vector<double> &x) {
  const int n = (int)x.size();
  std::vector<double> v((size_t)(m + 1) * n);
  for (int k = 0; k < n; k++) {
    v[k] = 1.0;
    if (m >= 1) v[(size_t)n + k] = x[k];
    for (int i = 2; i <= m; i++)
      v[(size_t)i * n + k] = (2 - 1.0 / i) * x[k] * v[(size_t)(i - 1) * n + k] - (1 - 1.0 / i) * v[(size_t)(i - 2) * n + k];
  }
  return v;
}

// Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (row-major),
// eigenvalues ascending, eigenvectors in the columns of V.
inline void jacobi_eigen(int n, std::vector<ld> A, std::vector<ld> &d, std::vector<ld> &V) {
  V.assign((size_t)n * n, 0);
  for (int i = 0; i < n; i++) V[(size_t)i * n + i] = 1;
  for (int sweep = 0; sweep < 100; sweep++) {
    ld off = 0, tot = 0;
    for (int p = 0; p < n; p++)
      for (int q = 0; q < n; q++) {
        const ld a = A[(size_t)p * n + q] * A[(size_t)p * n + q];
        tot += a;
        if (p != q) off += a;
      }
    if (off <= 1e-38L * (tot > 0 ? tot : 1)) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const ld apq = A[(size_t)p * n + q];
        if (apq == 0) continue;
        const ld app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        const ld theta = (aqq - app) / (2 * apq);
        const ld t = (theta >= 0 ? 1 : -1) / (fabsl(theta) + sqrtl(theta * theta + 1));
        const ld c = 1 / sqrtl(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) {  // A <- A J
          const ld akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * akp - s * akq;
          A[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {  // A <- J^T A
          const ld apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * apk - s * aqk;
          A[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          const ld vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return A[(size_t)a * n + a] < A[(size_t)b * n + b]; });
  d.resize(n);
  std::vector<ld> W((size_t)n * n);
  for (int c = 0; c < n; c++) {
    d[c] = A[(size_t)idx[c] * n + idx[c]];
    for (int r = 0; r < n; r++) W[(size_t)r * n + c] = V[(size_t)r * n + idx[c]];
  }
  V.swap(W);
}

// DVR of the normalised associated Legendre functions Pt_j^K, j = K..jmax:
// points = eigenvalues of cos(theta) in that basis (Gauss-Jacobi(K, K) nodes),
// Pb[j][k] = Pt_j^K(x_k), weights w_k with sum_k w_k Pt_j Pt_j' = delta.
// K = 0 uses the Gauss-Legendre grid of the J = 0 propagator.
inline void assoc_legendre_dvr(int K, int jmax, std::vector<double> &x, std::vector<double> &w,
                               std::vector<double> &Pb) {
  const int n = jmax - K + 1;
  if (K == 0) {
    gauss_legendre(n, x, w);
    Pb = legendre_p2(jmax, x);
    for (int j = 0; j <= jmax; j++)
      for (int k = 0; k < n; k++) Pb[(size_t)j * n + k] *= std::sqrt(j + 0.5);
    return;
  }
  std::vector<ld> X((size_t)n * n, 0);
  for (int a = 0; a + 1 < n; a++) {
    const ld j = K + a;
    const ld off = sqrtl(((j + 1) * (j + 1) - (ld)K * K) / ((2 * j + 1) * (2 * j + 3)));
    X[(size_t)a * n + a + 1] = X[(size_t)(a + 1) * n + a] = off;
  }
  std::vector<ld> d, U;
  jacobi_eigen(n, X, d, U);
  // Pt_K^K(x) = c_K (1 - x^2)^(K/2), c_K = sqrt((2K+1)/2) sqrt((2K)!) / (2^K K!)
  const ld logc = 0.5L * logl((2 * K + 1) / 2.0L) + 0.5L * lgammal(2 * K + 1) - K * logl(2.0L) - lgammal(K + 1);
  x.resize(n);
  w.resize(n);
  Pb.assign((size_t)n * n, 0.0);
  for (int k = 0; k < n; k++) {
    const ld xk = d[k];
    const ld pkk = expl(logc) * powl(1 - xk * xk, K / 2.0L);
    ld u0 = U[(size_t)0 * n + k];
    const ld sgn = (u0 * pkk) < 0 ? -1 : 1;
    const ld om = (u0 / pkk) * (u0 / pkk);
    x[k] = (double)xk;
    w[k] = (double)om;
    for (int a = 0; a < n; a++) Pb[(size_t)a * n + k] = (double)(sgn * U[(size_t)a * n + k] / sqrtl(om));
  }
}

// (J - j)^2 in the parity-adapted body-fixed basis K = 0..min(J, j): diagonal
// J(J+1) + j(j+1) - 2K^2, (K, K+1) entry -lambda+(J,K) lambda+(j,K) sqrt(1 + delta_K0).
// Returns U (row K, column m) and the exact l of each eigenvalue l(l+1).
inline void coriolis_eigen(int J, int j, std::vector<double> &U, std::vector<int> &l) {
  const int g = std::min(J, j) + 1;
  std::vector<ld> B((size_t)g * g, 0);
  for (int K = 0; K < g; K++) {
    B[(size_t)K * g + K] = (ld)J * (J + 1) + (ld)j * (j + 1) - 2.0L * K * K;
    if (K + 1 < g) {
      ld c = -sqrtl((ld)J * (J + 1) - (ld)K * (K + 1)) * sqrtl((ld)j * (j + 1) - (ld)K * (K + 1));
      if (K == 0) c *= sqrtl(2.0L);
      B[(size_t)K * g + K + 1] = B[(size_t)(K + 1) * g + K] = c;
    }
  }
  std::vector<ld> d, V;
  jacobi_eigen(g, B, d, V);
  U.resize((size_t)g * g);
  l.resize(g);
  for (int m = 0; m < g; m++) {
    const ld lam = d[m] < 0 ? 0 : d[m];
    l[m] = (int)llroundl((-1 + sqrtl(1 + 4 * lam)) / 2);
    for (int K = 0; K < g; K++) U[(size_t)K * g + m] = (double)V[(size_t)K * g + m];
  }
}

// FP32 pair (fc, fs) for the complex number c + i s of modulus rho, chosen
// among the floats within 4 ulp of c and of s so that fc^2 + fs^2 is as close
// to rho^2 as they allow (ties: the pair nearest the value).  A factor that
// multiplies the same points every step then scales their norm by 1 + O(1e-10)
// per application instead of by the 1 + O(1e-8) of a plain rounding, while its
// phase is off by at most a few FP32 ulps.
inline void round_unit_complex(double c, double s, float &fc, float &fs) {
  const double rho2 = c * c + s * s;
  const float c0 = (float)c, s0 = (float)s;
  float cands_c[9], cands_s[9];
  cands_c[4] = c0;
  cands_s[4] = s0;
  for (int k = 1; k <= 4; k++) {
    cands_c[4 + k] = std::nextafter(cands_c[3 + k], INFINITY);
    cands_c[4 - k] = std::nextafter(cands_c[5 - k], -INFINITY);
    cands_s[4 + k] = std::nextafter(cands_s[3 + k], INFINITY);
    cands_s[4 - k] = std::nextafter(cands_s[5 - k], -INFINITY);
  }
  double best = INFINITY, bdist = INFINITY;
  fc = c0;
  fs = s0;
  for (float a : cands_c)
    for (float b : cands_s) {
      const double err = std::fabs((double)a * a + (double)b * b - rho2);
      const double dist = std::fabs((double)a - c) + std::fabs((double)b - s);
      if (err < best || (err == best && dist < bdist)) {
        best = err;
        bdist = dist;
        fc = a;
        fs = b;
      }
    }
}

// Rotation by theta as three shears (lifting steps),
//   x1 = x - t y,   y' = y + s x1,   x' = x1 - t y',   t = tan(theta/2), s = sin theta,
// whose matrix [[1 - ts, -t(2 - st)], [s, 1 - st]] has determinant 1 whatever
// t and s are: FP32 constants change the angle slightly but never scale the
// norm, and the rotation by -theta with (-t, -s) is its exact inverse.  The
// composite's angle follows from its trace, cos = 1 - t s, so (t, s) is chosen
// among the floats within 8 ulp of tan(theta/2) (each with the three floats
// nearest (1 - cos theta) / t) to make t s closest to 2 sin^2(theta/2).
// |theta| <= pi/2 keeps |t| <= 1 and the intermediate values bounded.
inline void lift_consts(double theta, float &t, float &s) {
  t = s = 0.f;
  if (theta == 0.0) return;
  const bool neg = theta < 0;
  const ld th = neg ? -(ld)theta : (ld)theta;
  const ld sh = sinl(th / 2), ch = cosl(th / 2);
  const ld a = 2 * sh * sh, tex = sh / ch;
  float t0 = (float)tex, tb = t0, sb = (float)(2 * sh * ch);
  ld best = INFINITY;
  for (int k = 0; k < 8; k++) t0 = std::nextafter(t0, -INFINITY);
  for (int k = 0; k <= 16; k++, t0 = std::nextafter(t0, INFINITY)) {
    if (!(t0 > 0)) continue;
    const float s0 = (float)(a / t0);
    const float cs[3] = {std::nextafter(s0, -INFINITY), s0, std::nextafter(s0, INFINITY)};
    for (float sc : cs) {
      const ld e = fabsl((ld)t0 * sc - a);
      if (e < best) {
        best = e;
        tb = t0;
        sb = sc;
      }
    }
  }
  t = neg ? -tb : tb;
  s = neg ? -sb : sb;
}

// theta reduced to (-pi/2, pi/2]: exp(i theta) = sigma exp(i theta_r), sigma = +-1.
inline double lift_reduce(double theta, int &sigma) {
  const ld pi = 3.141592653589793238462643383279502884L;
  ld r = remainderl((ld)theta, 2 * pi);  // [-pi, pi]
  sigma = 1;
  if (r > pi / 2) {
    r -= pi;
    sigma = -1;
  } else if (r <= -pi / 2) {
    r += pi;
    sigma = -1;
  }
  return (double)r;
}

// FP32 copy of an orthogonal g x g matrix U (row K, column m), rounded so that
// every column keeps unit length to second order.  The mixing applies U^T and
// then U to the same point every step; with plain rounding, U_f = U + E, the
// composite's norm changes by 4 sum_m (U^T E)_mm |z_m|^2 per step, coherently
// over thousands of steps.  Here each column's entries are moved between
// their two nearest floats (at most 2 ulp from the value) until
// sum_K U[K][m] E[K][m] is as small as the choices allow.
inline void round_orthogonal_columns(const double *U, int g, std::vector<float> &Q) {
  Q.resize((size_t)g * g);
  for (size_t i = 0; i < Q.size(); i++) Q[i] = (float)U[i];
  for (int m = 0; m < g; m++) {
    double s = 0;
    for (int K = 0; K < g; K++) s += U[(size_t)K * g + m] * ((double)Q[(size_t)K * g + m] - U[(size_t)K * g + m]);
    for (int it = 0; it < 4 * g; it++) {
      double best = std::fabs(s);
      int bk = -1;
      float bv = 0;
      for (int K = 0; K < g; K++) {
        const double u = U[(size_t)K * g + m];
        const float q = Q[(size_t)K * g + m];
        const double lim = 2.0 * std::fabs((double)std::nextafter((float)u, INFINITY) - (double)(float)u);
        for (float v : {std::nextafter(q, INFINITY), std::nextafter(q, -INFINITY)}) {
          if (std::fabs((double)v - u) > lim) continue;
          const double s2 = s + u * ((double)v - (double)q);
          if (std::fabs(s2) < best) {
            best = std::fabs(s2);
            bk = K;
            bv = v;
          }
        }
      }
      if (bk < 0) break;
      const double u = U[(size_t)bk * g + m];
      s += u * ((double)bv - (double)Q[(size_t)bk * g + m]);
      Q[(size_t)bk * g + m] = bv;
    }
  }
}

}  // namespace qdt
