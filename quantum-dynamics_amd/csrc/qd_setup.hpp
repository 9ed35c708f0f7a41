// qd_setup.hpp -- the set-up the reference does in MATLAB before it calls the
// propagator, restated in C++ for the C ABI (qd_make_radial, qd_make_theta,
// qd_dividing_surface, qd_init_packet).  Header-only and free of HIP, so the
// CPU tests build it on its own.
//
//   gauss_legendre        common/GaussLegendre.F:4-38, common/GaussLegendre2.m:4-32
//   legendre_p2           common/LegendreP2.m:4-23 (stored (m+1) x n, H3/main.m:101-104)
//   assoc_legendre_norm   the normalised associated Legendre functions of the
//                         body-fixed K blocks (the J > 0 extension)
//   diatomic_state        H3/H2VibRotWaveFunction.m:4-47 (radial sine DVR,
//                         Colbert-Miller, on the grid indexed from 1)
//   packet                H3/InitWavePacket.m:2-31
//   eta_sq_plane          common/EtaSq.m:4-18
//   eta_sq_partial_wave   the same projection in partial wave l > 0
//   dividing_surface      H3/main.m:77 (MATLAB int32() rounds half away from zero)
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

namespace qds {

// ---- angular grid -------------------------------------------------------------
// Gauss-Legendre nodes (ascending) and weights on [-1, 1]: Newton on P_n in
// long double from cos(pi (k - 1/4) / (n + 1/2)); w = 2 / ((1 - x^2) P_n'(x)^2);
// made exactly symmetric as tabulated half-grids are.
inline void gauss_legendre(int n, double *x, double *w) {
  if (n == 1) {
    x[0] = 0.0;
    w[0] = 2.0;
    return;
  }
  typedef long double ld;
  const ld pi = 3.141592653589793238462643383279502884L;
  std::vector<ld> xs(n), ws(n);
  for (int k = 1; k <= n; k++) {
    ld z = cosl(pi * (k - 0.25L) / (n + 0.5L)), dp = 0;
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
  std::vector<int> ord(n);
  for (int i = 0; i < n; i++) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return xs[a] < xs[b]; });
  std::vector<double> xd(n), wd(n);
  for (int i = 0; i < n; i++) {
    xd[i] = (double)xs[ord[i]];
    wd[i] = (double)ws[ord[i]];
  }
  for (int i = 0; i < n; i++) {
    x[i] = 0.5 * (xd[i] - xd[n - 1 - i]);
    w[i] = 0.5 * (wd[i] + wd[n - 1 - i]);
  }
  if (n % 2) x[n / 2] = 0.0;
}

// P_l(x_k), l = 0..m, into legendre[l + (m+1) k]: P_i = (2 - 1/i) x P_{i-1} - (1 - 1/i) P_{i-2}
inline void legendre_p2(int m, const double *x, int n, double *legendre) {
  for (int k = 0; k < n; k++) {
    double *col = legendre + (size_t)(m + 1) * k;
    col[0] = 1.0;
    if (m >= 1) col[1] = x[k];
    for (int i = 2; i <= m; i++) col[i] = (2.0 - 1.0 / i) * x[k] * col[i - 1] - (1.0 - 1.0 / i) * col[i - 2];
  }
}

// Normalised associated Legendre function Pt_j^K(x) (int_{-1}^{1} Pt^2 = 1):
// Pt_K^K = c_K (1 - x^2)^(K/2) with c_K = sqrt((2K+1)/2 (2K)!) / (2^K K!) > 0, and
// x Pt_j = a_{j+1} Pt_{j+1} + a_j Pt_{j-1},  a_j = sqrt((j^2 - K^2) / ((2j-1)(2j+1))).
// For K = 0 it is sqrt(j + 1/2) P_j(x).
inline double assoc_legendre_norm(int j, int K, double x) {
  if (j < K) return 0.0;
  const double logc = 0.5 * std::log((2 * K + 1) / 2.0) + 0.5 * std::lgamma(2.0 * K + 1) - K * std::log(2.0) -
                      std::lgamma(K + 1.0);
  double p0 = 0.0, p1 = std::exp(logc) * std::pow(1.0 - x * x, 0.5 * K);
  auto a = [K](int jj) { return std::sqrt(((double)jj * jj - (double)K * K) / ((2.0 * jj - 1) * (2.0 * jj + 1))); };
  for (int jj = K; jj < j; jj++) {
    const double p2 = (x * p1 - a(jj) * p0) / a(jj + 1);
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

// ---- symmetric eigenproblem --------------------------------------------------------
// The k-th smallest eigenpair (0-based) of the symmetric n x n matrix A (row
// major, overwritten): Householder reduction to tridiagonal form, the
// eigenvalue by bisection on the Sturm sequence, its eigenvector by inverse
// iteration on the tridiagonal matrix, transformed back.
inline void sym_eigenpair(std::vector<double> &A, int n, int k, double *lambda, std::vector<double> &vec) {
  std::vector<double> d(n), e(std::max(n - 1, 1));
  std::vector<std::vector<double>> hv;  // Householder vectors (of rows k+1..n-1)
  for (int c = 0; c + 2 < n; c++) {
    const int m = n - c - 1;
    std::vector<double> v(m);
    double nrm = 0;
    for (int i = 0; i < m; i++) {
      v[i] = A[(size_t)(c + 1 + i) * n + c];
      nrm += v[i] * v[i];
    }
    nrm = std::sqrt(nrm);
    d[c] = A[(size_t)c * n + c];
    if (nrm == 0.0) {
      e[c] = 0.0;
      hv.push_back(std::vector<double>(m, 0.0));
      continue;
    }
    const double alpha = v[0] > 0 ? -nrm : nrm;
    e[c] = alpha;
    v[0] -= alpha;
    double vn = 0;
    for (double t : v) vn += t * t;
    vn = std::sqrt(vn);
    for (double &t : v) t /= vn;
    // trailing block B <- H B H, H = I - 2 v v^T:  w = 2 B v, w -= (v^T w) v, B -= v w^T + w v^T
    std::vector<double> w(m, 0.0);
    for (int i = 0; i < m; i++) {
      const double *row = &A[(size_t)(c + 1 + i) * n + c + 1];
      double s = 0;
      for (int j = 0; j < m; j++) s += row[j] * v[j];
      w[i] = 2 * s;
    }
    double vw = 0;
    for (int i = 0; i < m; i++) vw += v[i] * w[i];
    for (int i = 0; i < m; i++) w[i] -= vw * v[i];
    for (int i = 0; i < m; i++) {
      double *row = &A[(size_t)(c + 1 + i) * n + c + 1];
      for (int j = 0; j < m; j++) row[j] -= v[i] * w[j] + w[i] * v[j];
    }
    hv.push_back(v);
  }
  if (n >= 2) {
    d[n - 2] = A[(size_t)(n - 2) * n + n - 2];
    d[n - 1] = A[(size_t)(n - 1) * n + n - 1];
    e[n - 2] = A[(size_t)(n - 1) * n + n - 2];
  } else {
    d[0] = A[0];
  }
  // eigenvalue k by bisection: count(l) = eigenvalues below l (Sturm sequence)
  auto count = [&](double l) {
    int c = 0;
    double q = d[0] - l;
    if (q < 0) c++;
    for (int i = 1; i < n; i++) {
      if (q == 0) q = 1e-300;
      q = d[i] - l - e[i - 1] * e[i - 1] / q;
      if (q < 0) c++;
    }
    return c;
  };
  double lo = d[0], hi = d[0];
  for (int i = 0; i < n; i++) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0) + (i + 1 < n ? std::fabs(e[i]) : 0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  for (int it = 0; it < 200 && hi - lo > 0; it++) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (count(mid) > k) hi = mid;
    else lo = mid;
  }
  const double lam = 0.5 * (lo + hi);
  *lambda = lam;
  // inverse iteration: (T - lam I) y = b, Gaussian elimination with partial pivoting
  std::vector<double> y(n, 1.0 / std::sqrt((double)n));
  const double shift = lam + 1e-14 * std::max(1.0, std::fabs(lam));
  for (int it = 0; it < 4; it++) {
    // rows i: sub a[i], diag b[i], super c[i], second super f[i] (from pivoting)
    std::vector<double> a(n, 0.0), b(n), cc(n, 0.0), f(n, 0.0), rhs = y;
    for (int i = 0; i < n; i++) {
      b[i] = d[i] - shift;
      if (i > 0) a[i] = e[i - 1];
      if (i + 1 < n) cc[i] = e[i];
    }
    for (int i = 0; i + 1 < n; i++) {
      // eliminate a[i+1] with row i, swapping rows if row i+1 has the larger pivot
      if (std::fabs(a[i + 1]) > std::fabs(b[i])) {
        std::swap(b[i], a[i + 1]);
        std::swap(cc[i], b[i + 1]);
        std::swap(f[i], cc[i + 1]);
        std::swap(rhs[i], rhs[i + 1]);
      }
      if (b[i] == 0.0) b[i] = 1e-300;
      const double m = a[i + 1] / b[i];
      a[i + 1] = 0.0;
      b[i + 1] -= m * cc[i];
      cc[i + 1] -= m * f[i];
      rhs[i + 1] -= m * rhs[i];
    }
    if (b[n - 1] == 0.0) b[n - 1] = 1e-300;
    for (int i = n - 1; i >= 0; i--) {
      double s = rhs[i];
      if (i + 1 < n) s -= cc[i] * y[i + 1];
      if (i + 2 < n) s -= f[i] * y[i + 2];
      y[i] = s / b[i];
    }
    double nrm = 0;
    for (double t : y) nrm += t * t;
    nrm = std::sqrt(nrm);
    for (double &t : y) t /= nrm;
  }
  // back to the original basis: x = H_0 H_1 ... H_{n-3} y
  for (int c = (int)hv.size() - 1; c >= 0; c--) {
    const std::vector<double> &v = hv[c];
    double s = 0;
    for (size_t i = 0; i < v.size(); i++) s += v[i] * y[c + 1 + i];
    for (size_t i = 0; i < v.size(); i++) y[c + 1 + i] -= 2 * s * v[i];
  }
  vec = y;
}

// ---- the diatomic state ---------------------------------------------------------
// Eigenpair v of the diatomic on the r grid: the radial sine-DVR kinetic
// matrix indexed from 1 (T_ii = pi^2/3 - 1/(2 i^2), T_ij = (-1)^(i-j) (2/(i-j)^2 -
// 2/(i+j)^2), over 2 mu dr^2) plus V(r) + j(j+1)/(2 mu r^2) on the diagonal;
// psi normalised to sum |psi|^2 dr = 1, with its largest-magnitude value positive.
inline void diatomic_state(const double *r, int n, double dr, double mu, const double *V, int j, int v, double *e,
                           double *psi) {
  std::vector<double> H((size_t)n * n);
  const double s = 1.0 / (2 * mu * dr * dr);
  const double pi = 3.141592653589793238462643383279502884;
  for (int a = 1; a <= n; a++)
    for (int b = 1; b <= n; b++) {
      double t;
      if (a == b) t = pi * pi / 3 - 1.0 / (2.0 * a * a);
      else {
        const double dm = (double)(a - b), dp = (double)(a + b);
        t = ((a - b) % 2 ? -1.0 : 1.0) * (2.0 / (dm * dm) - 2.0 / (dp * dp));
      }
      H[(size_t)(a - 1) * n + (b - 1)] = t * s;
    }
  for (int a = 0; a < n; a++) H[(size_t)a * n + a] += V[a] + j * (j + 1.0) / (2 * mu * r[a] * r[a]);
  std::vector<double> vec;
  sym_eigenpair(H, n, v, e, vec);
  int imax = 0;
  for (int a = 1; a < n; a++)
    if (std::fabs(vec[a]) > std::fabs(vec[imax])) imax = a;
  const double sg = vec[imax] < 0 ? -1.0 : 1.0;
  for (int a = 0; a < n; a++) psi[a] = sg * vec[a] / std::sqrt(dr);
}

// ---- the incoming packet and its normalisation ---------------------------------------
// g(R) = (pi delta^2)^(-1/4) exp(-(R - R0)^2 / (2 delta^2) - i k0 R)
inline std::complex<double> packet_g(double R, double R0, double delta, double k0) {
  const double pi = 3.141592653589793238462643383279502884;
  const double a = std::pow(1.0 / (pi * delta * delta), 0.25) * std::exp(-(R - R0) * (R - R0) / (2 * delta * delta));
  return std::complex<double>(a * std::cos(-k0 * R), a * std::sin(-k0 * R));
}

// |eta(E)|^2 = sqrt(pi) (2 mu delta / k) exp(-delta^2 (k - k0)^2), k = sqrt(2 mu E)
// (E the translational energy; 0 for E <= 0)
inline double eta_sq_plane(double delta, double k0, double mu, double E) {
  if (!(E > 0)) return 0.0;
  const double pi = 3.141592653589793238462643383279502884;
  const double k = std::sqrt(2 * mu * E);
  return std::sqrt(pi) * 2 * mu * delta / k * std::exp(-delta * delta * (k - k0) * (k - k0));
}

// w_l(x) = x h_l^(1)(x) by w_{l+1} = (2l+1)/x w_l - w_{l-1}, w_0 = -i e^{ix}, w_1 = -(1 + i/x) e^{ix}
inline std::complex<double> riccati_hankel_plus(int l, double x) {
  const std::complex<double> ex(std::cos(x), std::sin(x));
  std::complex<double> w0 = std::complex<double>(0, -1) * ex;
  if (l == 0) return w0;
  std::complex<double> w1 = -(1.0 + std::complex<double>(0, 1.0 / x)) * ex;
  for (int m = 1; m < l; m++) {
    const std::complex<double> w2 = ((2.0 * m + 1) / x) * w1 - w0;
    w0 = w1;
    w1 = w2;
  }
  return w1;
}

// |eta_l(E)|^2 = (mu / k) |int w_l(k R) g(R) dR|^2: the packet's projection onto
// the energy-normalised incoming free wave of orbital angular momentum l, in
// EtaSq.m's normalisation (equal to eta_sq_plane at l = 0).  Trapezoid over
// R0 +- 10 delta (n_quad points), where the Gaussian weight has decayed.
inline double eta_sq_partial_wave(double R0, double delta, double k0, double mu, int l, double E,
                                  int n_quad = 4001) {
  if (!(E > 0)) return 0.0;
  const double k = std::sqrt(2 * mu * E);
  const double a = R0 - 10 * delta, b = R0 + 10 * delta, h = (b - a) / (n_quad - 1);
  std::complex<double> s(0, 0);
  for (int q = 0; q < n_quad; q++) {
    const double R = a + q * h;
    if (R <= 0) continue;
    const std::complex<double> f = riccati_hankel_plus(l, k * R) * packet_g(R, R0, delta, k0);
    s += (q == 0 || q == n_quad - 1) ? 0.5 * f : f;
  }
  s *= h;
  return mu / k * std::norm(s);
}

// int32((rd - rmin) / dr) with MATLAB's rounding (half away from zero)
inline int dividing_surface(double rd, double rmin, double dr) {
  const double v = (rd - rmin) / dr;
  const int mag = (int)std::floor(std::fabs(v) + 0.5);
  return v >= 0 ? mag : -mag;
}

}  // namespace qds
