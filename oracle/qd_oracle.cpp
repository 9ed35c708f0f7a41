/*
 * qd_oracle.cpp -- CPU restatement of the reference propagator.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see qd_oracle.h).
 *
 * What it restates, routine by routine:
 *   TimeEvolution::time_evolution / evolution_dt  common/timeEvol.C:619-685
 *   phase tables                                   common/timeEvol.C:438-517
 *   pre-half-step (Complex sqrt via asin-arg)      common/timeEvol.C:519-538,
 *                                                  common/complex.h:60-68,126-159
 *   dump                                           common/timeEvol.C:687-735
 *   Legendre weights                               common/timeEvol.C:737-760
 *   flux surface, time->energy, CRP                common/timeEvol.C:762-878,
 *                                                  common/CRPAux.F, common/Gradient3.F
 *   derived radial tables                          common/MatlabStructures.C:17-36
 *   momentum grid                                  common/fftwinterface.C:110-125
 *   CUDA-variant switches                          common-cuda/evolutionCUDA2.cu:487-552,898-994
 * The reference's FFT is FFTW 3.3.4 (not vendored); here a plain mixed-radix
 * Cooley-Tukey DFT with long-double twiddles stands in for it.  The complex
 * arithmetic follows the reference's Complex class operation by operation so
 * that the table-building code rounds the same way.
 *
 * Compiled with -ffp-contract=off so that no multiply-add is fused: the
 * compiled Fortran kernels of the reference (plain x86-64, no FMA) round the
 * same way and the golden vectors match bit for bit.
 */
#include "qd_oracle.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <map>
#include <mutex>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kPi = M_PI;  /* complex.h:12-14 uses M_PI */

/* ---- complex arithmetic of common/complex.h ---- */
struct Cx {
  double re, im;
};
inline Cx cx(double r, double i) { return Cx{r, i}; }
inline Cx mul(const Cx &a, const Cx &b) {  /* complex.h:122-123 */
  return Cx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline Cx mul(const Cx &a, double r) { return Cx{a.re * r, a.im * r}; }  /* :118-119 */
inline Cx divr(const Cx &a, double r) { return Cx{a.re / r, a.im / r}; } /* :126-127 */
inline double abs2(const Cx &c) { return c.re * c.re + c.im * c.im; }   /* :45-46 */
inline Cx conj(const Cx &c) { return Cx{c.re, -c.im}; }
inline Cx cexp_ref(const Cx &c) {  /* :133-134: exp(re) * Complex(cos(im), sin(im)) */
  const double e = std::exp(c.re);
  return Cx{std::cos(c.im) * e, std::sin(c.im) * e};
}
inline double arg_ref(const Cx &c) {  /* :60-68 */
  const double angle = std::asin(c.im / std::sqrt(abs2(c)));
  if (c.re < 0 && c.im > 0) return kPi - angle;
  if (c.re < 0 && c.im < 0) return -(kPi + angle);
  return angle;
}
inline Cx sqrt_ref(const Cx &c) {  /* :148-159: pow(c, 0.5) */
  const double theta = 0.5 * arg_ref(c);
  const double r = std::pow(abs2(c), 0.5 / 2.0);
  return Cx{r * std::cos(theta), r * std::sin(theta)};
}
inline Cx cdiv_ref(const Cx &a, const Cx &b) {  /* :130-131: c1/abs2(c2) * conj(c2) */
  return mul(divr(a, abs2(b)), conj(b));
}

/* ---- mixed-radix DFT ---- */
struct FftPlan {
  int n = 0;
  std::vector<double> wre, wim;  /* exp(-2 pi i j / n), j < n */
  std::vector<int> factors;
};

std::mutex g_plan_mutex;
std::map<int, FftPlan *> g_plans;

const FftPlan *get_plan(int n) {
  std::lock_guard<std::mutex> lock(g_plan_mutex);
  auto it = g_plans.find(n);
  if (it != g_plans.end()) return it->second;
  FftPlan *p = new FftPlan;
  p->n = n;
  p->wre.resize(n);
  p->wim.resize(n);
  for (int j = 0; j < n; j++) {
    const long double a = -2.0L * 3.141592653589793238462643383279502884L * j / n;
    p->wre[j] = (double)cosl(a);
    p->wim[j] = (double)sinl(a);
  }
  int m = n;
  while (m > 1) {
    int f;
    if (m % 4 == 0) f = 4;
    else if (m % 2 == 0) f = 2;
    else if (m % 3 == 0) f = 3;
    else if (m % 5 == 0) f = 5;
    else {
      f = 7;
      while (m % f) f += 2;
    }
    p->factors.push_back(f);
    m /= f;
  }
  g_plans[n] = p;
  return p;
}

/* Decimation in time: out[0..n) = DFT of in[0], in[s], ..., in[(n-1)s].
 * tw_step maps this sub-transform's roots onto the full table. */
void fft_rec(const FftPlan &P, int level, int n, int sign, const Cx *in, int s, Cx *out,
             int tw_step) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  const int p = P.factors[level];
  const int m = n / p;
  for (int q = 0; q < p; q++) fft_rec(P, level + 1, m, sign, in + q * s, s * p, out + q * m, tw_step * p);
  Cx ybuf[16], tbuf[16];
  std::vector<Cx> yv, tv;
  Cx *y = ybuf, *t = tbuf;
  if (p > 16) {
    yv.resize(p);
    tv.resize(p);
    y = yv.data();
    t = tv.data();
  }
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      const int e = (q * k * tw_step) % P.n;
      const Cx w = Cx{P.wre[e], sign < 0 ? P.wim[e] : -P.wim[e]};
      y[q] = q == 0 ? out[q * m + k] : mul(out[q * m + k], w);
    }
    for (int u = 0; u < p; u++) {
      Cx acc = y[0];
      for (int q = 1; q < p; q++) {
        const int e = ((q * u) % p) * (P.n / p);
        const Cx w = Cx{P.wre[e], sign < 0 ? P.wim[e] : -P.wim[e]};
        const Cx z = mul(y[q], w);
        acc.re += z.re;
        acc.im += z.im;
      }
      t[u] = acc;
    }
    for (int u = 0; u < p; u++) out[k + m * u] = t[u];
  }
}

void fft_strided(int n, int sign, Cx *data, int stride, Cx *work_in, Cx *work_out) {
  const FftPlan &P = *get_plan(n);
  for (int j = 0; j < n; j++) work_in[j] = data[(size_t)j * stride];
  fft_rec(P, 0, n, sign, work_in, 1, work_out, 1);
  for (int j = 0; j < n; j++) data[(size_t)j * stride] = work_out[j];
}

/* 2-D DFT of a column-major n1 x n2 slab: along i for each j, then along j. */
void fft_2d_slab(int n1, int n2, int sign, Cx *slab, std::vector<Cx> &w1, std::vector<Cx> &w2) {
  const int nm = n1 > n2 ? n1 : n2;
  if ((int)w1.size() < nm) w1.resize(nm);
  if ((int)w2.size() < nm) w2.resize(nm);
  for (int j = 0; j < n2; j++) fft_strided(n1, sign, slab + (size_t)j * n1, 1, w1.data(), w2.data());
  for (int i = 0; i < n1; i++) fft_strided(n2, sign, slab + i, n1, w1.data(), w2.data());
}

thread_local char g_err[256];

/* Gradient stencil coefficients.  For 3..11 points the reference's integer
 * numerators and common denominator (Gradient3.F:31-167) are used so the
 * rounding follows it; wider stencils use the closed form
 * c_j = (-1)^(j+1) (h!)^2 / (j (h-j)! (h+j)!)  (gradients2.cu:18-35). */
struct Stencil {
  int h = 0;
  double num[31];
  double den = 1;
  bool integer_form = false;
};

bool make_stencil(int npoints, Stencil &st) {
  if (npoints < 3 || npoints > 61 || (npoints % 2) == 0) return false;
  st.h = npoints / 2;
  static const double t3[] = {1}, t5[] = {8, -1}, t7[] = {45, -9, 1}, t9[] = {672, -168, 32, -3},
                      t11[] = {2100, -600, 150, -25, 2};
  const double *t = nullptr;
  switch (npoints) {
    case 3: t = t3; st.den = 2; break;
    case 5: t = t5; st.den = 12; break;
    case 7: t = t7; st.den = 60; break;
    case 9: t = t9; st.den = 840; break;
    case 11: t = t11; st.den = 2520; break;
    default: break;
  }
  if (t) {
    st.integer_form = true;
    for (int j = 0; j < st.h; j++) st.num[j] = t[j];
    return true;
  }
  st.integer_form = false;
  st.den = 1;
  for (int j = 1; j <= st.h; j++) {
    long double c = 1.0L;  /* (h!)^2 / ((h-j)! (h+j)!) = prod_{i=1..j} (h-j+i)/(h+i) */
    for (int i = 1; i <= j; i++) c *= (long double)(st.h - j + i) / (long double)(st.h + i);
    c /= j;
    st.num[j - 1] = (double)((j % 2) ? c : -c);
  }
  return true;
}

int set_err(int code, const char *msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}

}  // namespace

extern "C" {

void qdo_fft_1d(int n, int sign, double *data, int stride) {
  std::vector<Cx> a(n), b(n);
  fft_strided(n, sign, reinterpret_cast<Cx *>(data), stride, a.data(), b.data());
}

void qdo_fft_2d(int n1, int n2, int sign, double *data) {
  std::vector<Cx> w1, w2;
  fft_2d_slab(n1, n2, sign, reinterpret_cast<Cx *>(data), w1, w2);
}

int qdo_momentum(int nm, double xl, double *p) {
  if (nm / 2 * 2 != nm) return QD_ERR_ODD_N;  /* fftwinterface.C:112 */
  const double two_Pi_xl = 2.0 * kPi / xl;
  const int nm_half = nm / 2;
  double p1 = 0;
  for (int i = 0; i <= nm_half; i++) {  /* running sum, as the reference does */
    p[i] = p1;
    p1 += two_Pi_xl;
  }
  p1 = two_Pi_xl * (-nm_half + 1);
  for (int i = nm_half + 1; i < nm; i++) {
    p[i] = p1;
    p1 += two_Pi_xl;
  }
  return QD_OK;
}

void qdo_forward_legendre(const double *cpsi_, double *lpsi_, int nr, int ntheta, int nleg,
                          const double *wlegp) {
  const Cx *cpsi = reinterpret_cast<const Cx *>(cpsi_);
  Cx *lpsi = reinterpret_cast<Cx *>(lpsi_);
  memset(lpsi, 0, sizeof(Cx) * (size_t)nr * nleg);  /* DCZeros, LegTransform.F:15 */
#pragma omp parallel for schedule(static, 1)
  for (int l = 0; l < nleg; l++) {
    Cx *out = lpsi + (size_t)l * nr;
    for (int k = 0; k < ntheta; k++) {
      const double w = wlegp[k + (size_t)ntheta * l];
      const Cx *in = cpsi + (size_t)k * nr;
      for (int i = 0; i < nr; i++) {
        out[i].re += in[i].re * w;
        out[i].im += in[i].im * w;
      }
    }
  }
}

void qdo_backward_legendre(double *cpsi_, const double *lpsi_, int nr, int ntheta, int nleg,
                           const double *legp) {
  Cx *cpsi = reinterpret_cast<Cx *>(cpsi_);
  const Cx *lpsi = reinterpret_cast<const Cx *>(lpsi_);
  memset(cpsi, 0, sizeof(Cx) * (size_t)nr * ntheta);  /* LegTransform.F:43 */
#pragma omp parallel for schedule(static, 1)
  for (int k = 0; k < ntheta; k++) {
    Cx *out = cpsi + (size_t)k * nr;
    for (int l = 0; l < nleg; l++) {
      const double p = legp[l + (size_t)nleg * k];
      const Cx *in = lpsi + (size_t)l * nr;
      for (int i = 0; i < nr; i++) {
        out[i].re += in[i].re * p;
        out[i].im += in[i].im * p;
      }
    }
  }
}

int qdo_gradient_coefficients(int npoints, double *c) {
  Stencil st;
  if (!make_stencil(npoints, st)) return QD_ERR_STENCIL;
  for (int j = 0; j < st.h; j++) c[j] = st.num[j] / st.den;
  return QD_OK;
}

int qdo_gradient3d(int n1, int n2, int n3, int n2p, int base, double dx, const double *f_,
                   double *v_, double *g_, int npoints) {
  Stencil st;
  if (!make_stencil(npoints, st)) return set_err(QD_ERR_STENCIL, "NPoints error");
  const int s = n2p - base;  /* 0-based surface row */
  if (s - st.h < 0 || s + st.h >= n2) return set_err(QD_ERR_SURFACE, "surface stencil outside r grid");
  const Cx *f = reinterpret_cast<const Cx *>(f_);
  Cx *v = reinterpret_cast<Cx *>(v_);
  Cx *g = reinterpret_cast<Cx *>(g_);
#pragma omp parallel for schedule(static, 1)
  for (int k = 0; k < n3; k++) {
    for (int i = 0; i < n1; i++) {
      auto F = [&](int j) { return f[i + (size_t)n1 * (j + (size_t)n2 * k)]; };
      v[i + (size_t)n1 * k] = F(s);
      Cx acc{0, 0};
      /* The reference sums from the outermost pair inwards (Gradient3.F:80-84). */
      for (int j = st.h; j >= 1; j--) {
        const Cx a = F(s + j), b = F(s - j);
        const Cx d{a.re - b.re, a.im - b.im};
        const double c = st.num[j - 1];
        if (st.integer_form && c == 1.0) {
          acc.re += d.re;
          acc.im += d.im;
        } else if (st.integer_form && c == -1.0) {
          acc.re -= d.re;
          acc.im -= d.im;
        } else {
          acc.re += c * d.re;
          acc.im += c * d.im;
        }
      }
      const double den = st.den * dx;
      g[i + (size_t)n1 * k] = Cx{acc.re / den, acc.im / den};
    }
  }
  return QD_OK;
}

void qdo_psi_time_to_energy(int n1, int ntheta, int ne, const double *e_, const double *psit_,
                            const double *dpsit_, double *fai_, double *dfai_) {
  const Cx *e = reinterpret_cast<const Cx *>(e_);
  const Cx *psit = reinterpret_cast<const Cx *>(psit_);
  const Cx *dpsit = reinterpret_cast<const Cx *>(dpsit_);
  Cx *fai = reinterpret_cast<Cx *>(fai_);
  Cx *dfai = reinterpret_cast<Cx *>(dfai_);
  const size_t nx = (size_t)n1 * ntheta;
#pragma omp parallel for schedule(static, 1)
  for (int ie = 0; ie < ne; ie++) {
    Cx *f = fai + nx * ie;
    Cx *df = dfai + nx * ie;
    for (size_t x = 0; x < nx; x++) {
      const Cx a = mul(e[ie], psit[x]);
      const Cx b = mul(e[ie], dpsit[x]);
      f[x].re += a.re;
      f[x].im += a.im;
      df[x].re += b.re;
      df[x].im += b.im;
    }
  }
}

void qdo_calculate_crp(int n1, int ntheta, int ne, double dr1, double mu2, const double *w,
                       const double *eta_sq, const double *fai_, const double *dfai_, double *crp) {
  const Cx *fai = reinterpret_cast<const Cx *>(fai_);
  const Cx *dfai = reinterpret_cast<const Cx *>(dfai_);
  const double dr1mu2 = dr1 / mu2;
  const size_t nx = (size_t)n1 * ntheta;
#pragma omp parallel for schedule(static, 1)
  for (int ie = 0; ie < ne; ie++) {
    Cx ce{0, 0};
    for (int k = 0; k < ntheta; k++) {
      Cx c3{0, 0};
      for (int i = 0; i < n1; i++) {
        const size_t x = i + (size_t)n1 * k + nx * ie;
        const Cx z = mul(conj(fai[x]), dfai[x]);
        c3.re += z.re;
        c3.im += z.im;
      }
      ce.re += w[k] * c3.re;
      ce.im += w[k] * c3.im;
    }
    crp[ie] = ce.im / eta_sq[ie] * dr1mu2;
  }
}

/* ------------------------------------------------------------------ */
/* Whole propagation                                                    */
/* ------------------------------------------------------------------ */

int qdo_time_evolution(qd_problem *P, const qdo_options *opt_in, qdo_diag *diag) {
  qdo_options opt{};
  if (opt_in) opt = *opt_in;
  const auto t_start = std::chrono::steady_clock::now();
#ifdef _OPENMP
  if (opt.nthreads > 0) omp_set_num_threads(opt.nthreads);
#endif
  if (!P || !P->pot || !P->psi || !P->r1.r || !P->r2.r || !P->theta.x || !P->theta.w ||
      !P->theta.legendre)
    return set_err(QD_ERR_NULL, "missing array");
  const int n1 = P->r1.n, n2 = P->r2.n, nth = P->theta.n, m = P->theta.m + 1;
  if (n1 <= 0 || n2 <= 0 || nth <= 0 || m <= 0) return set_err(QD_ERR_ARG, "bad sizes");
  if (n1 % 2 || n2 % 2) return set_err(QD_ERR_ODD_N, "odd grid size");
  const bool cuda = opt.variant == QD_VARIANT_CUDA;
  const double dt = P->time.time_step;
  const double dr1 = P->r1.dr, dr2 = P->r2.dr;
  const size_t n12 = (size_t)n1 * n2;
  const size_t npt = n12 * nth;

  /* MatlabStructures.C:17-36: psq2m = p^2/2mu, one2mr2 = 1/(2 mu r^2). */
  std::vector<double> psq1(n1), psq2(n2), I1(n1), I2(n2);
  qdo_momentum(n1, n1 * dr1, psq1.data());
  qdo_momentum(n2, n2 * dr2, psq2.data());
  for (int i = 0; i < n1; i++) psq1[i] = psq1[i] * psq1[i] / (2 * P->r1.mass);
  for (int i = 0; i < n2; i++) psq2[i] = psq2[i] * psq2[i] / (2 * P->r2.mass);
  {
    const double m2 = P->r1.mass + P->r1.mass;
    for (int i = 0; i < n1; i++) {
      /* CUDA variant builds R_i as r_left + i dr (cumath.h:153-160). */
      const double r = cuda ? P->r1.r[0] + i * dr1 : P->r1.r[i];
      I1[i] = 1.0 / (m2 * r * r);
    }
    const double m22 = P->r2.mass + P->r2.mass;
    for (int j = 0; j < n2; j++) {
      const double r = cuda ? P->r2.r[0] + j * dr2 : P->r2.r[j];
      I2[j] = 1.0 / (m22 * r * r);
    }
  }

  Cx *psi = reinterpret_cast<Cx *>(P->psi);
  const double *pot = P->pot;
  const double *w = P->theta.w;
  const double *leg = P->theta.legendre;

  /* setup_exp_ipot_dt (timeEvol.C:438-463): exp(Complex(0,-dt) * v). */
  std::vector<Cx> expV(npt);
  {
    const Cx Idt{0.0, -dt};
#pragma omp parallel for schedule(static)
    for (size_t x = 0; x < npt; x++) expV[x] = cexp_ref(mul(Idt, pot[x]));
  }
  /* setup_exp_irot_dt_2 (:465-491): exp(l(l+1)(I1+I2) * Complex(0,-dt/2)). */
  std::vector<Cx> expRot((size_t)n12 * m);
  {
    const Cx Idt2{0.0, -dt / 2};
#pragma omp parallel for schedule(static, 1)
    for (int l = 0; l < m; l++)
      for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++)
          expRot[i + (size_t)n1 * j + n12 * l] = cexp_ref(mul(Idt2, l * (l + 1) * (I1[i] + I2[j])));
  }
  /* setup_exp_ikin_dt (:493-517): exp((T1+T2) * Complex(0,-dt)) / n1 / n2.
   * The CUDA variant scales by 1/(n1 n2) after the inverse FFT instead (:682-690). */
  std::vector<Cx> expKin(n12);
  {
    const Cx Idt{0.0, -dt};
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        Cx c = cexp_ref(mul(Idt, psq1[i] + psq2[j]));
        if (!cuda) c = divr(divr(c, (double)n1), (double)n2);
        expKin[i + (size_t)n1 * j] = c;
      }
  }
  /* setup_weight_legendre (:737-760): wp(k,l) = (l+0.5) w_k P_l(x_k). */
  std::vector<double> wleg((size_t)nth * m);
  for (int l = 0; l < m; l++) {
    const double f = l + 0.5;
    for (int k = 0; k < nth; k++) wleg[k + (size_t)nth * l] = f * w[k] * leg[l + (size_t)m * k];
  }
  /* setup_dump (:687-711): applied only when both arrays exist (timeEvol.h:111). */
  const bool apply_dump = P->dump1 && P->dump2;
  std::vector<double> dump;
  if (apply_dump) {
    dump.resize(n12);
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) dump[i + (size_t)n1 * j] = P->dump1[i] * P->dump2[j];
  }

  /* CRP data (setup_CRP_data, :762-805). */
  const int ne = P->crp.n_energies;
  const bool do_crp = P->crp.calculate_CRP != 0;
  std::vector<Cx> psi_s, dpsi_s, fai, dfai, e_dt, e_t;
  Stencil st;
  if (do_crp) {
    if (!P->crp.energies || !P->crp.eta_sq || !P->crp.CRP) return set_err(QD_ERR_NULL, "missing CRP array");
    if (!make_stencil(P->crp.n_gradient_points, st)) return set_err(QD_ERR_STENCIL, "NPoints error");
    const int s = P->crp.n_dividing_surface - (cuda ? 0 : 1);
    if (s - st.h < 0 || s + st.h >= n2) return set_err(QD_ERR_SURFACE, "surface stencil outside r grid");
    psi_s.resize((size_t)n1 * nth);
    dpsi_s.resize((size_t)n1 * nth);
    fai.assign((size_t)n1 * nth * ne, Cx{0, 0});
    dfai.assign((size_t)n1 * nth * ne, Cx{0, 0});
    e_dt.resize(ne);
    e_t.resize(ne);
    const Cx I{0.0, 1.0};
    for (int i = 0; i < ne; i++) {
      e_dt[i] = cexp_ref(mul(mul(I, P->crp.energies[i]), dt));
      e_t[i] = Cx{dt, 0.0};
    }
  }

  Cx *lpsi_buf = new (std::nothrow) Cx[n12 * m];
  if (!lpsi_buf) return set_err(QD_ERR_NOMEM, "out of memory");
  Cx *lpsi = lpsi_buf;

  const int total_steps = P->time.total_steps;
  int &steps = P->time.steps;
  double e_kin = 0, e_rot = 0, e_pot = 0, module = 0;

  auto apply_dump_fn = [&]() {
    if (!apply_dump) return;
#pragma omp parallel for schedule(static, 1)
    for (int k = 0; k < nth; k++)
      for (size_t x = 0; x < n12; x++) {
        Cx &c = psi[x + n12 * k];
        c.re *= dump[x];
        c.im *= dump[x];
      }
  };
  auto module_fn = [&]() {  /* module_for_psi, :220-243 */
    double s = 0.0;
#pragma omp parallel for schedule(static, 1) reduction(+ : s)
    for (int k = 0; k < nth; k++) {
      double sk = 0.0;
      for (size_t x = 0; x < n12; x++) sk += abs2(psi[x + n12 * k]);
      s += w[k] * sk;
    }
    return s * dr1 * dr2;
  };

  std::vector<double> crp_tmp(do_crp ? ne : 0);
  bool nan_seen = false;

  for (int i_step = 0; i_step < total_steps; i_step++) {
    if (cuda) apply_dump_fn();  /* evolutionCUDA2.cu:503: dump before the step */

    if (i_step == 0 && steps == 0) {
      if (cuda) {
        /* evolution_with_potential(-dt/2): psi *= exp(-i V (-dt/2)) */
        const Cx Ih{0.0, dt / 2};
#pragma omp parallel for schedule(static)
        for (size_t x = 0; x < npt; x++) psi[x] = mul(psi[x], cexp_ref(mul(Ih, pot[x])));
      } else {
        /* pre_evolution_with_potential_dt_2 (:519-538): psi /= sqrt(exp(-iVdt)). */
#pragma omp parallel for schedule(static)
        for (size_t x = 0; x < npt; x++) psi[x] = cdiv_ref(psi[x], sqrt_ref(expV[x]));
      }
    }

    /* evolution_dt (:619-647) */
#pragma omp parallel for schedule(static)
    for (size_t x = 0; x < npt; x++) psi[x] = mul(psi[x], expV[x]);

    qdo_forward_legendre(P->psi, reinterpret_cast<double *>(lpsi), (int)n12, nth, m, wleg.data());

#pragma omp parallel for schedule(static, 1)
    for (int l = 0; l < m; l++)
      for (size_t x = 0; x < n12; x++) lpsi[x + n12 * l] = mul(lpsi[x + n12 * l], expRot[x + n12 * l]);

#pragma omp parallel
    {
      std::vector<Cx> w1, w2;
#pragma omp for schedule(static, 1)
      for (int l = 0; l < m; l++) fft_2d_slab(n1, n2, -1, lpsi + n12 * l, w1, w2);
    }

#pragma omp parallel for schedule(static, 1)
    for (int l = 0; l < m; l++)
      for (size_t x = 0; x < n12; x++) lpsi[x + n12 * l] = mul(lpsi[x + n12 * l], expKin[x]);

    if (opt.diagnostics) {  /* kinetic_energy_for_legendre_psi(0), :390-436 */
      double s = 0.0;
#pragma omp parallel for schedule(static, 1) reduction(+ : s)
      for (int l = 0; l < m; l++) {
        double sl = 0.0;
        for (int j = 0; j < n2; j++)
          for (int i = 0; i < n1; i++)
            sl += abs2(lpsi[i + (size_t)n1 * j + n12 * l]) * (psq1[i] + psq2[j]);
        s += sl / (l + 0.5);
      }
      e_kin = s * dr1 * dr2 * (double)(n1 * n2);
    }

#pragma omp parallel
    {
      std::vector<Cx> w1, w2;
#pragma omp for schedule(static, 1)
      for (int l = 0; l < m; l++) fft_2d_slab(n1, n2, +1, lpsi + n12 * l, w1, w2);
    }
    if (cuda) {  /* cublasZdscal(1/(n1 n2)) after the inverse FFT */
      const double sc = 1.0 / (double)(n1 * n2);
#pragma omp parallel for schedule(static)
      for (size_t x = 0; x < n12 * m; x++) lpsi[x] = mul(lpsi[x], sc);
    }

#pragma omp parallel for schedule(static, 1)
    for (int l = 0; l < m; l++)
      for (size_t x = 0; x < n12; x++) lpsi[x + n12 * l] = mul(lpsi[x + n12 * l], expRot[x + n12 * l]);

    if (opt.diagnostics) {  /* rotational_energy(0), :323-360 */
      double s = 0.0;
#pragma omp parallel for schedule(static, 1) reduction(+ : s)
      for (int l = 0; l < m; l++) {
        double sl = 0.0;
        for (int j = 0; j < n2; j++)
          for (int i = 0; i < n1; i++) sl += abs2(lpsi[i + (size_t)n1 * j + n12 * l]) * (I1[i] + I2[j]);
        s += sl * l * (l + 1) / (l + 0.5);
      }
      e_rot = s * dr1 * dr2;
    }

    qdo_backward_legendre(P->psi, reinterpret_cast<const double *>(lpsi), (int)n12, nth, m, leg);

    if (opt.diagnostics) {  /* potential_energy, :256-283 */
      double s = 0.0;
#pragma omp parallel for schedule(static, 1) reduction(+ : s)
      for (int k = 0; k < nth; k++) {
        double sk = 0.0;
        for (size_t x = 0; x < n12; x++) sk += abs2(psi[x + n12 * k]) * pot[x + n12 * k];
        s += w[k] * sk;
      }
      e_pot = s * dr1 * dr2;
    }

    if (!cuda) apply_dump_fn();  /* dump_psi after the step (:667) */

    if (opt.diagnostics) {
      module = module_fn();
      if (!std::isfinite(module)) nan_seen = true;
    }
    if (opt.verbose)
      printf(" Step: %d e_kin: %.15g e_rot: %.15g e_pot: %.15g module: %.15g\n", i_step, e_kin,
             e_rot, e_pot, module);

    if (cuda) steps++;

    if (do_crp) {
      /* calculate_reaction_probabilities (:870-878) */
      const int s_idx = P->crp.n_dividing_surface;
      qdo_gradient3d(n1, n2, nth, s_idx, cuda ? 0 : 1, dr2, P->psi, reinterpret_cast<double *>(psi_s.data()),
                     reinterpret_cast<double *>(dpsi_s.data()), P->crp.n_gradient_points);
      if (cuda) {
        /* exp(i t E) dt with t = (k+1) dt (evolutionCUDA2.cu:143-145,550) */
        const double t = (i_step + 1) * dt;
        for (int i = 0; i < ne; i++) e_t[i] = mul(cexp_ref(mul(Cx{0.0, t}, P->crp.energies[i])), dt);
      } else {
        for (int i = 0; i < ne; i++) e_t[i] = mul(e_t[i], e_dt[i]);  /* update_exp_ienergy_t */
      }
      qdo_psi_time_to_energy(n1, nth, ne, reinterpret_cast<const double *>(e_t.data()),
                             reinterpret_cast<const double *>(psi_s.data()),
                             reinterpret_cast<const double *>(dpsi_s.data()),
                             reinterpret_cast<double *>(fai.data()), reinterpret_cast<double *>(dfai.data()));
      const bool last = i_step == total_steps - 1;
      const bool every = opt.crp_every > 0 && ((i_step + 1) % opt.crp_every == 0);
      if (last || every) {
        qdo_calculate_crp(n1, nth, ne, dr1, P->r2.mass, w, P->crp.eta_sq,
                          reinterpret_cast<const double *>(fai.data()),
                          reinterpret_cast<const double *>(dfai.data()), P->crp.CRP);
        if (cuda)  /* Zdotc(dfai, fai): Im(conj(dphi) phi) = -CPU value (:972-977) */
          for (int i = 0; i < ne; i++) P->crp.CRP[i] = -P->crp.CRP[i];
      }
    }
    if (!cuda) steps++;
  }

  delete[] lpsi_buf;
  if (diag) {
    diag->module = opt.diagnostics ? module : module_fn();
    diag->e_kin = e_kin;
    diag->e_rot = e_rot;
    diag->e_pot = e_pot;
    diag->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  }
  if (nan_seen) return set_err(QD_ERR_NAN, "non-finite norm");
  return QD_OK;
}

const char *qdo_last_error(void) { return g_err; }

}  // extern "C"
