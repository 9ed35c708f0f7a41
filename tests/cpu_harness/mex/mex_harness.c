/* Runs the INTEGRATION.md MEX binding (mexFunction) on a MATLAB-like struct
 * built from stand-in mxArrays, against a stand-in qd_time_evolution that
 * checks the qd_problem it receives and acts like a propagation of
 * total_steps steps (advances time.steps, writes psi and CRP.CRP).  Exit
 * status 0 when every check holds.  Test code only. */
#include <setjmp.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "mex.h"
#include "qd.h"

struct mxArray_tag {
  const char *name;
  void *data;
  int nfields;
  mxArray *fields[12];
};

mxArray *mxGetField(const mxArray *pm, size_t index, const char *f) {
  (void)index;
  if (!pm) return NULL;
  for (int i = 0; i < pm->nfields; i++)
    if (strcmp(pm->fields[i]->name, f) == 0) return pm->fields[i];
  return NULL;
}
double *mxGetPr(const mxArray *a) { return (double *)a->data; }
void *mxGetData(const mxArray *a) { return a->data; }
static jmp_buf err_jmp;
static char err_msg[256];
void mexErrMsgTxt(const char *msg) {
  snprintf(err_msg, sizeof err_msg, "%s", msg);
  longjmp(err_jmp, 1);
}
mxArray *mxCreateDoubleScalar(double v) { (void)v; return NULL; }
int mexCallMATLAB(int a, mxArray *b[], int c, mxArray *d[], const char *e) { (void)a; (void)b; (void)c; (void)d; (void)e; return 0; }

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

/* ---- stand-in of the library ---- */
static int calls, bad;
#define CHECK(c) do { if (!(c)) { fprintf(stderr, "check failed: %s\n", #c); bad++; } } while (0)
static double R[4] = {1, 2, 3, 4}, r[2] = {0.5, 1.0}, x[3] = {-0.5, 0, 0.5}, w[3] = {0.5, 1, 0.5};
static double leg[9], pot[24], psi[48], d1[4] = {1, 1, 1, 0.5}, d2[2] = {1, 0.9};
static double en[2] = {0.1, 0.2}, eta[2] = {1.5, 2.5}, crpv[2];
static int32_t expect_steps;
static int with_dumps;
int qd_time_evolution(qd_problem *p, const qd_options *o) {
  calls++;
  CHECK(o && o->precision == QD_PREC_FP64 && o->variant == QD_VARIANT_CPU);
  CHECK(p->r1.n == 4 && p->r1.r == R && p->r1.dr == 1.0 && p->r1.mass == 1836.0);
  CHECK(p->r2.n == 2 && p->r2.r == r && p->r2.dr == 0.5 && p->r2.mass == 918.0);
  CHECK(p->theta.n == 3 && p->theta.m == 2 && p->theta.x == x && p->theta.w == w && p->theta.legendre == leg);
  CHECK(p->pot == pot && p->psi == psi);
  CHECK(p->time.total_steps == 5 && p->time.time_step == 10.0);
  CHECK(p->time.steps == expect_steps);
  CHECK(with_dumps ? (p->dump1 == d1 && p->dump2 == d2) : (!p->dump1 && !p->dump2));
  CHECK(p->crp.n_dividing_surface == 1 && p->crp.n_gradient_points == 3 && p->crp.n_energies == 2);
  CHECK(p->crp.calculate_CRP == 1 && p->crp.energies == en && p->crp.eta_sq == eta && p->crp.CRP == crpv);
  if (p->time.steps < 0) return QD_ERR_ARG;
  p->psi[0] = 42.0;
  p->crp.CRP[1] = 0.25;
  p->time.steps += p->time.total_steps;
  return QD_OK;
}
const char *qd_last_error(void) { return "stand-in error"; }

static mxArray *field(const char *name, void *data) {
  static mxArray pool[64];
  static int used;
  mxArray *a = &pool[used++];
  a->name = name;
  a->data = data;
  a->nfields = 0;
  return a;
}
static void add(mxArray *s, mxArray *f) { s->fields[s->nfields++] = f; }

int main(void) {
  int32_t n1 = 4, n2 = 2, nt = 3, m = 2, total = 5, steps = 0, nds = 1, ngp = 3, ne = 2, calc = 1;
  double dr1 = 1.0, dr2 = 0.5, m1 = 1836.0, m2 = 918.0, dt = 10.0;
  mxArray *s = field("", NULL);
  mxArray *f;
  f = field("r1", NULL); add(f, field("n", &n1)); add(f, field("r", R)); add(f, field("dr", &dr1));
  add(f, field("mass", &m1)); add(s, f);
  f = field("r2", NULL); add(f, field("n", &n2)); add(f, field("r", r)); add(f, field("dr", &dr2));
  add(f, field("mass", &m2)); add(s, f);
  f = field("theta", NULL); add(f, field("n", &nt)); add(f, field("m", &m)); add(f, field("x", x));
  add(f, field("w", w)); add(f, field("legendre", leg)); add(s, f);
  add(s, field("pot", pot));
  add(s, field("psi", psi));
  f = field("time", NULL); add(f, field("total_steps", &total)); add(f, field("time_step", &dt));
  add(f, field("steps", &steps)); add(s, f);
  f = field("CRP", NULL); add(f, field("n_dividing_surface", &nds)); add(f, field("n_gradient_points", &ngp));
  add(f, field("n_energies", &ne)); add(f, field("calculate_CRP", &calc)); add(f, field("energies", en));
  add(f, field("eta_sq", eta)); add(f, field("CRP", crpv)); add(s, f);
  const mxArray *prhs[1] = {s};

  /* first call: no dumps, steps 0 -> 5, psi and CRP.CRP written in place */
  expect_steps = 0;
  with_dumps = 0;
  if (setjmp(err_jmp)) { fprintf(stderr, "unexpected error: %s\n", err_msg); return 1; }
  mexFunction(0, NULL, 1, prhs);
  CHECK(calls == 1 && steps == 5 && psi[0] == 42.0 && crpv[1] == 0.25);
  /* second call resumes: the struct's steps (5) reaches the library, and comes back as 10 */
  f = field("dump1", NULL); add(f, field("dump", d1)); add(s, f);
  f = field("dump2", NULL); add(f, field("dump", d2)); add(s, f);
  expect_steps = 5;
  with_dumps = 1;
  mexFunction(0, NULL, 1, prhs);
  CHECK(calls == 2 && steps == 10);
  /* an error from the library reaches mexErrMsgTxt with its message */
  steps = -1;
  expect_steps = -1;
  if (setjmp(err_jmp) == 0) {
    mexFunction(0, NULL, 1, prhs);
    CHECK(0 && "no error raised");
  } else {
    CHECK(strcmp(err_msg, "stand-in error") == 0);
  }
  /* wrong argument count */
  if (setjmp(err_jmp) == 0) {
    mexFunction(0, NULL, 0, prhs);
    CHECK(0 && "nrhs not checked");
  }
  if (bad) return 1;
  printf("mex binding ok\n");
  return 0;
}
