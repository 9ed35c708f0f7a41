/*
 * qd_oracle.h -- CPU oracle of the wavepacket propagator.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (libqd_hip.so) never links or calls it.
 *
 * Each function restates one reference routine; the citation is next to it.
 */
#ifndef QD_ORACLE_H
#define QD_ORACLE_H

#include "../include/qd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qdo_options {
  int32_t variant;        /* QD_VARIANT_CPU or QD_VARIANT_CUDA */
  int32_t diagnostics;    /* 1: evaluate norm and energies every step, as the
                             reference prints them (timeEvol.C:631-646,671) */
  int32_t crp_every;      /* CPU variant: P(E) every step when 1 (timeEvol.C:673);
                             0: only after the last step (same final value) */
  int32_t nthreads;       /* OpenMP threads, 0 = runtime default */
  int32_t verbose;        /* print the per-step log lines */
  int32_t reserved[3];
} qdo_options;

/* Last diagnostics of qdo_time_evolution (valid when diagnostics == 1). */
typedef struct qdo_diag {
  double module;          /* norm after the step */
  double e_kin, e_rot, e_pot;
  double seconds;         /* wall time of the whole call */
} qdo_diag;

/* Whole propagation, restating TimeEvolution::time_evolution
 * (common/timeEvol.C:649-685) or EvolutionCUDA::time_evolution
 * (common-cuda/evolutionCUDA2.cu:487-552) when variant == QD_VARIANT_CUDA. */
int qdo_time_evolution(qd_problem *p, const qdo_options *opt, qdo_diag *diag);

/* Unnormalised complex DFT of length n, sign -1 (FFTW_FORWARD) or +1, in place
 * on n interleaved complex values with element stride `stride` (complex units). */
void qdo_fft_1d(int n, int sign, double *data, int stride);
/* 2-D DFT of one n1 x n2 column-major slab (fftw_plan_dft_2d, fftwinterface.C:71-76). */
void qdo_fft_2d(int n1, int n2, int sign, double *data);

/* FFTWInterface::get_momentum_for_fftw (common/fftwinterface.C:110-125). */
int qdo_momentum(int n, double xl, double *p);

/* LegTransform.F:6-31 and :34-59 (complex x real, K ascending). */
void qdo_forward_legendre(const double *cpsi, double *lpsi, int nr, int ntheta,
                          int nleg, const double *wlegp);
void qdo_backward_legendre(double *cpsi, const double *lpsi, int nr, int ntheta,
                           int nleg, const double *legp);

/* Gradient3.F:9-167 generalised to odd stencils 3..61 (gradients2.cu:37-610).
 * n2p is the surface index in the convention `base` (1 = Fortran, 0 = C). */
int qdo_gradient3d(int n1, int n2, int n3, int n2p, int base, double dx,
                   const double *f, double *v, double *g, int npoints);

/* Central first-derivative coefficients c_1..c_h of a (2h+1)-point stencil
 * (gradients2.cu:18-35 closed form). */
int qdo_gradient_coefficients(int npoints, double *c);

/* CRPAux.F:4-31. */
void qdo_psi_time_to_energy(int n1, int ntheta, int ne, const double *exp_iet_dt,
                            const double *psit, const double *dpsit,
                            double *fai, double *dfai);
/* CRPAux.F:34-66. */
void qdo_calculate_crp(int n1, int ntheta, int ne, double dr1, double mu2,
                       const double *w, const double *eta_sq, const double *fai,
                       const double *dfai, double *crp);

#ifdef __cplusplus
}
#endif

#endif
