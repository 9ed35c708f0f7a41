/*
 * qd.h -- C ABI of the MI355X time-dependent wavepacket propagator.
 *
 * The drop-in boundary is the reference's MEX entry
 *   void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
 * (common/TimeEvolutionMex.C:18-78 for the CPU propagator,
 *  common-cuda/TimeEvolutionMexCUDA.C:20-83 for the CUDA one), which takes one
 * MATLAB struct and mutates three of its fields in place: psi, time.steps and
 * CRP.CRP.  The structs below carry exactly the fields that entry reads
 * (common/MatlabStructures.h:12-159), as plain pointers and sizes.  All arrays
 * are owned by the caller; the library never frees them.
 *
 * Array layout is MATLAB's column-major order with R fastest
 * (common/mat.h:18-41):  element (i, j, k) of an n1 x n2 x ntheta array sits
 * at i + n1*(j + n2*k).  psi is complex, stored as interleaved (re, im)
 * doubles, so it holds 2*n1*n2*ntheta doubles.
 */
#ifndef QD_H
#define QD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes (replace the reference's insist()/MCrash aborts) ---- */
enum {
  QD_OK = 0,
  QD_ERR_NULL = 1,        /* a required array is NULL (reference: null dereference) */
  QD_ERR_ODD_N = 2,       /* r1.n or r2.n odd (common/fftwinterface.C:112) */
  QD_ERR_STENCIL = 3,     /* n_gradient_points not an odd number in [3, 61]
                             (common/Gradient3.F:25, common-cuda/gradients2.cu:566-569) */
  QD_ERR_SURFACE = 4,     /* dividing-surface stencil leaves the r grid
                             (common-cuda/evolutionCUDA2.cu:915) */
  QD_ERR_NAN = 5,         /* non-finite norm detected */
  QD_ERR_DEVICE = 6,      /* HIP runtime error */
  QD_ERR_ARG = 7,         /* invalid size or option */
  QD_ERR_NOMEM = 8,       /* device or host allocation failed */
  QD_ERR_NOT_BUILT = 9    /* requested path is not available in this build */
};

/* ---- precision / pipeline / variant selectors ---- */
enum {
  QD_PREC_FP64 = 0,       /* every array and operation in FP64 */
  QD_PREC_MIXED = 1       /* FP32 psi/FFT/phases, BF16 hi+lo theta GEMM with FP32
                             accumulation, FP64 flux accumulation and P(E) */
};
enum {
  QD_PIPE_FUSED = 0,      /* four HBM sweeps per step (theta-fwd, R pass, r pass, theta-bwd) */
  QD_PIPE_PLAIN = 1       /* one kernel per reference operator (timeEvol.C:619-647) */
};
enum {
  QD_VARIANT_CPU = 0,     /* semantics of TimeEvolutionMex (common/timeEvol.C) */
  QD_VARIANT_CUDA = 1     /* semantics of TimeEvolutionMexCUDA (common-cuda/evolutionCUDA2.cu) */
};

/* RadialCoordinate (common/MatlabStructures.h:12-39, .C:17-36). */
typedef struct qd_radial {
  int32_t n;              /* grid points, even */
  const double *r;        /* r[n], uniform grid; FFT period is n*dr */
  double dr;              /* spacing */
  double mass;            /* reduced mass in electron masses */
} qd_radial;

/* AngleCoordinate (common/MatlabStructures.h:41-66, .C:38-50). */
typedef struct qd_angle {
  int32_t n;              /* number of theta DVR points */
  int32_t m;              /* highest Legendre degree; the FBR has m+1 functions */
  const double *x;        /* x[n] = cos(theta_k), ascending */
  const double *w;        /* w[n] Gauss-Legendre weights */
  const double *legendre; /* (m+1) x n column-major: legendre[l + (m+1)*k] = P_l(x_k) */
} qd_angle;

/* EvolutionTime (common/MatlabStructures.h:68-90). */
typedef struct qd_time {
  int32_t total_steps;    /* steps to take in this call */
  int32_t steps;          /* in/out: steps taken so far (0 => apply the initial -dt/2 phase) */
  double time_step;       /* dt in atomic units */
} qd_time;

/* CummulativeReactionProbabilities (common/MatlabStructures.h:131-159). */
typedef struct qd_crp {
  int32_t n_dividing_surface; /* r index of the flux surface (1-based in the CPU
                                 variant, common/Gradient3.F:45; 0-based in the CUDA
                                 variant, common-cuda/gradients2.cu:601) */
  int32_t n_gradient_points;  /* odd stencil size */
  int32_t n_energies;
  int32_t calculate_CRP;      /* 0: propagate only */
  const double *energies;     /* total energies (Eh), n_energies */
  const double *eta_sq;       /* |eta(E)|^2 normalisation, n_energies */
  double *CRP;                /* out: P(E), n_energies */
} qd_crp;

/* The MATLAB struct passed as prhs[0] (common/TimeEvolutionMex.C:30-68). */
typedef struct qd_problem {
  qd_radial r1;           /* R (atom - diatom centre of mass) */
  qd_radial r2;           /* r (diatom bond) */
  qd_angle theta;
  const double *pot;      /* V(R_i, r_j, theta_k), n1*n2*ntheta, Eh */
  double *psi;            /* in/out, 2*n1*n2*ntheta interleaved complex */
  qd_time time;           /* time.steps is in/out */
  const double *dump1;    /* dump1.dump[n1] or NULL (absorption off, timeEvol.h:111) */
  const double *dump2;    /* dump2.dump[n2] or NULL */
  qd_crp crp;
} qd_problem;

/* Options outside the reference struct.  A zeroed struct gives the defaults:
 * FP64, fused pipeline, CPU-variant semantics, P(E) at the end of the call. */
typedef struct qd_options {
  int32_t precision;      /* QD_PREC_* */
  int32_t pipeline;       /* QD_PIPE_* */
  int32_t variant;        /* QD_VARIANT_* */
  int32_t flux_block;     /* steps per time block of the flux accumulation; 0 = 64 */
  int32_t crp_every;      /* >0: evaluate P(E) (and call the callback) every N steps */
  int32_t use_graph;      /* 1: replay captured per-block step graphs */
  int32_t device;         /* HIP device ordinal */
  int32_t profile;        /* 1: time every kernel class with HIP events (qd_get_timing) */
  int32_t theta_split;    /* mixed theta GEMM: bf16 pieces per operand, 2 (hi+lo, three
                             products) or 3 (six products); 0 = 3 */
  int32_t reserved;
} qd_options;

/* Per-step callback, the equivalent of options.wave_to_matlab
 * (common/wave2matlab.C:27-34).  crp is P(E) at this step (n_energies values)
 * when P(E) was evaluated, else NULL.  A nonzero return stops the propagation. */
typedef int (*qd_step_cb)(int32_t steps, const double *crp, void *user);

typedef struct qd_handle qd_handle;

/* One-shot drop-in for TimeEvolutionMex(struct) / TimeEvolutionMexCUDA(struct)
 * (common/TimeEvolutionMex.C:18-78, common-cuda/TimeEvolutionMexCUDA.C:20-83):
 * copies psi and the tables to the GPU, takes time.total_steps steps, writes
 * psi, time.steps and (if crp.calculate_CRP) crp.CRP back.  As in the
 * reference, the flux accumulators start from zero on every call. */
int qd_time_evolution(qd_problem *problem, const qd_options *options);

/* Handle API: the handle owns the device state, including the flux
 * accumulators, so repeated qd_propagate calls continue P(E) across calls
 * (the reference restarts them, timeEvol.C:762-805). */
int qd_create(const qd_problem *problem, const qd_options *options, qd_handle **out);
int qd_propagate(qd_handle *h, int32_t n_steps);
int qd_set_callback(qd_handle *h, qd_step_cb cb, void *user);
int qd_get_psi(qd_handle *h, double *psi_out);          /* 2*n1*n2*ntheta doubles */
int qd_get_crp(qd_handle *h, double *crp_out);          /* n_energies doubles */
int qd_get_norm(qd_handle *h, double *norm_out);        /* sum_k w_k sum |psi|^2 dr1 dr2 */
int qd_get_steps(qd_handle *h, int32_t *steps_out);
/* Average device time of one step over the last qd_propagate call (ms), and of
 * the dominant kernel class selected by `which` (0 = step, 1 = theta GEMMs,
 * 2 = R pass, 3 = r pass, 4 = flux). */
int qd_get_timing(qd_handle *h, int32_t which, double *ms_out);
int qd_destroy(qd_handle *h);

/* Last error message of the calling thread (empty string if none). */
const char *qd_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* QD_H */
