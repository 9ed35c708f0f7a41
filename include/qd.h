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
  int32_t use_graph;      /* 1: run whole flux blocks (flux_block steps and their
                             accumulation) as one replayed hipGraph */
  int32_t device;         /* HIP device ordinal */
  int32_t profile;        /* 1: time every kernel class with HIP events (qd_get_timing) */
  int32_t theta_split;    /* mixed theta GEMM: bf16 pieces per operand, 2 (hi+lo, three
                             products) or 3 (six products); 0 = 3 */
  int32_t autotune;       /* 1: time the candidate kernel configurations on the
                             handle's own arrays at creation and keep the fastest */
  int32_t fuse_theta;     /* mixed precision, angular basis of at most 128: 0 = fuse the
                             theta-backward GEMM of each step with the dump, the flux
                             rows, the V phase and the next theta-forward GEMM (the
                             state between steps is then the FBR array; psi is rebuilt
                             from it when read, or, in the CUDA propagator's order,
                             whose FBR state holds the dumped psi, written by the last
                             step before it is read); -1 = never */
  int32_t diagnostics_every; /* >0: every N-th step also evaluates the reference's
                             per-step diagnostics -- module, e_kin, e_rot, e_pot
                             (timeEvol.C:619-647,671) -- at the same points of the
                             step, on the device (qd_get_diagnostics).  A single
                             J = 0 grid only (qd_create); such steps run one kernel
                             per operator.  0 = never */
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
/* Diagnostics of the last step that evaluated them (options.diagnostics_every):
 * out[0] = the step count after it, out[1] = module (sum_k w_k sum |psi|^2
 * dr1 dr2, timeEvol.C:220-243, after the dump), out[2] = e_kin (:390-436, in
 * momentum space after the kinetic phase), out[3] = e_rot (:323-360, after the
 * second rotational half phase), out[4] = e_pot (:256-283, after the backward
 * theta transform, before the dump).  QD_ERR_ARG if none has been evaluated. */
int qd_get_diagnostics(qd_handle *h, double *out);
/* Average device time of one step over the last qd_propagate call (ms), and,
 * with options.profile = 1, of one launch of the kernel class selected by
 * `which` (0 = step, 1 = theta GEMMs, 2 = R pass, 3 = r pass, 4 = flux-surface
 * extraction, 5 = other, 6 = one flux accumulation, i.e. the fold of the
 * time block into phi(E), dphi(E); flushes made by qd_get_crp after the call
 * are included). */
int qd_get_timing(qd_handle *h, int32_t which, double *ms_out);
int qd_destroy(qd_handle *h);

/* Replace the handle's psi (same layout as qd_get_psi); time.steps and the
 * flux accumulators are kept. */
int qd_set_psi(qd_handle *h, const double *psi_in);

/* Checkpoint of the whole propagation state -- the wavefunction as the handle
 * holds it (psi, or the FBR array of the fused theta step, with psi beside it
 * in the CUDA propagator's order), the energy-resolved
 * flux accumulators phi(E) and dphi(E) with the surface slices of the open time
 * block, the step counts and the pending first half step -- into a caller
 * buffer of qd_checkpoint_size bytes.  Loading it into a handle created from
 * the same problem and options continues the propagation bit for bit as if it
 * had not stopped (the reference keeps these only in memory, timeEvol.C:762-805,
 * and restarts them on every call). */
int qd_checkpoint_size(qd_handle *h, int64_t *bytes);
int qd_checkpoint_save(qd_handle *h, void *buf, int64_t bytes);
int qd_checkpoint_load(qd_handle *h, const void *buf, int64_t bytes);

/* ---- J > 0: body-fixed (J, K) blocks ----------------------------------
 * No counterpart in the reference, which propagates J = 0 only (the rotational
 * term of common/timeEvol.C:487 is l(l+1)(1/2mu1R^2 + 1/2mu2r^2)).  Block
 * (J, K) holds n_K = jmax - K + 1 theta points: the DVR of the normalised
 * associated Legendre functions P_j^K, j = K..jmax (Gauss-Legendre for K = 0).
 * The K blocks of one J are Coriolis coupled through (J - j)^2 in the
 * parity-adapted body-fixed basis.  A J = 0 block uses the reference's own
 * Legendre matrices and reproduces qd_create on the same grid. */
typedef struct qd_jk_block {
  int32_t J, K;           /* 0 <= K <= min(J, jmax); every J needs K = 0..min(J, jmax) */
  const double *pot;      /* V on the block's grid, n1*n2*n_K (blocks of one K share it) */
  const double *psi;      /* initial psi, 2*n1*n2*n_K interleaved */
  const double *eta_sq;   /* |eta_J(E)|^2 normalising this block's flux, n_energies
                             values (the packet's projection onto the incoming wave
                             of partial wave J, qd_init_packet), or NULL for
                             base.crp.eta_sq */
} qd_jk_block;

typedef struct qd_jk_problem {
  qd_problem base;        /* r1, r2, time, dump1, dump2, crp; theta, pot and psi unused */
  int32_t jmax;           /* highest j of the rotational basis */
  int32_t coriolis;       /* 1: Coriolis-coupled; 0: centrifugal-sudden (K blocks uncoupled) */
  int32_t n_blocks;
  int32_t reserved;
  const qd_jk_block *blocks;
} qd_jk_problem;

/* theta points x (ascending cos theta) and weights w of block K (n_K values each). */
int qd_jk_theta_grid(int32_t K, int32_t jmax, double *x, double *w);
/* Build a handle over all blocks; qd_propagate advances them together,
 * qd_get_psi returns the blocks back to back in the given order, qd_get_crp the
 * sum over blocks, qd_get_crp_blocks P(E) per block (n_blocks x n_energies). */
int qd_create_jk(const qd_jk_problem *problem, const qd_options *options, qd_handle **out);
int qd_get_crp_blocks(qd_handle *h, double *crp_out);

/* ---- test hooks: one stage of the step on the handle's current arrays ----
 * stage 0: theta forward (psi -> FBR array, no V phase); 1: R pass with phases,
 * mixing and kinetic factor; 2: r pass with phases and kinetic factor;
 * 3: theta backward (FBR array -> psi, no dump); 4: forward FFT along R only;
 * 5: forward FFT along r only; 6: inverse FFT along R only; 7: inverse FFT
 * along r only.  The FBR array has n1*n2*(FBR columns) complex values.  With
 * the split mixing R pass selected, stage 1 uses psi as its scratch array. */
int qd_debug_stage(qd_handle *h, int32_t stage);
/* kernel configuration in use: [0] autotuned, [1] theta k-steps of the
 * matrix-stationary kernel (0: tiled kernel), [2] its 16-row tiles per work
 * group, [3] register-FFT R pass, [4] register-FFT r pass, [5] R-pass lines
 * per work group, [6] lines per mixing bundle, [7] r-pass lines per group,
 * [8] lines per bundle of the prefetching FP32 mixing R pass (0: not used),
 * [9] its unit-size bucket, [10] fused theta step in use (1: the dense
 * transforms, 2: their parity halves, 3: the parity halves with each tile
 * shared by a pair of waves, one per component), [11] its 16-row
 * tiles per work group, [12] that mixing R pass runs two work groups per CU,
 * [13] threads per work group of the fused theta step, [14] the split mixing
 * R pass (K -> l mixing, the register FFT pass without mixing, l -> K mixing,
 * three launches) is in use.  vals holds 15 entries. */
int qd_get_tuning(qd_handle *h, int32_t *vals);
/* test hook: force one kernel choice the autotuner makes.  what 0: the
 * prefetching FP32 mixing R pass with one (value 0) or two (value 1) work
 * groups per CU; QD_ERR_ARG when the handle does not have that kernel.
 * what 1: its consecutive bundles per turn of a work group (1..64, default 8).
 * what 2: the split mixing R pass on (value 1) or off (value 0); QD_ERR_ARG
 * when the handle does not have it, or has no other R pass for its groups.
 * what 3: the fused theta step on the dense transforms (value 0; bases of at
 * most 64) or on their parity halves (1: 512 threads per work group, 256 for
 * bases of 65..128; 2: 768 threads, bases of at most 64; 3: 512 threads as
 * wave pairs, bases of 65..128); QD_ERR_ARG when the handle does not have that
 * kernel.
 * what 4: threads per work group of the register-FFT r pass, 256 or 512
 * (batches of twice the consecutive R points); QD_ERR_ARG when it does not fit.
 * what 5: the fused step's flux-surface extraction run by the next step's
 * register-FFT r pass (value 1, the default) or as a launch of its own after
 * each step (0).
 * what 6: the wave-pair fused theta kernel's two waves of a tile meet through
 * LDS counters (value 1, the default) or at work-group barriers (0); the same
 * bits either way.  QD_ERR_ARG on a handle without the wave-pair kernel. */
int qd_debug_select(qd_handle *h, int32_t what, int32_t value);
int qd_debug_get_fbr(qd_handle *h, double *out);
int qd_debug_set_fbr(qd_handle *h, const double *in);

/* ---- set-up -------------------------------------------------------------------
 * What the reference's MATLAB drivers compute before they call the propagator,
 * restated in C++ so that a caller can take packet parameters instead of a
 * precomputed psi.  Host code only: these need no GPU. */

/* r[n] = linspace(lo, hi, n) and *dr = r[1] - r[0] (the drivers' grids, H3/main.m:45-60). */
int qd_make_radial(int32_t n, double lo, double hi, double *r, double *dr);
/* theta.x and theta.w (n Gauss-Legendre points, ascending; common/GaussLegendre.F:4-38,
 * GaussLegendre2.m:4-32) and theta.legendre ((m+1) x n, legendre[l + (m+1)*k] = P_l(x_k);
 * LegendreP2.m:4-23 transposed as H3/main.m:101-104 does).  legendre may be NULL. */
int qd_make_theta(int32_t n, int32_t m, double *x, double *w, double *legendre);
/* n_dividing_surface = int32((rd - rmin) / dr), rounding half away from zero as
 * MATLAB's int32() does (H3/main.m:77). */
int qd_dividing_surface(double rd, double rmin, double dr, int32_t *index);

/* The initial state of InitWavePacket.m (H3/InitWavePacket.m:2-31). */
typedef struct qd_packet {
  double R0, delta, k0;   /* Gaussian in R: centre, width, mean momentum (r1.r0, r1.delta, r1.k0) */
  int32_t v0, j0;         /* vibrational level and rotation of the diatomic */
  int32_t J, K;           /* total angular momentum and its body-fixed projection; J = 0, K = 0
                             is the reference's problem */
  const double *vdiat;    /* diatomic potential on r2.r, n2 values (Eh; H3/H2PES.m) */
} qd_packet;

/* psi (2*n1*n2*ntheta interleaved) = g(R_i) phi(r_j) Theta(x_k) with
 *   g(R) = (pi delta^2)^(-1/4) exp(-(R - R0)^2 / (2 delta^2) - i k0 R),
 *   phi  = eigenstate v0 of the diatomic in rotation j0 on the r grid, by the
 *          radial sine DVR of H3/H2VibRotWaveFunction.m:19-46 (indexed from 1;
 *          normalised to sum phi^2 dr = 1, its largest value positive),
 *   Theta = sqrt(j0 + 1/2) P_j0(x_k) for J = 0 on theta.x (the reference's
 *          Gauss-Legendre points), and the normalised associated Legendre
 *          Pt_j0^K(x_k) for J > 0 on the block's points (qd_jk_theta_grid);
 * *e_diatomic = e_v (Eh), and eta_sq[e] = |eta(energies[e] - e_v)|^2, the flux
 * normalisation: common/EtaSq.m:4-18 for J = 0, and for J > 0 the packet's
 * projection onto the energy-normalised incoming wave of orbital angular
 * momentum l = J (which reduces to EtaSq.m at l = 0; needs j0 = 0, where l = J).
 * eta_sq may be NULL when n_energies = 0; psi may be NULL to get only e_v and eta. */
int qd_init_packet(const qd_radial *r1, const qd_radial *r2, const qd_angle *theta, const qd_packet *packet,
                   int32_t n_energies, const double *energies, double *psi, double *e_diatomic, double *eta_sq);

/* Last error message of the calling thread (empty string if none). */
const char *qd_last_error(void);

/* ---- learned J interpolation ------------------------------------------------
 * No counterpart in the reference, which sums P^J(E) over propagated J only.
 * A 3 x 128 tanh MLP in FP32 (inputs n_in <= 4, e.g. J, E and the reduced
 * centrifugal barrier; one output, P^J(E)), trained on the GPU by hand-written
 * forward, backward and Adam kernels on P^J(E) of the propagated J, used to
 * predict P^J(E) of J that are skipped.  Parameters are one flat FP32 vector:
 * W1[128][n_in], b1, W2[128][128], b2, W3[128][128], b3, w4[128], b4. */
typedef struct qd_mlp qd_mlp;
int qd_mlp_n_params(int32_t n_in, int64_t *n_out);
int qd_mlp_create(int32_t n_in, const float *params, int32_t device, qd_mlp **out);
/* training set: X[n][n_in] features, T[n] targets */
int qd_mlp_set_data(qd_mlp *h, const float *X, const float *T, int32_t n);
/* mean squared error and its gradient at the current parameters (no update) */
int qd_mlp_loss_grad(qd_mlp *h, float *loss, float *grad);
/* iters full-batch Adam steps (b1 0.9, b2 0.999, eps 1e-8); use_graph = 1
 * replays one captured iteration as a hipGraph; last_loss may be NULL */
int qd_mlp_train(qd_mlp *h, int32_t iters, float lr, int32_t use_graph, float *last_loss);
int qd_mlp_predict(qd_mlp *h, const float *X, int32_t n, float *Y);
/* qd_mlp_predict that allocates nothing (no counterpart in the reference): for
 * callers that cannot allocate device memory, e.g. a tight serving loop that
 * predicts P^J(E) of skipped J.  X and Y are host arrays as above; they are
 * staged through a fixed device array of QD_MLP_NOALLOC_MAX samples (1.25 MiB,
 * one per GPU, held until the process ends) instead of per-call buffers.
 * n > QD_MLP_NOALLOC_MAX returns QD_ERR_ARG.  Calls that overlap on one GPU
 * share that array and race. */
#define QD_MLP_NOALLOC_MAX 65536
int qd_mlp_predict_noalloc(qd_mlp *h, const float *X, int32_t n, float *Y);
int qd_mlp_get_params(qd_mlp *h, float *params);
int qd_mlp_destroy(qd_mlp *h);
const char *qd_mlp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* QD_H */
