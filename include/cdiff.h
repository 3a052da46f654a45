/*
 * cdiff.h -- C ABI of libcdiff, the MI355X (gfx950) implementation of
 * CurvilinearDiffusion.jl's accelerated pseudo-transient (APT) solver.
 *
 * The reference's plug-in point is Julia dispatch on the backend type of
 * PseudoTransientSolver{N,T,BE} (src/pseudo_transient/PseudoTransientSolver.jl:28,44,
 * backend chosen by `backend=` at :49).  A backend marker type whose methods
 * `ccall` these entry points replaces, one for one:
 *
 *   cdiff_create / cdiff_set_metrics   PseudoTransientSolver(mesh, bcs; backend, mean, T)
 *                                      (PseudoTransientSolver.jl:48-103, phys_dims :122-153,
 *                                      get_metric_cache :105-120)
 *   cdiff_set_density, cdiff_set_source  the rho argument and solver.source_term
 *                                      (PseudoTransientSolver.jl:31,184; drivers copy! into
 *                                      source_term, e.g. test/integrated/2d/gaussian/run.jl:166)
 *   cdiff_update_conductivity          update_conductivity!(solver, mesh, T, rho, cp, kappa)
 *                                      (src/conductivity.jl:15-30,76-84)
 *   cdiff_step                         nonlinear_thermal_conduction_step!(solver, mesh, T, rho,
 *                                      cp, kappa, dt; kwargs...) -> (stats, next_dt)
 *                                      (src/nonlinear_thermal_conduction.jl:51-72) which runs
 *                                      PseudoTransientScheme.step! (PseudoTransientSolver.jl:180-351)
 *   cdiff_apply_bcs                    applybcs!(bcs, mesh, u, nhalo)
 *                                      (src/boundary_conditions/boundary_operators.jl:26-30)
 *   cdiff_next_dt                      next_dt(u_new, u_prev, dt; tol, maximum_u_change,
 *                                      timestep_growth_factor) (src/timestep_control.jl:33-54)
 *   cdiff_get_field                    reads of solver.u, .u_prev, .res, .q, .q', .alpha,
 *                                      .dtau_rho, .theta_r_dtau, .source_term, .cache
 *                                      (PseudoTransientSolver.jl:28-46, to_vtk :380-412)
 *
 * Conventions.
 *  - Every array is FP64, column-major (i fastest) and full-size: the mesh
 *    halo is included, exactly as the Julia arrays.  Host arrays are borrowed
 *    for the duration of a call.
 *  - Device memory, the device-resident temperature and the persistent solver
 *    state (fluxes q, the iteration count of the previous step, alpha) belong
 *    to the handle.  A handle is bound to the device current at creation and
 *    is not re-entrant; calls on one handle must not overlap.
 *  - Every call returns a cdiff_status; cdiff_last_error(h) gives the message,
 *    worded as the reference's error() text where one exists.
 */
#ifndef CDIFF_H
#define CDIFF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cdiff_solver cdiff_solver;

typedef enum {
  CDIFF_OK = 0,
  CDIFF_NONFINITE = 1,      /* "Non-finite error detected! ..."   PseudoTransientSolver.jl:302-307 */
  CDIFF_MAX_ITER = 2,       /* "Maximum iteration limit reached"  PseudoTransientSolver.jl:309-314 */
  CDIFF_INVALID_ALPHA = 3,  /* validate_scalar(alpha, ...)        validity_checks.jl:1-42 */
  CDIFF_INVALID_SOURCE = 4, /* validate_scalar(source_term, ...)  validity_checks.jl:1-42 */
  CDIFF_BAD_ARG = 5,        /* argument / @assert failures         PseudoTransientSolver.jl:225-227 */
  CDIFF_HIP_ERROR = 6,      /* a HIP runtime call failed */
  CDIFF_NOT_READY = 7       /* metrics not set */
} cdiff_status;

/* Boundary condition kinds (boundary_operators.jl:11-17). */
typedef enum {
  CDIFF_BC_NEUMANN = 0,
  CDIFF_BC_DIRICHLET = 1,
  CDIFF_BC_PERIODIC = 2,
  CDIFF_BC_FIXED_NEG_SYMMETRY = 3
} cdiff_bc_kind;

typedef struct {
  int32_t kind;  /* cdiff_bc_kind */
  int32_t pad;
  double value;  /* DirichletBC value (bc.val.u in 2D, bc.val in 3D) */
} cdiff_bc;

/* Face means (src/averaging.jl:7-8); `mean=` of the constructor. */
typedef enum { CDIFF_MEAN_HARMONIC = 0, CDIFF_MEAN_ARITHMETIC = 1 } cdiff_mean;

/* Conductivity law kappa(rho, T) = kappa0 * |rho|^rho_exp * |T|^t_exp, and 0
 * when T is not finite.  Covers every closure the reference drivers pass
 * (constant, 2.5|T|^3, T^3, T^(5/2)); Julia closures cannot cross ccall.
 * Integer exponents 0..3 are evaluated by repeated multiplication (Julia
 * literal powers), others with pow(). */
typedef struct {
  double kappa0;
  double t_exp;
  double rho_exp;
} cdiff_kappa;

/* Precision of the pseudo-iterations. */
typedef enum {
  CDIFF_PREC_FP64 = 0,   /* the reference's arithmetic, FP64 throughout */
  CDIFF_PREC_FP32_DC = 1 /* FP32 APT sweeps inside an FP64 defect-correction loop */
} cdiff_precision;

/* step! keyword arguments (PseudoTransientSolver.jl:188-202) and next_dt
 * keywords (timestep_control.jl:33-42).  cdiff_default_opts fills the
 * reference defaults. */
typedef struct {
  int32_t max_iter;              /* 1500 */
  int32_t norm_kind;             /* 0: sqrt(sum r^2 / N) (CPU L2_norm, residuals.jl:8-22);
                                    1: norm(r, Inf)/sqrt(N) (GPU L2_norm, residuals.jl:1-6) */
  double rtol;                   /* 1e-5 */
  double atol;                   /* sqrt(eps(Float64)) */
  double cfl;                    /* 1/sqrt(3) */
  int32_t apply_cutoff;          /* false */
  int32_t calculate_next_dt;     /* true */
  int32_t subcycle_conductivity; /* true */
  int32_t enforce_positivity;    /* false */
  double tol;                    /* 1e-3 */
  double maximum_u_change;       /* 0.05 */
  double timestep_growth_factor; /* 1.1 */
  int32_t precision;             /* cdiff_precision, CDIFF_PREC_FP64 */
  int32_t inner_iters;           /* FP32_DC: FP32 sweeps per FP64 correction (0 = auto) */
} cdiff_opts;

/* stats = (rel_err, abs_err, niter, time) of step! (PseudoTransientSolver.jl:349). */
typedef struct {
  double rel_err;
  double abs_err;
  int32_t niter;   /* pseudo-iterations of this step */
  int32_t nchecks; /* residual evaluations */
  double time;     /* seconds, device-synchronised wall time of the step */
  int32_t outer;   /* FP32_DC: FP64 defect-correction rounds */
  int32_t warm_start; /* 1 if the step started from the learned prediction */
  int32_t fp64_fallback; /* FP32_DC: 1 if the defect correction gave up (diverged or did not converge)
                          * and the step was redone by the FP64 APT from T^n; the other fields are
                          * then the FP64 APT's */
  int32_t warm_consulted; /* 1 if the learned warm start was consulted at this step (its prediction
                           * used or not); 0 when none is attached or it sat the step out */
} cdiff_stats;

/* Fields for cdiff_get_field / cdiff_set_field. */
typedef enum {
  CDIFF_FIELD_U = 0,       /* solver.u (device temperature) */
  CDIFF_FIELD_U_PREV = 1,  /* solver.u_prev */
  CDIFF_FIELD_RES = 2,     /* solver.res of the last check */
  CDIFF_FIELD_ALPHA = 3,   /* solver.alpha */
  CDIFF_FIELD_DTAU_RHO = 4,
  CDIFF_FIELD_THETA_R_DTAU = 5,
  CDIFF_FIELD_SOURCE = 6,
  CDIFF_FIELD_Q = 8,       /* + axis: solver.q.{x,y,z} (damped flux) */
  CDIFF_FIELD_QPRIME = 12, /* + axis: solver.q'.{x,y,z} (plain flux) */
  CDIFF_FIELD_CACHE = 16,  /* + axis: solver.cache.{alpha,beta,gamma} */
  CDIFF_FIELD_DENSITY = 20
} cdiff_field;

void cdiff_default_opts(cdiff_opts *opts);

/* Kernel variants and the parameters of the FP32 defect correction.  No
 * counterpart in the reference, whose only configuration surface is step!'s
 * keywords (PseudoTransientSolver.jl:188-202): cdiff_default_tuning gives the
 * values measured best on MI355X.  The FP64 variants give bit-identical
 * results, and so do the correction sweep's layouts and depths (the tests
 * compare them); the defect-correction parameters change only the path to
 * the same FP64 tolerance.  0 selects the automatic choice where noted. */
typedef struct {
  /* FP64 APT */
  int32_t sweep2d;          /* 0: LDS tiles (default), 1: row-marching strips */
  int32_t tile_rows;        /* rows per 2D tile: 4, 8 (default) or 16 */
  int32_t march_rows64;     /* rows per chunk of the row-marching strips, 0 = auto */
  int32_t pairs;            /* 1: two iterations per 2D launch where possible (default), 0: one */
  int32_t sweep3d;          /* 0: 64 x 8 columns through LDS (default), 1: one wavefront per row */
  int32_t planes3d;         /* planes per 3D chunk, 0 = auto */
  int32_t seg_min, seg_max; /* launches between host reads of the convergence flag: 4, 256 */
  /* FP32 defect correction */
  int32_t dc_coefficient_form; /* 1: coefficient-form sweeps where no face is periodic (default), 0: conductivity form */
  int32_t dc_newton;        /* 1: d(alpha)/du terms in the correction operator (default) */
  int32_t dc_fused3d;       /* 1: fused 3D outer pass (default), 0: separate kernels */
  int32_t dc_depth2d;       /* iterations per 2D correction launch: 2, 4, 6 or 8 (default) */
  int32_t dc_layout2d;      /* 0: 64-column tiles, 1: wave-aligned tiles, 2: marching strips (default) */
  int32_t dc_march_rows;    /* rows per marching chunk, 0 = auto (the height that minimises rounds of resident
                               wavefronts x (rows + 2 depth); 4096^2: 179, all wavefronts in one round) */
  int32_t dc_launch_iters;  /* iterations per launch, 1 .. depth; 0 = the depth */
  int32_t dc_depth3d;       /* iterations per 3D correction launch: 1 or 2 (default) */
  int32_t dc_planes3d;      /* planes per chunk of the one-iteration 3D sweep, 0 = auto */
  int32_t dc_planes3d_pair; /* planes per chunk of the two-iteration 3D sweep, 0 = auto */
  int32_t dc_outer_planes3d; /* planes per chunk of the fused 3D outer pass, 0 = auto (about one block per CU) */
  int32_t dc_check;         /* correction sweeps between checks, 0 = auto (2D: the depth times max(1, 12 / depth); 3D: 2) */
  int32_t dc_max_outer;     /* correction rounds before the step is redone by the FP64 APT: 100 */
  int32_t verbose;          /* 1: per-round diagnostics of the defect correction on stderr */
  int32_t exact_pow;        /* 1: pow() for T^(n+1/2) instead of T^n sqrt(T) (within 2 ulp) */
  int32_t graphs;           /* 1: 2D FP64 blocks of 16 pseudo-iterations without a check replayed
                               from captured HIP graphs; 0: launched one by one (default: the
                               same speed or faster where measured, 128^2 to 4096^2) */
  int32_t warm_policy;      /* learned warm start: 0 consults the network where the saving expected
                               from it exceeds a consultation's cost, and sits out after
                               consultations that did not pay (default); 1 consults it at every
                               step it applies to (evaluation, tests) */
  double dc_eta;            /* inner solves stop at eta x the outer residual; 0 = auto (1e-3 2D, 1e-2 3D) */
  double dc_floor;          /* ... or at floor x atol: 0.3 */
  double dc_cfl_scale;      /* pseudo-time CFL of the correction sweeps over the step's: 1.2 */
  double dc_dtau_scale;     /* factor on their dtau: 1.3 */
  double dc_theta_scale;    /* factor on their theta: 1.0 */
} cdiff_tuning;

void cdiff_default_tuning(cdiff_tuning *t);
/* Replace the solver's tuning (any time between steps; CDIFF_BAD_ARG for a
 * value out of range, the tuning then unchanged). */
int cdiff_set_tuning(cdiff_solver *h, const cdiff_tuning *t);
int cdiff_get_tuning(const cdiff_solver *h, cdiff_tuning *t);

/* Create a solver for a mesh of `ndim` (2 or 3) with full (halo-inclusive)
 * dimensions dims_full[0..ndim-1] and mesh halo `nhalo` (>= 1).  bcs[2*ndim] in
 * (ilo, ihi, jlo, jhi, klo, khi) order -- the NamedTuple position of the
 * reference (boundary_operators.jl:19-28).  Density defaults to 1, the source
 * to 0. */
int cdiff_create(int ndim, const int64_t *dims_full, int nhalo, int mean, const cdiff_bc *bcs,
                 const cdiff_kappa *kappa, double cp, cdiff_solver **out);

/* Mesh metrics (all full-size FP64 arrays):
 *   cell[a*ndim + c]             = d(xi_a)/d(x_c) at cell centres (xi.x1, xi.x2, eta.x1, ...)
 *   edge[b*(ndim*ndim+1) + a*ndim + c] = conservative metric xi_hat_a . e_c on the
 *                                  b+1/2 faces; edge[b*(ndim*ndim+1) + ndim*ndim] = J there
 *   L, spacing[ndim]             = phys_dims (PseudoTransientSolver.jl:122-153)
 * The metric cache alpha_c/beta_c/gamma_c (mesh_metric_cache.jl:39-250) and
 * the divergence coefficients are computed here once, in FP64. */
int cdiff_set_metrics(cdiff_solver *h, const double *const *cell, const double *const *edge, double L,
                      const double *spacing);

/* The same metrics built on the device from the node coordinates: nodes[c]
 * (c < ndim) is the ghost-inclusive node grid of coordinate c, shape
 * dims_full + 1 per axis, first index fastest (the nodes CurvilinearGrids.jl
 * derives its metric arrays from; consumed at PseudoTransientSolver.jl:122-153,
 * mesh_metric_cache.jl:39-250, flux_divergence.jl:11-14).  Cell metrics from
 * the mean node differences of the cell's edges, face metrics from the two
 * adjacent cells (normal derivative) and the face's own nodes (tangential);
 * L and spacing as for cdiff_set_metrics. */
int cdiff_set_nodes(cdiff_solver *h, const double *const *nodes, double L, const double *spacing);

/* 2D axisymmetric (RZ) geometry, x = r, y = z: r is the full-size array of
 * cell-centre radii (> 0 on the domain), NULL for planar (the default).  The
 * operator becomes div(alpha grad T) in cylindrical coordinates: the planar
 * operator plus alpha (1/r) dT/dr.  Must precede cdiff_set_metrics /
 * cdiff_set_nodes.  No counterpart in the reference APT: its RZ operator
 * exists only as a derivation (scripts/derivation_noncons_2d_rz.jl:111-126). */
int cdiff_set_axisymmetric(cdiff_solver *h, const double *r);

/* The conductivity law and heat capacity of the following steps: the kappa and
 * c_p arguments of nonlinear_thermal_conduction_step! and update_conductivity!
 * (src/nonlinear_thermal_conduction.jl:51-72, src/conductivity.jl:15). */
int cdiff_set_kappa(cdiff_solver *h, const cdiff_kappa *kappa, double cp);

int cdiff_set_density(cdiff_solver *h, const double *rho);   /* NULL resets to 1 */
int cdiff_set_source(cdiff_solver *h, const double *source); /* NULL resets to 0 */

/* Twilight-zone forcing f = A + B * alpha added to the update and the residual
 * (update.jl:23-71, residuals.jl:49-99): B carries the alpha dependence of
 * define_forcing_wavy (twilight_zone.jl:175-191).  NULL, NULL removes it. */
int cdiff_set_forcing(cdiff_solver *h, const double *A, const double *B);

/* update_conductivity!(solver, mesh, T, rho, cp, kappa): alpha = |kappa/(rho cp)|
 * over the full array (src/conductivity.jl:15-30,76-84). */
int cdiff_update_conductivity(cdiff_solver *h, const double *T);

/* One time step.  T (full, host) is read, and overwritten with solver.u at the
 * end, like copy!(T, solver.u) (PseudoTransientSolver.jl:229-230,346). */
int cdiff_step(cdiff_solver *h, double *T, double dt, const cdiff_opts *opts, cdiff_stats *stats,
               double *next_dt);

/* Device-resident variants: the temperature stays in HBM between steps.
 * cdiff_set_temperature / cdiff_get_temperature move it; cdiff_step_resident
 * is cdiff_step without the two host copies. */
int cdiff_set_temperature(cdiff_solver *h, const double *T);
int cdiff_get_temperature(cdiff_solver *h, double *T);
int cdiff_step_resident(cdiff_solver *h, double dt, const cdiff_opts *opts, cdiff_stats *stats,
                        double *next_dt);

/* applybcs!(bcs, mesh, u, 1) on a host array (uploaded, filled, downloaded). */
int cdiff_apply_bcs(cdiff_solver *h, double *u);

/* next_dt(u_new, u_prev, dt; ...) on host arrays of the solver's full size. */
int cdiff_next_dt(cdiff_solver *h, const double *u_new, const double *u_prev, double dt, double tol,
                  double maximum_u_change, double timestep_growth_factor, double *out);

/* next_dt on device arrays, allocating nothing.  No counterpart in the
 * reference: a variant of cdiff_next_dt for callers that cannot allocate
 * device memory (code under HIP graph capture, tight driver loops, harnesses
 * that forbid hipMalloc).  u_new, u_prev: device arrays of the n elements of
 * the full arrays (halo and corners included: next_dt takes its maxima over
 * every element, timestep_control.jl:43-62); the next time step is written to
 * the device double *out.  Enqueued on `stream` (a hipStream_t, NULL for the
 * null stream) without waiting.  The reduction's partial maxima live in a
 * fixed scratch array, one per GPU (256 KiB, held until the process ends),
 * so n is at most 134,217,728 (CDIFF_BAD_ARG beyond), and calls that overlap
 * on one GPU race on it.  On failure, errbuf (may be NULL) receives the
 * message. */
int cdiff_next_dt_device(const double *u_new, const double *u_prev, int64_t n, double dt, double tol,
                         double maximum_u_change, double timestep_growth_factor, double *out, void *stream,
                         char *errbuf, size_t errlen);

/* Copy a solver field (cdiff_field) to a host array of the full size. */
int cdiff_get_field(cdiff_solver *h, int field, double *out);

/* Overwrite the persistent state: field CDIFF_FIELD_Q+axis or CDIFF_FIELD_ALPHA. */
int cdiff_set_field(cdiff_solver *h, int field, const double *in);

/* The error-check state solver.last_err_check (PseudoTransientSolver.jl:45,83,321). */
int cdiff_get_last_iter(cdiff_solver *h, int32_t *out);
int cdiff_set_last_iter(cdiff_solver *h, int32_t v);

/* Zero q and q', set last_err_check = 0 (a freshly constructed solver). */
int cdiff_reset_state(cdiff_solver *h);

/* Measurement helper (no counterpart in the reference): run `nsweeps`
 * back-to-back fused sweeps of the current state on the handle's stream
 * (FP64, or the FP32 correction sweep with precision = CDIFF_PREC_FP32_DC),
 * with (check bit 0 set) or without the residual, bracketed by HIP events
 * (after one untimed launch of the same kind, so that the launches are timed
 * and not their enqueueing); returns the mean time per pseudo-iteration in
 * milliseconds.  With check bit
 * 1 set, the iterations run two per launch (2D without periodic faces; even
 * nsweeps); with bit 2 and precision = CDIFF_PREC_FP32_DC, the coefficient-form
 * correction sweep runs at its full depth per launch (after an FP32 step;
 * nsweeps a multiple of the depth).  Clobbers the iterate: call it after the
 * steps being measured. */
int cdiff_bench_sweeps(cdiff_solver *h, double dt, int precision, int check, int nsweeps, double *mean_ms);

const char *cdiff_last_error(const cdiff_solver *h);

/* ---------------------------------------------------------------------------
 * Learned warm start (no counterpart in the reference).  A 4-level
 * convolutional encoder-decoder (32 channels per level: the 8 features of
 * each 2x2 block of cells in, 3x3 convolutions with ReLU at 1/2, 1/4 and 1/8
 * of the grid, 2x2 mean pooling down, nearest up-sampling with additive
 * skips back up, a 32 -> 4 head giving each cell of a block its value; BF16
 * weights and activations, FP32 accumulation on MFMA) predicts
 * (T^{n+1} - T^n) / (dt max|r(T^n)|) per cell from features of the solver
 * state at the start of a step: T^n, the previous increment, the cell
 * Jacobian, the face diffusion numbers dt*alpha_face*g, and the explicit-Euler
 * increment dt*r(T^n).  With a network attached, an FP32_DC or FP64 step of a
 * 2D mesh without periodic faces starts from T^n + prediction when that start
 * has the smaller FP64 residual, and from T^n otherwise; it always iterates
 * to the same tolerance.  The network is trained on the GPU by its own
 * kernels on step pairs captured from the solver. */
typedef struct cdiff_net cdiff_net;

int cdiff_net_num_params(void);
/* New network, weights initialised from `seed` (He initialisation). */
int cdiff_net_create(uint64_t seed, cdiff_net **out);
int cdiff_net_destroy(cdiff_net *net);
int cdiff_net_set_params(cdiff_net *net, const float *params, int n);
int cdiff_net_get_params(cdiff_net *net, float *params, int n);
/* Training store in device memory: `capacity` samples of height x width
 * cells (features + target). */
int cdiff_net_dataset(cdiff_net *net, int height, int width, int capacity);
/* Capture from a solver whose resident temperature is T^n: stage 0 stores the
 * features of a step of size dt in `slot`; stage 1, after that step, the
 * target (T^{n+1} - T^n) / (dt max|r(T^n)|). */
int cdiff_net_capture(cdiff_net *net, cdiff_solver *h, int slot, int stage, double dt);
/* The training loss of a sample (prediction p, target t, e = p - t):
 * kind 0, the mean squared error |e|^2 / (height*width) (the default);
 * kind 1, |A e|^2 / |A t|^2 with A v = v + the face diffusion numbers
 * dt*alpha_face*g times the differences of v across the faces: the squared
 * ratio of the FP64 residuals of the predicted start and of T^n, to first
 * order. */
int cdiff_net_set_loss(cdiff_net *net, int kind);
/* `epochs` passes of Adam (lr, beta 0.9/0.999) over samples [0, nsamples) in
 * a seeded order; the mean loss of the last epoch in *loss. */
int cdiff_net_train(cdiff_net *net, int nsamples, int epochs, float lr, uint64_t seed, double *loss);
/* Prediction and loss of a stored sample (delta: height*width floats,
 * row-major by j, or NULL). */
int cdiff_net_eval(cdiff_net *net, int slot, float *delta, double *loss);
/* Loss and its gradient for host-provided features (bf16 bit patterns,
 * (height+2)*(width+2)*8 with a zero border) and target (height*width):
 * grad has cdiff_net_num_params() floats, delta height*width (either may be
 * NULL).  Test hook for the backward kernels. */
int cdiff_net_loss_grad(cdiff_net *net, int height, int width, const uint16_t *features, const float *target,
                        float *grad, float *delta, double *loss);
/* Prediction (height*width floats, row-major by j) for host-provided
 * features, through the inference kernels the warm start uses (MFMA head);
 * no counterpart in the reference, whose solver has no learned part. */
int cdiff_net_predict(cdiff_net *net, int height, int width, const uint16_t *features, float *delta);
const char *cdiff_net_last_error(const cdiff_net *net);
/* Activity threshold: the prediction is 0 (and carries no gradient) on cells
 * whose 3x3 neighbourhood has |dt r(T^n)| <= gate * dt max|r(T^n)|
 * (default 1e-5).  Part of the trained model: set it before training. */
int cdiff_net_set_gate(cdiff_net *net, float gate);
/* Attach (or detach, net = NULL) the warm start to a solver. */
int cdiff_set_warm_start(cdiff_solver *h, cdiff_net *net);
const char *cdiff_status_string(int status);
int cdiff_destroy(cdiff_solver *h);

#ifdef __cplusplus
}
#endif
#endif
