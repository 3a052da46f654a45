/* apt_oracle.h -- CPU restatement of the reference APT solver (test
 * infrastructure only; see apt_oracle.c). */
#ifndef APT_ORACLE_H
#define APT_ORACLE_H
#include <stdint.h>

enum { ORA_OK = 0, ORA_NONFINITE = 1, ORA_MAX_ITER = 2, ORA_INVALID_ALPHA = 3, ORA_INVALID_SOURCE = 4,
       ORA_BAD_ARG = 5, ORA_INVALID_SCALAR = 6 };
enum { ORA_BC_NEUMANN = 0, ORA_BC_DIRICHLET = 1, ORA_BC_PERIODIC = 2, ORA_BC_FIXED_NEG_SYMMETRY = 3 };
enum { ORA_MEAN_HARMONIC = 0, ORA_MEAN_ARITHMETIC = 1 };

typedef struct {
  int32_t kind;
  int32_t pad;
  double value;
} ora_bc;

/* Mirrors the fields of PseudoTransientSolver (PseudoTransientSolver.jl:28-46).
 * All arrays are owned by the caller. */
typedef struct ora_solver {
  int32_t ndim;
  int32_t h;
  int64_t n[3];
  int64_t N;
  int64_t lo[3], hi[3];
  int32_t mean;
  int32_t last_err_check;
  ora_bc bcs[6];
  double L;
  double spacing[3];
  double kappa0, t_exp, rho_exp, cp;
  int32_t t_exp_is_int, rho_exp_is_int;
  double *u, *u_prev, *S, *res, *alpha, *theta, *dtau, *rho;
  double *q[3], *qp[3];
  double *cell[9];  /* cell[a*ndim + c] = d(xi_a)/d(x_c) at cell centres */
  double *edge[30]; /* edge[b*(ndim*ndim+1) + a*ndim + c] = hat metric; +ndim*ndim = J */
  double *cache[3];
  double *tzA, *tzB; /* twilight-zone forcing f = tzA + tzB*alpha, or NULL */
  double *tmp_u;
  char err[512];
  /* 2D axisymmetric (RZ, x = r): cell-centre radius, or NULL (planar).  Not
   * in the reference APT; the operator gains (1/r) dT/dr (see
   * ora_update_metric_cache). */
  double *rz_r;
} ora_solver;

typedef struct {
  int32_t max_iter;
  int32_t norm_kind; /* 0 = CPU RMS, 1 = GPU Inf/sqrt(N) */
  double rtol, atol, cfl;
  int32_t apply_cutoff, calculate_next_dt, subcycle_conductivity, enforce_positivity;
  double tol, maximum_u_change, timestep_growth_factor;
} ora_opts;

typedef struct {
  double rel_err, abs_err;
  int32_t niter, nchecks;
  double time;
} ora_stats;

int ora_init(ora_solver *s);
double ora_kappa(const ora_solver *s, double rho, double T);
int ora_applybcs(ora_solver *s, double *u, int nhalo);
void ora_update_conductivity(ora_solver *s, const double *T);
void ora_update_iteration_params(ora_solver *s, double Vpdtau, double dt);
void ora_compute_flux(ora_solver *s);
double ora_flux_divergence(const ora_solver *s, double *const *q, int64_t i, int64_t j, int64_t k);
void ora_compute_update(ora_solver *s, double dt);
void ora_update_residual(ora_solver *s, double dt);
double ora_l2_norm(const ora_solver *s, const double *a, int kind);
void ora_cutoff(ora_solver *s, double *A);
double ora_next_dt(const ora_solver *s, const double *unew, const double *uold, double dt, double tol,
                   double maximum_u_change, double growth);
int ora_validate_scalar(ora_solver *s, const double *a, const char *name, int positivity);
void ora_update_metric_cache(ora_solver *s);
void ora_schedule(int last_iter_count, int *min_err_check, int *err_interval);
int ora_is_check(int iter, int min_err_check, int err_interval);
int ora_step(ora_solver *s, double *T, double dt, const ora_opts *o, ora_stats *st, double *next_dt_out);
int ora_num_threads(void);
void ora_set_num_threads(int n);

#endif
