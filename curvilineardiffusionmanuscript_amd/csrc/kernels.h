// kernels.h -- declarations of the non-sweep kernels of libcdiff.
#pragma once
#include "cdiff_dev.h"

namespace cdiff {

struct MetricIn {
  const double* cell[9];
  const double* edge[30];
};

struct MetricOut {
  double* cell[9];
  double* edge[30];
};
// cell and face metrics from the ghost-inclusive node grid (metrics.hip)
__global__ void k_node_metrics2d(Geo g, const double* X, const double* Y, int64_t s1, MetricOut out);
__global__ void k_node_metrics3d(Geo g, const double* X, const double* Y, const double* Z, int64_t s1, int64_t s2,
                                 MetricOut out);

struct Coefs {
  const double* c[9];
};

__global__ void k_metric_coefs(Geo g, MetricIn m, double* out0, double* out1, double* out2, double* out3, double* out4,
                               double* out5, double* out6, double* out7, double* out8, const double* rz);
__global__ void k_alpha_full(Geo g, KappaLaw kap, const double* rho, const double* dom_src, const double* halo_src,
                             double* alpha);
__global__ void k_iter_params_full(Geo g, IterConst ic, const double* rho, const double* alpha, double* dtau,
                                   double* theta);
template <typename T>
__global__ void k_convert(Geo g, const double* src, T* dst);
__global__ void k_copy(int64_t n, const double* src, double* dst);
__global__ void k_cell_jacobian2d(Geo g, const double* xx, const double* xy, const double* ex, const double* ey,
                                  double* jac);
__global__ void k_merge(Geo g, const double* dom_src, const double* rest, double* out);
__global__ void k_fill(int64_t n, double v, double* dst);
__global__ void k_validate(Geo g, const double* a, int positivity, int shift, Ctrl* ctrl);
__global__ void k_ctrl_reset(Ctrl* ctrl, double rtol, double atol);
// k_ctrl_reset(ctrl, 0, 0), done when the outer block `follow` is done or has failed
__global__ void k_ctrl_reset_follow(Ctrl* ctrl, const Ctrl* follow);
__global__ void k_ctrl_validated(Ctrl* ctrl);
__global__ void k_finalize(Ctrl* ctrl, const double* psum, const double* pmax, int nparts, int iter, int max_iter,
                           double rtol, double atol, double nfull, int norm_kind, int check);
__global__ void k_halo_fill(Geo g, Ghost gh, double* u, int outer);
__global__ void k_epilogue_u(Geo g, const double* u, double* T, int positivity, int cutoff);
__global__ void k_next_dt_partial(Geo g, const double* unew, const double* uold, const double* red, double tol,
                                  int pass, double* part);
__global__ void k_reduce_max(const double* part, int n, double* out);
__global__ void k_next_dt_final(const double* mx, double dt, double muc, double growth, double* out);
__global__ void k_plain_flux(Geo g, int axis, int harmonic, const double* u, const double* alpha, double* qp);
// fused step-boundary passes (step_kernels.hip): STEP_THREADS threads per
// block, at most STEP_BLOCKS blocks (one partial maximum per block)
constexpr int STEP_THREADS = 256;
constexpr int STEP_BLOCKS = 2048;
__global__ void k_step_begin(Geo g, const double* T, double* u0, double* u1, double* up, const double* alpha,
                             int positivity, Ctrl* ctrl, double* part);
__global__ void k_step_end(Geo g, const double* u, double* T, int positivity, int cutoff, const double* up,
                           const double* umax, double tol, double* part);
void launch_dc_end(const Geo& g, const KappaLaw& kap, const double* rho, int harmonic, const double* u,
                   double* ucopy, double* alpha, double* const q[3], hipStream_t stream);
__global__ void k_residual_field(Geo g, double dt, const double* u, const double* uprev, const double* q0,
                                 const double* q1, const double* q2, Coefs cf, const double* S, const double* tzA,
                                 const double* tzB, const double* alpha, double* res);

}  // namespace cdiff
