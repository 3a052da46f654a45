// dc.h -- kernels of the FP32/FP64 defect-correction step.
#pragma once
#include "cdiff_dev.h"

namespace cdiff {

// an inner correction solve whose residual has not halved within this many
// sweeps stops (k_inner_finalize)
constexpr int DC_STALL_SWEEPS = 192;

struct DcCoefs {
  const double* c[9];
};

__global__ void k_dc_update(Geo g, Ghost gh, KappaLaw kap, const double* rho, const double* u_old, const float* e,
                            double* u_new, double* a64, float* a32, float* ub32, float* da32);
__global__ void k_dc_residual(Geo g, int harmonic, double dt, const double* u, const double* a, const double* uprev,
                              DcCoefs cf, const double* S, const double* tzA, const double* tzB, float* r32,
                              double* psum, double* pmax);
__global__ void k_inner_finalize(Ctrl* cin, const Ctrl* cout, const double* psum, const double* pmax, int nparts,
                                 double nfull, int norm_kind, double eta, int iter, int cap, double floor_frac,
                                 double rtol, double atol, int out_buf);
__global__ void k_fill32(int64_t n, float v, float* dst);
__global__ void k_ctrl_set_init(Ctrl* ctrl, const Ctrl* src, double rtol, double atol);
__global__ void k_lin_metrics(Geo g, const double* g11, const double* g22, const double* c12, const double* ac,
                              const double* bc, float* m0, float* m1, float* m2, float* m3, float* m4);
struct LinOut {
  float* fx[3];
  float* fy[3];
  float* fz[3];
  float* cc;
  double dtau_scale, theta_scale;  // of dtau and theta (1: the reference's)
};
__global__ void k_lin_metrics3d(Geo g, DcCoefs cf, float* m0, float* m1, float* m2, float* m3, float* m4, float* m5,
                                float* m6, float* m7, float* m8);
// Fused outer pass of the 2D step (dc_outer2d.hip): u_new = u_old + e with
// local BC rules (no periodic faces), alpha, the FP64 residual (into r32 and
// block partials over 64 x DC_OUTER2D_TY tiles, one block of 64 x
// DC_OUTER2D_TY threads each), and the correction-sweep coefficients.
constexpr int DC_OUTER2D_TY = 8;  // (16: 640 us against 509 at 4096^2)
struct DcOuter2D {
  Geo g;
  Ghost gh;
  KappaLaw kap;
  IterConst ic;
  int harmonic, newton, ntiles, chunks;
  double t_exp;
  double dtau_scale, theta_scale;  // of the correction sweep's pseudo-time step and damping (1: the reference's)
  const double* rho;
  const double* u_old;
  const float* e;  // nullable: no correction
  // when e_sel is set, the correction is e_buf0 or e_buf1 as e_sel->final_buf
  // says (the buffer the correction solve's last check saw written)
  const Ctrl* e_sel;
  const float* e_buf0;
  const float* e_buf1;
  double* u_new;
  const double* uprev;
  const double* c[5];  // g11 g22 c12 ac bc
  const double *S, *tzA, *tzB;
  float* r32;
  float* fx[3];
  float* fy[3];
  float* cc;
  double* psum;
  double* pmax;
};
// unit: no density array and cp = 1
void launch_dc_outer2d(const DcOuter2D& A, int nblk, bool unit, hipStream_t stream);
// Fused outer pass of the 3D step (dc_outer3d.hip): the same work over 64 x
// DC_OUTER3D_LY columns marching up chunks of `ch` planes.
constexpr int DC_OUTER3D_LY = 8;
struct DcOuter3D {
  Geo g;
  Ghost gh;
  KappaLaw kap;
  IterConst ic;
  int harmonic, newton, ntiles, chunks, stripsI, ch;
  double t_exp;
  double dtau_scale, theta_scale;  // of the correction sweep's pseudo-time step and damping (1: the reference's)
  const double* rho;
  const double* u_old;
  const float* e;  // nullable: no correction
  // when e_sel is set, the correction is e_buf0 or e_buf1 as e_sel->final_buf
  // says (the buffer the correction solve's last check saw written)
  const Ctrl* e_sel;
  const float* e_buf0;
  const float* e_buf1;
  double* u_new;
  const double* uprev;
  const double* c;  // g11 g22 g33 g12 g13 g23 ac bc gc: array q at c + q * cstride
  int64_t cstride;
  const double *S, *tzA, *tzB;
  float* r32;
  float* f;  // [axis][P, M, W]: array (axis, c) at f + (3 axis + c) * fstride
  int64_t fstride;
  float* cc;
  double* psum;
  double* pmax;
};
void launch_dc_outer3d(const DcOuter3D& A, int nblk, bool unit, hipStream_t stream);
__global__ void k_lin_coefs(Geo g, Ghost gh, int harmonic, IterConst ic, double t_exp, int newton, const double* rho,
                            const double* u, const double* a, LinOut out);
__global__ void k_to_f32(int64_t n, const double* src, float* dst);

}  // namespace cdiff
