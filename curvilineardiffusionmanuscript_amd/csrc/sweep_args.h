// sweep_args.h -- kernel argument blocks of the fused APT sweeps.
#pragma once
#include "cdiff_dev.h"

namespace cdiff {

template <typename T>
struct Sweep2DArgs {
  Geo g;
  Ghost gh;
  KappaLaw kap;
  IterConst ic;
  // nullable: the iteration constants in device memory instead of `ic`, so
  // that one captured graph of launches serves every time step
  const IterConst* ic_dev;
  int harmonic;
  int ntiles, chunks, ch;
  const T* __restrict__ u_src;
  T* __restrict__ u_dst;
  const T* __restrict__ qx_src;
  T* __restrict__ qx_dst;
  const T* __restrict__ qy_src;
  T* __restrict__ qy_dst;
  const T* __restrict__ u_prev;
  const T* __restrict__ g11;
  const T* __restrict__ g22;
  const T* __restrict__ c12;
  const T* __restrict__ ac;
  const T* __restrict__ bc;
  const T* __restrict__ S;    // nullable: source term
  const T* __restrict__ rho;  // nullable: density (1 when null)
  const T* __restrict__ tzA;  // nullable: forcing f = tzA + tzB * alpha
  const T* __restrict__ tzB;
  const T* __restrict__ frozen_alpha;  // nullable: subcycle_conductivity = false
  const T* __restrict__ frozen_theta;
  const T* __restrict__ frozen_dtau;
  // correction mode (defect correction): the state is the correction e, the
  // right-hand side `base` replaces u_prev/dt + S, and theta/dtau come from
  // the frozen alpha (alpha_only) instead of stored arrays
  const T* __restrict__ base;
  int alpha_only;
  // Newton correction (optional): the FP64 iterate u-bar and d(alpha)/du at
  // u-bar; the face flux of the correction then carries the linearised
  // conductivity change, -(d abar) * delta(u-bar)
  const T* __restrict__ ubar;
  const T* __restrict__ dalpha;
  // two-iteration launches with a check: the iterate before the last
  // iteration (domain cells), for the step epilogue (nullable)
  T* __restrict__ u_pre;
  double* part_sum;
  double* part_max;
  const Ctrl* ctrl;
};

#ifndef SWEEP3D_BJ
#define SWEEP3D_BJ 8  // waves (rows) per 3D block; rows 1..BJ-2 are owned
#endif

template <typename T>
struct Sweep3DArgs {
  Geo g;
  Ghost gh;
  KappaLaw kap;
  IterConst ic;
  int harmonic;
  int ntiles, chunks, ch, stripsI;
  const T* __restrict__ u_src;
  T* __restrict__ u_dst;
  const T* __restrict__ qx_src;
  T* __restrict__ qx_dst;
  const T* __restrict__ qy_src;
  T* __restrict__ qy_dst;
  const T* __restrict__ qz_src;
  T* __restrict__ qz_dst;
  const T* __restrict__ u_prev;
  const T* coef[9];  // g11 g22 g33, 0.25*g_xi.eta, 0.25*g_xi.zeta, 0.25*g_eta.zeta, alpha_c beta_c gamma_c
  int64_t coef_stride;  // coef[c] == coef[0] + c * coef_stride (one block)
  const T* __restrict__ S;
  const T* __restrict__ rho;
  const T* __restrict__ tzA;
  const T* __restrict__ tzB;
  const T* __restrict__ frozen_alpha;
  const T* __restrict__ frozen_theta;
  const T* __restrict__ frozen_dtau;
  // correction mode (defect correction): the state is the correction e, the
  // right-hand side `base` replaces u_prev/dt + S, and theta/dtau come from
  // the frozen alpha (alpha_only) instead of stored arrays
  const T* __restrict__ base;
  int alpha_only;
  double* part_sum;
  double* part_max;
  const Ctrl* ctrl;
};

template <typename T>
void launch_sweep3d(const Sweep3DArgs<T>& A, bool check, int nblocks, hipStream_t s);
// the same sweep over 64 x 8 columns with one-cell rings (sweep3d_col.hip)
template <typename T>
void launch_sweep3d_col(const Sweep3DArgs<T>& A, bool check, int nblocks, hipStream_t s);

template <typename T>
void launch_sweep2d_pair(const Sweep2DArgs<T>& A, bool check, int nblocks, hipStream_t s);

template <typename T>
void launch_sweep2d_tile(const Sweep2DArgs<T>& A, bool check, int nblocks, hipStream_t s);

template <typename T>
void launch_sweep2d(const Sweep2DArgs<T>& A, bool check, int nblocks, hipStream_t s);

// FP32 correction sweep in coefficient form (sweep2d_lin.hip).
struct Lin2DArgs {
  Geo g;
  int ntiles, chunks;  // tiles of lin2d_tile(kmax, layout) cells; chunks along j
  int nit;             // iterations of this launch, 1..kmax
  int rows_per_chunk;  // marching sweep: rows of a chunk
  int strips;          // marching sweep: strips of 64 - 2 kmax columns
  int zero_src;        // the correction and its fluxes start from 0 (e_src, q*_src not read)
  float inv_dt;
  const float* __restrict__ e_src;
  float* __restrict__ e_dst;
  const float* __restrict__ qx_src;
  float* __restrict__ qx_dst;
  const float* __restrict__ qy_src;
  float* __restrict__ qy_dst;
  const float* __restrict__ base;  // right-hand side (FP64 residual)
  const float* __restrict__ cc;    // dtau / (1 + dtau/dt) per domain cell
  const float* m[5];               // g11 + ac/2, ac/2 - g11, g22 + bc/2, bc/2 - g22, c12
  const float* fx[3];              // P, M, W of the xi faces
  const float* fy[3];              // P, M, W of the eta faces
  double* part_sum;
  double* part_max;
  const Ctrl* ctrl;
};

// kmax: 2, 4, 6 or 8 iterations per launch at most.  layout 0: tiles of
// 64 x (8 or 16) cells, region cells spread over the threads; layout 1:
// region rows aligned to wavefronts, tiles of (64 - 2 kmax) x (32 - 2 kmax)
void lin2d_tile(int kmax, int layout, int* cols, int* rows);
void launch_sweep2d_lin(const Lin2DArgs& A, int kmax, int layout, bool check, int nblocks, hipStream_t s);
// the marching sweep (sweep2d_march.hip): strips of 64 - 2 kmax columns, one
// wavefront per strip and chunk of rows, exactly kmax iterations per launch;
// partial sums: lin2d_march_parts(strips, chunks) entries
int lin2d_march_blocks(int strips, int chunks);
int lin2d_march_parts(int strips, int chunks);
// wavefronts of the marching sweep resident on one CU at once (0 if unknown)
int lin2d_march_resident_waves(int kmax);
void launch_sweep2d_march(const Lin2DArgs& A, int kmax, bool check, hipStream_t s);

// FP32 correction sweep in coefficient form, 3D (sweep3d_lin.hip): 64 x 8
// columns marching over chunks of `ch` planes.
struct Lin3DArgs {
  Geo g;
  int ntiles, chunks, stripsI, ch;
  int zero_src;  // the correction and its fluxes start from 0 (e_src, q_src not read)
  float inv_dt;
  const float* __restrict__ e_src;
  float* __restrict__ e_dst;
  const float* q_src[3];
  float* q_dst[3];
  const float* __restrict__ base;
  const float* __restrict__ cc;
  const float* m;  // m0..m5 (x_c, x_w, y_c, y_s, z_c, z_s), cxe, cxz, cez: array q at m + q * stride
  const float* f;  // [axis][P, M, W]: array (axis, c) at f + (3 axis + c) * stride
  int64_t stride;
  double* part_sum;
  double* part_max;
  const Ctrl* ctrl;
};
void launch_sweep3d_lin(const Lin3DArgs& A, bool check, int nblocks, hipStream_t s);
// two iterations per launch on 64 x 8 columns (sweep3d_lin2.hip)
void launch_sweep3d_lin2(const Lin3DArgs& A, bool check, int nblocks, hipStream_t s);

}  // namespace cdiff
