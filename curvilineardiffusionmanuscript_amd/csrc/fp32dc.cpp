// fp32dc.cpp -- the mixed-precision time step: FP32 APT sweeps inside an FP64
// defect-correction loop.
//
//   u <- T;  r <- r64(u)                      (FP64 residual of the reference
//                                              discrete equations, FP64 norm)
//   while |r| > atol (and rel > rtol):
//     solve  e/dt + div(q'_alpha(e)) = r     with FP32 APT sweeps, alpha frozen
//                                              at u, to |r_inner| <= eta |r|
//     u <- u + e;  r <- r64(u)
//
// The outer residual is exactly the quantity the reference tests against its
// tolerance (PseudoTransientSolver.jl:280-316, residuals.jl:24-47), evaluated
// in FP64, so the step ends at the same FP64 residual tolerance as the FP64
// APT.  The inner equation is the frozen-conductivity (Picard) linearisation
// of the residual around u: the outer iteration converges for the nonlinear
// kappa through the FP64 residual.  The correction e is small, so FP32 holds
// it to full relative precision; u itself never goes through FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <utility>
#include <vector>

#include "dc.h"
#include "kernels.h"
#include "solver.h"
#include "sweep_args.h"

using namespace cdiff;

#define HIP_TRY(expr)                                 \
  do {                                                \
    hipError_t _e = (expr);                           \
    if (_e != hipSuccess) return hip_fail(_e, #expr); \
  } while (0)

#define LAUNCH_CHECK(name)                                        \
  do {                                                            \
    hipError_t _e = hipGetLastError();                            \
    if (_e != hipSuccess) return hip_fail(_e, "launch of " name); \
  } while (0)

namespace {

dim3 grid_n(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

int cdiff_solver::ensure_fp32() {
  auto alloc32 = [&](float** p) -> int {
    if (*p) return CDIFF_OK;
    HIP_TRY(hipMalloc((void**)p, sizeof(float) * g.total));
    owned32.push_back(*p);
    hipLaunchKernelGGL(k_fill32, grid_n(g.total), dim3(256), 0, stream, g.total, 0.0f, *p);
    LAUNCH_CHECK("k_fill32");
    return CDIFF_OK;
  };
  int rc;
  for (int b = 0; b < 2; ++b)
    if ((rc = alloc32(&u32[b]))) return rc;
  for (int a = 0; a < ndim; ++a)
    for (int b = 0; b < 2; ++b)
      if ((rc = alloc32(&q32[a][b]))) return rc;
  if ((rc = alloc32(&base32))) return rc;
  if ((rc = alloc32(&a32))) return rc;
  dc_newton = tune.dc_newton != 0 && kap.t_exp != 0.0 && (ndim == 2 || lin_ok());
  if (dc_newton) {
    if ((rc = alloc32(&ub32))) return rc;
    if ((rc = alloc32(&da32))) return rc;
  }
  const int ncoef = ndim == 2 ? 5 : 9;
  if (!coef32[0]) {
    // one block, array c at coef32[0] + c * total (as the FP64 coefficients)
    float* blk = nullptr;
    HIP_TRY(hipMalloc((void**)&blk, sizeof(float) * g.total * ncoef));
    owned32.push_back(blk);
    for (int c = 0; c < ncoef; ++c) coef32[c] = blk + g.total * c;
    coefs32_valid = false;
  }
  if (!a64) {
    HIP_TRY(hipMalloc((void**)&a64, sizeof(double) * g.total));
    owned.push_back(a64);
  }
  if (!ctrl_in) {
    HIP_TRY(hipMalloc((void**)&ctrl_in, sizeof(Ctrl)));
    owned.push_back(ctrl_in);
  }
  const int64_t nfull = (int64_t)g.n[0] * g.n[1] * g.n[2];
  // one partial per 256 cells (per-cell residual kernels) or per tile of the
  // fused outer pass, whose block count is rounded up to a multiple of 8
  // (64 x DC_OUTER2D_TY tiles in 2D, 64 x DC_OUTER3D_LY columns of at least
  // one plane in 3D)
  const int tile_rows = ndim == 3 ? std::min(DC_OUTER2D_TY, DC_OUTER3D_LY) : DC_OUTER2D_TY;
  const int64_t tiles = (int64_t)((g.hi[0] - g.lo[0] + 64) / 64) * ((g.hi[1] - g.lo[1] + tile_rows) / tile_rows) *
                        (ndim == 3 ? (g.hi[2] - g.lo[2] + 1) : 1);
  const int nres = (int)std::max<int64_t>((nfull + 255) / 256, (tiles + 7) / 8 * 8);
  if (nres > dc_nparts) {
    HIP_TRY(hipMalloc((void**)&dc_psum, sizeof(double) * nres));
    owned.push_back(dc_psum);
    HIP_TRY(hipMalloc((void**)&dc_pmax, sizeof(double) * nres));
    owned.push_back(dc_pmax);
    dc_nparts = nres;
  }
  if (!coefs32_valid) {
    for (int c = 0; c < ncoef; ++c) {
      hipLaunchKernelGGL(k_to_f32, grid_n(g.total), dim3(256), 0, stream, g.total, (const double*)coef[c], coef32[c]);
      LAUNCH_CHECK("k_to_f32");
    }
    coefs32_valid = true;
  }
  if (rho && !aux32_valid) {
    if ((rc = alloc32(&rho32))) return rc;
    hipLaunchKernelGGL(k_to_f32, grid_n(g.total), dim3(256), 0, stream, g.total, (const double*)rho, rho32);
    LAUNCH_CHECK("k_to_f32");
  }
  aux32_valid = true;
  // coefficient-form sweep: no periodic faces
  lin = lin_ok();
  if (lin && ndim == 2) {
    // eight iterations per launch, marching up strips of 48 columns (4096^2:
    // 39.0 us per iteration; 54.9 at depth 6, 74.1 at depth 4).  On
    // wave-aligned 52 x 20 tiles six iterations take 79 us per iteration (100
    // at depth 4 on 64 x 16 tiles): tiles re-read their ring
    lin_k = tune.dc_depth2d;
    const int ni = g.hi[0] - g.lo[0] + 1, nj = g.hi[1] - g.lo[1] + 1;
    lin_layout = tune.dc_layout2d;
    if (lin_layout == 2 && lin_k == 2) lin_layout = 1;  // the marching sweep runs 4, 6 or 8 iterations
    // the marching sweep addresses its arrays with 32-bit byte offsets
    if (lin_layout == 2 && (int64_t)g.total * (int64_t)sizeof(float) > (int64_t)UINT32_MAX) lin_layout = 1;
    int cols = 64, rows = 16;
    // (launches of fewer than lin_k iterations use the wave-aligned tiles)
    lin2d_tile(lin_k, lin_layout == 0 ? 0 : 1, &cols, &rows);
    lin_chunks = (nj + rows - 1) / rows;
    lin_tiles = ((ni + cols - 1) / cols) * lin_chunks;
    if (lin_layout == 2) {
      // marching sweep: strips of 64 - 2 lin_k columns, one wavefront per
      // strip and chunk of rows; each wavefront marches 2 lin_k rows more
      // than its chunk holds.  The launch time follows
      //   rounds x (rows per chunk + 2 lin_k)  row steps,
      // rounds = the wavefronts over those resident at once (measured at
      // 4096^2 for chunks of 32 to 96 rows: a partly filled last round costs
      // nearly a full one), so the chunk height is the one that minimises it,
      // at least 2 lin_k rows (4096^2, depth 8, two wavefronts per SIMD: 23
      // chunks of 179 rows in one round, 195 row steps, 39.0 us per
      // iteration; chunks of 64 rows took three rounds, 240 row steps, 45.5)
      march_strips = (ni + 63 - 2 * lin_k) / (64 - 2 * lin_k);
      const int64_t slots = (int64_t)ncu * std::max(1, lin2d_march_resident_waves(lin_k));
      int64_t best = -1;
      march_rows = nj;
      for (int c = 1; c <= std::max(1, nj / (2 * lin_k)); ++c) {
        const int rws = (nj + c - 1) / c;
        if (c > 1 && (nj + rws - 1) / rws != c) continue;  // the same chunks as a smaller c
        const int64_t rounds = ((int64_t)c * march_strips + slots - 1) / slots;
        const int64_t cost = rounds * (rws + 2 * lin_k);
        if (best < 0 || cost < best) {
          best = cost;
          march_rows = rws;
        }
      }
      march_rows = std::max(2 * lin_k, march_rows);
      if (tune.dc_march_rows > 0) march_rows = std::max(2 * lin_k, (int)tune.dc_march_rows);
      march_chunks = (nj + march_rows - 1) / march_rows;
    }
  } else if (lin) {
    // 3D: one or two iterations per launch; 64 x 8 columns over chunks of planes
    lin_k = tune.dc_depth3d;
    const int ni = g.hi[0] - g.lo[0] + 1, nj = g.hi[1] - g.lo[1] + 1, nk = g.hi[2] - g.lo[2] + 1;
    // about two 512-thread blocks per CU (256^3: 64 planes per chunk)
    const int64_t cols = (int64_t)((ni + 63) / 64) * ((nj + 7) / 8);
    const int dflt = (int)std::min<int64_t>(256, std::max<int64_t>(4, (cols * nk + 2 * ncu - 1) / (2 * ncu)));
    lin3_ch = std::max(1, std::min(nk, tune.dc_planes3d > 0 ? (int)tune.dc_planes3d : dflt));
    lin3_strips = (ni + 63) / 64;
    lin_chunks = (nk + lin3_ch - 1) / lin3_ch;
    lin_tiles = lin3_strips * ((nj + 7) / 8) * lin_chunks;
    // the two-iteration sweep runs one 832-thread block per CU (256^3: 128 planes per chunk)
    const int dflt2 = (int)std::min<int64_t>(256, std::max<int64_t>(4, (cols * nk + ncu - 1) / ncu));
    lin3b_ch = std::max(1, std::min(nk, tune.dc_planes3d_pair > 0 ? (int)tune.dc_planes3d_pair : dflt2));
    lin3b_chunks = (nk + lin3b_ch - 1) / lin3b_ch;
    lin3b_tiles = lin3_strips * ((nj + 7) / 8) * lin3b_chunks;
    lin3b_nblocks = (lin3b_tiles + 7) / 8 * 8;
  }
  if (lin) {
    lin_nblocks = (lin_tiles + 7) / 8 * 8;
    const int march_parts = (ndim == 2 && lin_layout == 2) ? lin2d_march_parts(march_strips, march_chunks) : 0;
    const int nparts_need = std::max(std::max(lin_nblocks, lin3b_nblocks), march_parts);
    // one block each: the folded metrics m[c] at m[0] + c * total; the face
    // coefficients [axis][P, M, W] at fx[0] + (3 axis + c) * total
    auto alloc_block = [&](float** p, int n) -> int {
      HIP_TRY(hipMalloc((void**)p, sizeof(float) * g.total * n));
      owned32.push_back(*p);
      hipLaunchKernelGGL(k_fill32, grid_n(g.total * n), dim3(256), 0, stream, g.total * n, 0.0f, *p);
      LAUNCH_CHECK("k_fill32");
      return CDIFF_OK;
    };
    const int nm = ndim == 2 ? 5 : 9;
    if (!lin_m[0]) {
      float* blk = nullptr;
      if ((rc = alloc_block(&blk, nm))) return rc;
      for (int c = 0; c < nm; ++c) lin_m[c] = blk + g.total * c;
      lin_metrics_valid = false;
    }
    if (!lin_fx[0]) {
      float* blk = nullptr;
      if ((rc = alloc_block(&blk, 3 * ndim))) return rc;
      for (int c = 0; c < 3; ++c) {
        lin_fx[c] = blk + g.total * c;
        lin_fy[c] = blk + g.total * (3 + c);
        if (ndim == 3) lin_fz[c] = blk + g.total * (6 + c);
      }
    }
    if ((rc = alloc32(&lin_cc))) return rc;
    if (nparts_need > lin_nparts) {
      HIP_TRY(hipMalloc((void**)&lin_psum, sizeof(double) * nparts_need));
      owned.push_back(lin_psum);
      HIP_TRY(hipMalloc((void**)&lin_pmax, sizeof(double) * nparts_need));
      owned.push_back(lin_pmax);
      lin_nparts = nparts_need;
    }
    if (!lin_metrics_valid) {
      if (ndim == 2) {
        hipLaunchKernelGGL(k_lin_metrics, grid_n((int64_t)g.n[0] * g.n[1]), dim3(256), 0, stream, g,
                           (const double*)coef[0], (const double*)coef[1], (const double*)coef[2],
                           (const double*)coef[3], (const double*)coef[4], lin_m[0], lin_m[1], lin_m[2], lin_m[3],
                           lin_m[4]);
      } else {
        DcCoefs cf{};
        for (int t = 0; t < 9; ++t) cf.c[t] = coef[t];
        hipLaunchKernelGGL(k_lin_metrics3d, grid_n(g.total), dim3(256), 0, stream, g, cf, lin_m[0], lin_m[1],
                           lin_m[2], lin_m[3], lin_m[4], lin_m[5], lin_m[6], lin_m[7], lin_m[8]);
      }
      LAUNCH_CHECK("k_lin_metrics");
      lin_metrics_valid = true;
    }
  }
  // ghost rules of the correction: Dirichlet (and untouched) ghosts hold no
  // correction; copy / periodic / negated copies carry over unchanged
  gh32 = gh;
  for (int f = 0; f < 6; ++f) {
    if (gh32.mode[f] == G_CONST || gh32.mode[f] == G_KEEP) {
      gh32.mode[f] = G_CONST;
      gh32.val[f] = 0.0;
    }
  }
  return CDIFF_OK;
}

int cdiff_solver::launch_sweep32(double dt, double Vpdtau, int src, bool check, bool pair) {
  if (!u32[0]) return fail(CDIFF_BAD_ARG, "FP32 buffers not set up (run an FP32 step first)");
  const int dst = 1 - src;
  const float* r32 = rho ? rho32 : nullptr;
  if (ndim == 2) {
    Sweep2DArgs<float> A{};
    A.g = g;
    A.gh = gh32;
    A.kap = kap;
    A.ic.L = L;
    A.ic.Vpdtau = Vpdtau;
    A.ic.dt = dt;
    A.harmonic = mean == CDIFF_MEAN_HARMONIC;
    A.ntiles = ntiles;
    A.chunks = chunks;
    A.ch = ch;
    A.u_src = u32[src];
    A.u_dst = u32[dst];
    A.qx_src = q32[0][src];
    A.qx_dst = q32[0][dst];
    A.qy_src = q32[1][src];
    A.qy_dst = q32[1][dst];
    A.g11 = coef32[0];
    A.g22 = coef32[1];
    A.c12 = coef32[2];
    A.ac = coef32[3];
    A.bc = coef32[4];
    A.rho = r32;
    A.frozen_alpha = a32;
    A.alpha_only = 1;
    A.base = base32;
    if (dc_newton) {
      A.ubar = ub32;
      A.dalpha = da32;
    }
    A.part_sum = part_sum;
    A.part_max = part_max;
    A.ctrl = ctrl_in;
    if (pair)
      launch_sweep2d_pair<float>(A, check, nblocks, stream);
    else if (tile2d)
      launch_sweep2d_tile<float>(A, check, nblocks, stream);
    else
      launch_sweep2d<float>(A, check, nblocks, stream);
    LAUNCH_CHECK("sweep2d<float>");
    return CDIFF_OK;
  }
  Sweep3DArgs<float> A{};
  A.g = g;
  A.gh = gh32;
  A.kap = kap;
  A.ic.L = L;
  A.ic.Vpdtau = Vpdtau;
  A.ic.dt = dt;
  A.harmonic = mean == CDIFF_MEAN_HARMONIC;
  A.ntiles = ntiles;
  A.chunks = chunks;
  A.ch = ch;
  A.stripsI = strips3[0];
  A.u_src = u32[src];
  A.u_dst = u32[dst];
  A.qx_src = q32[0][src];
  A.qx_dst = q32[0][dst];
  A.qy_src = q32[1][src];
  A.qy_dst = q32[1][dst];
  A.qz_src = q32[2][src];
  A.qz_dst = q32[2][dst];
  for (int t = 0; t < 9; ++t) A.coef[t] = coef32[t];
  A.rho = r32;
  A.frozen_alpha = a32;
  A.alpha_only = 1;
  A.base = base32;
  A.part_sum = part_sum;
  A.part_max = part_max;
  A.ctrl = ctrl_in;
  A.coef_stride = g.total;
  if (col3d)
    launch_sweep3d_col<float>(A, check, nblocks, stream);
  else
    launch_sweep3d<float>(A, check, nblocks, stream);
  LAUNCH_CHECK("sweep3d<float>");
  return CDIFF_OK;
}

// Coefficients of the correction sweep at the FP64 iterate u[buf] (alpha in a64).
int cdiff_solver::lin_coefs(double dt, double Vpdtau, int buf) {
  LinOut out{};
  for (int c = 0; c < 3; ++c) {
    out.fx[c] = lin_fx[c];
    out.fy[c] = lin_fy[c];
    out.fz[c] = lin_fz[c];
  }
  out.cc = lin_cc;
  out.dtau_scale = dc_dtau_scale;
  out.theta_scale = dc_theta_scale;
  IterConst ic{L, Vpdtau, dt};
  hipLaunchKernelGGL(k_lin_coefs, grid_n(g.total), dim3(256), 0, stream, g, gh,
                     (int)(mean == CDIFF_MEAN_HARMONIC), ic, kap.t_exp, (int)dc_newton, (const double*)rho,
                     (const double*)u[buf], (const double*)a64, out);
  LAUNCH_CHECK("k_lin_coefs");
  lin_coefs_ready = true;
  return CDIFF_OK;
}

int cdiff_solver::launch_lin(double dt, int src, bool check, int nit, bool zero_src) {
  if (!lin || !lin_coefs_ready) return fail(CDIFF_BAD_ARG, "coefficient-form sweep not set up (run an FP32 step first)");
  if (nit < 1 || nit > lin_k) return fail(CDIFF_BAD_ARG, "iterations per launch out of range");
  const int dst = 1 - src;
  if (ndim == 3) {
    Lin3DArgs B{};
    B.g = g;
    const bool two = nit == 2;
    B.ntiles = two ? lin3b_tiles : lin_tiles;
    B.chunks = two ? lin3b_chunks : lin_chunks;
    B.stripsI = lin3_strips;
    B.ch = two ? lin3b_ch : lin3_ch;
    B.inv_dt = (float)(1.0 / dt);
    B.zero_src = zero_src;
    B.e_src = u32[src];
    B.e_dst = u32[dst];
    for (int a = 0; a < 3; ++a) {
      B.q_src[a] = q32[a][src];
      B.q_dst[a] = q32[a][dst];
    }
    B.base = base32;
    B.cc = lin_cc;
    B.m = lin_m[0];
    B.f = lin_fx[0];
    B.stride = g.total;
    B.part_sum = lin_psum;
    B.part_max = lin_pmax;
    B.ctrl = ctrl_in;
    if (two) {
      launch_sweep3d_lin2(B, check, lin3b_nblocks, stream);
      lin_last_parts = lin3b_nblocks;
    } else {
      launch_sweep3d_lin(B, check, lin_nblocks, stream);
      lin_last_parts = lin_nblocks;
    }
    LAUNCH_CHECK("sweep3d_lin");
    return CDIFF_OK;
  }
  Lin2DArgs A{};
  A.g = g;
  A.ntiles = lin_tiles;
  A.chunks = lin_chunks;
  A.nit = nit;
  A.zero_src = zero_src;
  A.inv_dt = (float)(1.0 / dt);
  A.e_src = u32[src];
  A.e_dst = u32[dst];
  A.qx_src = q32[0][src];
  A.qx_dst = q32[0][dst];
  A.qy_src = q32[1][src];
  A.qy_dst = q32[1][dst];
  A.base = base32;
  A.cc = lin_cc;
  for (int c = 0; c < 5; ++c) A.m[c] = lin_m[c];
  for (int c = 0; c < 3; ++c) {
    A.fx[c] = lin_fx[c];
    A.fy[c] = lin_fy[c];
  }
  A.part_sum = lin_psum;
  A.part_max = lin_pmax;
  A.ctrl = ctrl_in;
  if (lin_layout == 2 && nit == lin_k) {
    A.strips = march_strips;
    A.chunks = march_chunks;
    A.rows_per_chunk = march_rows;
    launch_sweep2d_march(A, lin_k, check, stream);
    lin_last_parts = lin2d_march_parts(march_strips, march_chunks);
    LAUNCH_CHECK("sweep2d_march");
    return CDIFF_OK;
  }
  launch_sweep2d_lin(A, lin_k, lin_layout == 0 ? 0 : 1, check, lin_nblocks, stream);
  lin_last_parts = lin_nblocks;
  LAUNCH_CHECK("sweep2d_lin");
  return CDIFF_OK;
}

bool cdiff_solver::lin_ok() const {
  if (!tune.dc_coefficient_form) return false;
  for (int f = 0; f < 2 * ndim; ++f)
    if (bcs[f].kind == CDIFF_BC_PERIODIC) return false;
  return true;
}

int cdiff_solver::dc_outer_pass(double dt, const double* u_old, const float* e, double* u_new, const double* uprev,
                                int iter, const cdiff_opts& o, bool finalize, const Ctrl* e_sel) {
  // (the fused passes keep alpha in LDS; the step's epilogue forms the final one)
  const int ni = g.hi[0] - g.lo[0] + 1, nj = g.hi[1] - g.lo[1] + 1;
  if (ndim == 3) {
    const int nk = g.hi[2] - g.lo[2] + 1;
    DcOuter3D B{};
    B.g = g;
    B.gh = gh;
    B.kap = kap;
    B.ic = IterConst{L, dc_Vpdtau, dt};
    B.dtau_scale = dc_dtau_scale;
    B.theta_scale = dc_theta_scale;
    B.harmonic = mean == CDIFF_MEAN_HARMONIC;
    B.newton = dc_newton;
    // planes per chunk: one block per CU resides at a time (registers), so
    // about one block per CU in all (256^3: 128 column blocks, chunks of 128
    // planes, 15.02 ms per step against 15.25 with chunks of 32), at least 8
    // planes (each chunk marches two planes more than it owns)
    const int cols3 = ((ni + 63) / 64) * ((nj + DC_OUTER3D_LY - 1) / DC_OUTER3D_LY);
    const int auto_ch = std::max(8, (nk + std::max(1, ncu / cols3) - 1) / std::max(1, ncu / cols3));
    B.ch = std::max(1, std::min(nk, tune.dc_outer_planes3d > 0 ? (int)tune.dc_outer_planes3d : auto_ch));
    B.stripsI = (ni + 63) / 64;
    B.chunks = (nk + B.ch - 1) / B.ch;
    B.ntiles = B.stripsI * ((nj + DC_OUTER3D_LY - 1) / DC_OUTER3D_LY) * B.chunks;
    B.t_exp = kap.t_exp;
    B.rho = rho;
    B.u_old = u_old;
    B.e = e;
    B.e_sel = e_sel;
    B.e_buf0 = u32[0];
    B.e_buf1 = u32[1];
    B.u_new = u_new;
    B.uprev = uprev;
    B.c = coef[0];
    B.cstride = g.total;
    B.S = S;
    B.tzA = tzA;
    B.tzB = tzB;
    B.r32 = base32;
    B.f = lin_fx[0];
    B.fstride = g.total;
    B.cc = lin_cc;
    B.psum = dc_psum;
    B.pmax = dc_pmax;
    const int nblk = (B.ntiles + 7) / 8 * 8;
    if (nblk > dc_nparts) return fail(CDIFF_HIP_ERROR, "partial-sum buffer too small");
    launch_dc_outer3d(B, nblk, rho == nullptr && kap.cp == 1.0, stream);
    LAUNCH_CHECK("k_dc_outer3d");
    if (finalize) {
      hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, stream, ctrl, (const double*)dc_psum,
                         (const double*)dc_pmax, nblk, iter, INT_MAX, o.rtol, o.atol, N_full, o.norm_kind, 1);
      LAUNCH_CHECK("k_finalize");
    }
    lin_coefs_ready = true;
    return CDIFF_OK;
  }
  DcOuter2D A{};
  A.g = g;
  A.gh = gh;
  A.kap = kap;
  A.ic = IterConst{L, dc_Vpdtau, dt};
  A.dtau_scale = dc_dtau_scale;
  A.theta_scale = dc_theta_scale;
  A.harmonic = mean == CDIFF_MEAN_HARMONIC;
  A.newton = dc_newton;
  A.chunks = (nj + DC_OUTER2D_TY - 1) / DC_OUTER2D_TY;
  A.ntiles = ((ni + 63) / 64) * A.chunks;
  A.t_exp = kap.t_exp;
  A.rho = rho;
  A.u_old = u_old;
  A.e = e;
  A.e_sel = e_sel;
  A.e_buf0 = u32[0];
  A.e_buf1 = u32[1];
  A.u_new = u_new;
  A.uprev = uprev;
  for (int c = 0; c < 5; ++c) A.c[c] = coef[c];
  A.S = S;
  A.tzA = tzA;
  A.tzB = tzB;
  A.r32 = base32;
  for (int c = 0; c < 3; ++c) {
    A.fx[c] = lin_fx[c];
    A.fy[c] = lin_fy[c];
  }
  A.cc = lin_cc;
  A.psum = dc_psum;
  A.pmax = dc_pmax;
  const int nblk = (A.ntiles + 7) / 8 * 8;
  if (nblk > dc_nparts) return fail(CDIFF_HIP_ERROR, "partial-sum buffer too small");
  launch_dc_outer2d(A, nblk, rho == nullptr && kap.cp == 1.0, stream);
  LAUNCH_CHECK("k_dc_outer2d");
  if (finalize) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, stream, ctrl, (const double*)dc_psum,
                       (const double*)dc_pmax, nblk, iter, INT_MAX, o.rtol, o.atol, N_full, o.norm_kind, 1);
    LAUNCH_CHECK("k_finalize");
  }
  lin_coefs_ready = true;
  return CDIFF_OK;
}

// norm (into ctrl) of the partial sums the last 2D outer pass left
int cdiff_solver::dc_finalize_outer(int iter, const cdiff_opts& o) {
  if (ndim != 2 || !lin) return fail(CDIFF_BAD_ARG, "no 2D outer pass to finalize");
  const int ni = g.hi[0] - g.lo[0] + 1, nj = g.hi[1] - g.lo[1] + 1;
  const int ntiles = ((ni + 63) / 64) * ((nj + DC_OUTER2D_TY - 1) / DC_OUTER2D_TY);
  const int nblk = (ntiles + 7) / 8 * 8;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, stream, ctrl, (const double*)dc_psum, (const double*)dc_pmax,
                     nblk, iter, INT_MAX, o.rtol, o.atol, N_full, o.norm_kind, 1);
  LAUNCH_CHECK("k_finalize");
  return CDIFF_OK;
}

// residual of the iterate in u[dstbuf] after u[dstbuf] = u[srcbuf] + e
int cdiff_solver::dc_residual(double dt, int srcbuf, const float* e, int dstbuf, int iter, const cdiff_opts& o) {
  const int64_t nfull = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (lin && (ndim == 2 || tune.dc_fused3d))
    return dc_outer_pass(dt, u[srcbuf], e, u[dstbuf], u_prev, iter, o, true);
  hipLaunchKernelGGL(k_dc_update, grid_n(nfull), dim3(256), 0, stream, g, gh, kap, (const double*)rho,
                     (const double*)u[srcbuf], e, u[dstbuf], a64, a32, (dc_newton && !lin) ? ub32 : nullptr,
                     (dc_newton && !lin) ? da32 : nullptr);
  LAUNCH_CHECK("k_dc_update");
  DcCoefs cf{};
  for (int t = 0; t < 9; ++t) cf.c[t] = coef[t];
  hipLaunchKernelGGL(k_dc_residual, grid_n(nfull), dim3(256), 0, stream, g, (int)(mean == CDIFF_MEAN_HARMONIC), dt,
                     (const double*)u[dstbuf], (const double*)a64, (const double*)u_prev, cf, (const double*)S,
                     (const double*)tzA, (const double*)tzB, base32, dc_psum, dc_pmax);
  LAUNCH_CHECK("k_dc_residual");
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, stream, ctrl, (const double*)dc_psum, (const double*)dc_pmax,
                     (int)((nfull + 255) / 256), iter, INT_MAX, o.rtol, o.atol, N_full, o.norm_kind, 1);
  LAUNCH_CHECK("k_finalize");
  // 3D coefficient form: the sweep coefficients of the new iterate
  if (lin) return lin_coefs(dt, dc_Vpdtau, dstbuf);
  return CDIFF_OK;
}

int cdiff_solver::step_fp32_dc(double dt, const cdiff_opts& o, double Vpdtau, cdiff_stats* st) {
  int rc = ensure_fp32();
  if (rc) return rc;
  // inner stop at eta times the outer residual, checked every ichk sweeps; in
  // 3D (one sweep per launch) shorter inner solves checked more often need
  // fewer sweeps in total (256^3 spots: 48 -> 32 sweeps per step); in 2D
  // a check every 8 to 12 sweeps: every eight-iteration launch checks
  // (4096^2 hot disk: 110 sweeps and 7.64 ms per step, against 125 sweeps and
  // 8.03 ms checking every second launch); at depth 6 every second launch
  const double eta = tune.dc_eta > 0.0 ? tune.dc_eta : (ndim == 3 ? 1e-2 : 1e-3);
  // the last correction only needs to reach the outer tolerance: inner
  // sweeps stop at floor_frac * atol at the latest
  const double floor_frac = tune.dc_floor;
  const int ichk = tune.dc_check > 0 ? (int)tune.dc_check : (ndim == 3 ? 2 : lin_k * std::max(1, 12 / lin_k));
  const int max_outer = tune.dc_max_outer;
  const int fixed_inner = o.inner_iters;
  const int64_t total_cap = (int64_t)o.max_iter;
  const bool pairs = pairs_ok();
  const int lin_nit = std::max(1, std::min(lin_k, tune.dc_launch_iters > 0 ? (int)tune.dc_launch_iters : lin_k));
  // pseudo-time step of the correction sweeps: the fixed point does not depend
  // on it, so it may differ from the FP64 APT's (scale 1 = the reference's)
  if (dc_cfl_scale < 0) dc_cfl_scale = tune.dc_cfl_scale;
  dc_Vpdtau = Vpdtau * dc_cfl_scale;
  // and on top of that a 1.3x larger dtau with theta unchanged: measured on
  // the 4096^2 and 1024^2 hot disks, 4096^2 and 2048^2 blobs and 256^3 spots,
  // 5-12 % fewer sweeps per step (a scale of 2 diverged on some of them)
  if (dc_dtau_scale_dflt < 0) {
    dc_dtau_scale_dflt = tune.dc_dtau_scale;
    dc_theta_scale_dflt = tune.dc_theta_scale;
  }
  dc_dtau_scale = dc_dtau_scale_dflt;
  dc_theta_scale = dc_theta_scale_dflt;
  // rtol > 0: the relative criterion divides by the reference's init_L2, the
  // residual after the FP64 APT's first pseudo-iteration from T^n
  // (PseudoTransientSolver.jl:280-293), whatever start the correction loop
  // takes.  One FP64 sweep with its check, normed on the inner control block;
  // the buffer it wrote gets T^n back.
  const bool want_init = o.rtol > 0.0;
  if (want_init) {
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(1), 0, stream, ctrl_in, 0.0, 0.0);
    LAUNCH_CHECK("k_ctrl_reset");
    if ((rc = launch_sweep64(o, dt, Vpdtau, cur, true, false))) return rc;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, stream, ctrl_in, (const double*)part_sum,
                       (const double*)part_max, nparts, 1, INT_MAX, 0.0, 0.0, N_full, o.norm_kind, 1);
    LAUNCH_CHECK("k_finalize");
    double* w = u[1 - cur];
    hipLaunchKernelGGL(k_copy, grid_n(g.total), dim3(256), 0, stream, g.total, (const double*)Tdev, w);
    LAUNCH_CHECK("k_copy");
    if (dirichlet_outer) {
      hipLaunchKernelGGL(k_halo_fill, grid_n((int64_t)g.n[0] * g.n[1] * g.n[2]), dim3(256), 0, stream, g, gh, w, 1);
      LAUNCH_CHECK("k_halo_fill");
    }
  }
  // u[0] holds T (prologue); the first FP64 residual, of T with its BCs, into
  // u[1] -- when the learned warm start's feature pass has just evaluated it,
  // only its norm (the start with the smaller residual is evaluated next)
  int a = 1;
  const bool r0_ready = warm_pending && warm_r0_ready && lin;
  if ((rc = r0_ready ? dc_finalize_outer(1, o) : dc_residual(dt, 0, nullptr, a, 1, o))) return rc;
  HIP_TRY(hipMemcpyAsync(&ctrl_h[2], ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  if (r0_ready && ctrl_h[2].done) {
    // T^n is already converged: its iterate with the BC ghosts in u[1]
    if ((rc = dc_residual(dt, 0, nullptr, a, 1, o))) return rc;
    HIP_TRY(hipMemcpyAsync(&ctrl_h[2], ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
  }
  st->warm_start = 0;
  const bool dc_debug = tune.verbose != 0;
  if (dc_debug) std::fprintf(stderr, "[dc] step dt=%.3e: residual of T^n %.3e\n", dt, ctrl_h[2].abs_);
  if (warm_pending && lin && !ctrl_h[2].done) {
    bool used = false;
    if ((rc = warm_start_apply(dt, o, &used))) return rc;
    st->warm_start = used ? 1 : 0;
    if (dc_debug) std::fprintf(stderr, "[dc] warm start %s: residual %.3e\n", used ? "used" : "rejected", ctrl_h[2].abs_);
  }
  warm_pending = false;
  if (want_init) {
    hipLaunchKernelGGL(k_ctrl_set_init, dim3(1), dim3(1), 0, stream, ctrl, (const Ctrl*)ctrl_in, o.rtol, o.atol);
    LAUNCH_CHECK("k_ctrl_set_init");
    HIP_TRY(hipMemcpyAsync(&ctrl_h[2], ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
  }
  int outer = 0;
  int64_t inner_total = 0;
  int inner_checks = 0;
  const int64_t nq = g.total;
  // the defect correction is abandoned (and the step redone by the FP64 APT)
  // when its outer residual turns non-finite, grows past 100x the first one
  // or has not converged after max_outer rounds
  dc_failed = false;
  const double r_first = ctrl_h[2].abs_;
  step_r_start = r_first;
  auto failed_now = [&]() {
    const Ctrl& c = ctrl_h[2];
    return c.status == CDIFF_NONFINITE || !(c.abs_ == c.abs_) || c.abs_ > 100.0 * r_first;
  };
  if (failed_now()) dc_failed = true;
  // A correction solve in flight: its iterations enqueued so far, the
  // segment the host reads next, and the buffer each launch wrote.
  struct Solve {
    int it = 0, seg = 0, ebuf = 0, cap = 0;
    bool pending = false;
    std::vector<std::pair<int, int>> ends;  // (last iteration, buffer written)
  };
  // enqueue iterations S.it + 1 .. S.it + len (at most S.cap) and the read of
  // the inner control block after them (ctrl_h[S.seg & 1], event ev[S.seg & 1])
  auto enqueue = [&](Solve& S, int len) -> int {
    const int cap = S.cap;
    auto is_chk = [&](int x) { return fixed_inner > 0 ? (x == cap) : (x % ichk == 0 || x == cap); };
    const int s = S.it + 1;
    const int e = (int)std::min<int64_t>(cap, (int64_t)s + len - 1);
    for (int t = s; t <= e;) {
      int tl = t;
      if (lin) {
        // up to lin_nit iterations; a check iteration ends the launch
        while (tl < e && tl - t + 1 < lin_nit && !is_chk(tl)) ++tl;
        if ((rc = launch_lin(dt, S.ebuf, is_chk(tl), tl - t + 1, t == 1))) return rc;
      } else {
        const bool pair = pairs && !is_chk(t) && t + 1 <= e;
        tl = pair ? t + 1 : t;
        if ((rc = launch_sweep32(dt, Vpdtau, S.ebuf, is_chk(tl), pair))) return rc;
      }
      const bool chk_last = is_chk(tl);
      S.ebuf ^= 1;
      S.ends.push_back({tl, S.ebuf});
      if (chk_last) {
        hipLaunchKernelGGL(k_inner_finalize, dim3(1), dim3(1024), 0, stream, ctrl_in, (const Ctrl*)ctrl,
                           (const double*)(lin ? lin_psum : part_sum), (const double*)(lin ? lin_pmax : part_max),
                           lin ? lin_last_parts : nparts, N_full, o.norm_kind, fixed_inner > 0 ? 0.0 : eta, tl, cap,
                           fixed_inner > 0 ? 0.0 : floor_frac, o.rtol, o.atol, S.ebuf);
        LAUNCH_CHECK("k_inner_finalize");
      }
      t = tl + 1;
    }
    S.it = e;
    HIP_TRY(hipMemcpyAsync(&ctrl_h[S.seg & 1], ctrl_in, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipEventRecord(ev[S.seg & 1], stream));
    return CDIFF_OK;
  };
  // the first segment runs to the check at or after the previous
  // correction's sweep count, later ones one check each; one segment is
  // always queued ahead of the host's read, so a correction that converges
  // leaves at most one check's launches (which return at once) behind it
  auto first_len = [&]() {
    return (int)std::min<int64_t>(
        INT_MAX / 2, std::max<int64_t>(ichk, (std::max<int64_t>(dc_inner_guess, 1) + ichk - 1) / ichk * ichk));
  };
  // The fused outer passes take the correction from the buffer that the
  // solve's last check recorded (k_inner_finalize), so the pass is enqueued
  // right behind the solve; behind the pass, the next solve's reset (done at
  // once when the pass converged or failed) and first segment, so the GPU
  // does not wait for the host between rounds.  The host reads the solve's
  // and the pass's control blocks together; a solve that diverged is dropped
  // and the pass redone without it (the next solve enqueued behind it is
  // then stale and is restarted).
  const bool fused = lin && (ndim == 2 || tune.dc_fused3d);
  Solve S;
  bool started = false;  // S's reset and first segment are already enqueued
  while (!ctrl_h[2].done && !dc_failed) {
    if (!started) {
      if (outer >= max_outer) {
        dc_failed = true;
        break;
      }
      // inner FP32 solve for the correction, from e = 0 and q = 0 (the
      // coefficient-form sweeps read no state on their first launch)
      if (!lin) {
        hipLaunchKernelGGL(k_fill32, grid_n(nq), dim3(256), 0, stream, nq, 0.0f, u32[0]);
        for (int ax = 0; ax < ndim; ++ax)
          for (int b = 0; b < 2; ++b)
            hipLaunchKernelGGL(k_fill32, grid_n(nq), dim3(256), 0, stream, nq, 0.0f, q32[ax][b]);
      }
      hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(1), 0, stream, ctrl_in, 0.0, 0.0);
      LAUNCH_CHECK("inner reset");
      // (the coefficient-form sweep's coefficients come with the residual pass)
      // the FP32 sweeps of a step never exceed max_iter: each correction solve
      // gets what is left of it, and none is left once it is used up
      const int64_t left = total_cap - inner_total;
      if (left <= 0) {
        dc_failed = true;
        break;
      }
      S = Solve{};
      S.cap = (int)std::min<int64_t>(fixed_inner > 0 ? fixed_inner : INT_MAX - 1, left);
      if ((rc = enqueue(S, first_len()))) return rc;
    }
    started = false;
    while (true) {
      if (S.pending) {
        HIP_TRY(hipEventSynchronize(ev[(S.seg - 1) & 1]));
        if (dc_debug)
          std::fprintf(stderr, "[dc]   inner check %d: residual %.3e\n", ctrl_h[(S.seg - 1) & 1].nchecks,
                       ctrl_h[(S.seg - 1) & 1].last_l2);
        if (ctrl_h[(S.seg - 1) & 1].done) break;
      }
      if (S.it >= S.cap) break;
      S.pending = true;
      ++S.seg;
      if ((rc = enqueue(S, ichk))) return rc;
    }
    const Ctrl outer_before = ctrl_h[2];  // the outer control block the pass starts from
    if (fused) {
      if ((rc = dc_outer_pass(dt, u[a], nullptr, u[1 - a], u_prev, outer + 2, o, true, ctrl_in))) return rc;
      HIP_TRY(hipMemcpyAsync(&ctrl_h[2], ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
    }
    HIP_TRY(hipMemcpyAsync(&ctrl_h[3], ctrl_in, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipEventRecord(ev[2], stream));
    // the next solve behind the pass, where the iteration budget cannot bind
    // within its first segment (so that it runs exactly as it would when
    // enqueued after the host's read)
    Solve N;
    bool next = false;
    if (fused && fixed_inner <= 0 && outer + 1 < max_outer) {
      const int flen = first_len();
      const int64_t left_next = total_cap - inner_total - S.it;  // S.it >= the solve's sweeps
      if (left_next > (int64_t)flen + ichk) {
        hipLaunchKernelGGL(k_ctrl_reset_follow, dim3(1), dim3(1), 0, stream, ctrl_in, (const Ctrl*)ctrl);
        LAUNCH_CHECK("inner reset");
        N.cap = (int)std::min<int64_t>(INT_MAX - 1, left_next);
        if ((rc = enqueue(N, flen))) return rc;
        next = true;
      }
    }
    HIP_TRY(hipEventSynchronize(ev[2]));
    const int ni = ctrl_h[3].final_iter > 0 ? ctrl_h[3].final_iter : S.it;
    dc_inner_guess = ni;
    inner_total += ni;
    inner_checks += ctrl_h[3].nchecks;
    if (ctrl_h[3].status != 0 && (dc_cfl_scale > 1.0 || dc_dtau_scale != 1.0 || dc_theta_scale != 1.0)) {
      // the sweeps diverged with the enlarged pseudo-time step: drop the
      // correction, fall back to the reference's step for this solver and
      // rebuild the coefficients of the current iterate
      if (dc_debug) std::fprintf(stderr, "[dc] correction sweeps diverged; pseudo-time scale 1 from here\n");
      dc_cfl_scale = 1.0;
      dc_Vpdtau = Vpdtau;
      dc_dtau_scale = dc_theta_scale = 1.0;
      dc_dtau_scale_dflt = dc_theta_scale_dflt = 1.0;
      if (fused) {
        // the pass above applied the dropped correction: its control block back
        ctrl_h[3] = outer_before;
        HIP_TRY(hipMemcpyAsync(ctrl, &ctrl_h[3], sizeof(Ctrl), hipMemcpyHostToDevice, stream));
      }
      if ((rc = dc_residual(dt, a, nullptr, 1 - a, outer + 2, o))) return rc;
      a = 1 - a;
      HIP_TRY(hipMemcpyAsync(&ctrl_h[2], ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
      continue;
    }
    ++outer;
    if (!fused) {
      int eb = S.ends.empty() ? 0 : S.ends.back().second;
      for (const auto& pr : S.ends)
        if (pr.first == ni) eb = pr.second;
      if ((rc = dc_residual(dt, a, u32[eb], 1 - a, outer + 1, o))) return rc;
      HIP_TRY(hipMemcpyAsync(&ctrl_h[2], ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
    }
    a = 1 - a;
    if (dc_debug)
      std::fprintf(stderr, "[dc] outer %d: inner %d (inner residual %.3e), outer residual %.3e\n", outer, ni,
                   ctrl_h[3].last_l2, ctrl_h[2].abs_);
    if (failed_now()) {
      dc_failed = true;
      break;
    }
    if (next) {
      // the next solve goes on from its first segment with the budget it has
      N.cap = (int)std::min<int64_t>(INT_MAX - 1, total_cap - inner_total);
      S = std::move(N);
      started = true;
    }
  }
  Ctrl& c = ctrl_h[2];
  if (!c.done) c.done = 1;
  st->rel_err = c.rel;
  st->abs_err = c.abs_;
  st->niter = (int)std::min<int64_t>(inner_total, INT_MAX);
  st->nchecks = c.nchecks + inner_checks;
  st->outer = outer;
  c.final_iter = st->niter;
  dc_final_buf = a;
  return CDIFF_OK;
}
