// dc_kernels.hip -- FP64 kernels of the mixed-precision defect-correction step.
//
// The outer loop keeps the iterate u in FP64 and evaluates the reference's
// residual r(u) = -mask*(u - u_prev)/dt - div(q'(u)) + S (+ f)
// (src/pseudo_transient/residuals.jl:24-47, flux_divergence.jl) in FP64; the
// inner loop runs FP32 APT sweeps on the correction equation
//   e/dt + div(q'_alpha(e)) = r,  alpha frozen at u,
// and the outer loop applies u <- u + e.  These are the FP64 pieces:
//   k_dc_update    u_new = u_old + e on the domain, ghosts = BCs of u_new,
//                  alpha(u_new) in FP64 (scratch) and FP32 (inner sweeps)
//   k_dc_residual  r(u_new) per domain cell, stored in FP32 for the inner
//                  solve, with per-block partials of sum(r^2) and max|r|
#include <limits>

#include "cdiff_dev.h"
#include "dc.h"
#include "lin_coef_ops.h"

namespace cdiff {

namespace {

__device__ __forceinline__ void unflat3(const Geo& g, int64_t e, int& i, int& j, int& k) {
  i = (int)(e % g.n[0]);
  const int64_t r = e / g.n[0];
  j = (int)(r % g.n[1]);
  k = (int)(r / g.n[1]);
}

}  // namespace

// One thread per cell of the full array.  Domain cells: u_new = u_old + e.
// First-layer ghost cells on a face: the BC rule applied to u_new (reading
// u_old + e of the source cell).  Everything else is copied.
__global__ void k_dc_update(Geo g, Ghost gh, KappaLaw kap, const double* __restrict__ rho,
                            const double* __restrict__ u_old, const float* __restrict__ e, double* __restrict__ u_new,
                            double* __restrict__ a64, float* __restrict__ a32, float* __restrict__ ub32,
                            float* __restrict__ da32) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (t >= ntot) return;
  int i, j, k;
  unflat3(g, t, i, j, k);
  const int idx[3] = {i, j, k};
  int nout = 0, ax = 0, layer = 0;
  for (int d = 0; d < g.ndim; ++d) {
    if (idx[d] < g.lo[d]) {
      ++nout;
      ax = d;
      layer = g.lo[d] - idx[d];
    } else if (idx[d] > g.hi[d]) {
      ++nout;
      ax = d;
      layer = idx[d] - g.hi[d];
    }
  }
  const int64_t c = at(g, i, j, k);
  double v;
  if (nout == 0) {
    v = u_old[c] + (e ? (double)e[c] : 0.0);
  } else if (nout == 1 && layer == 1) {
    const int f = 2 * ax + (idx[ax] < g.lo[ax] ? 0 : 1);
    const int m = gh.mode[f];
    if (m == G_CONST) {
      v = gh.val[f];
    } else if (m == G_KEEP) {
      v = u_old[c];
    } else {
      int s[3] = {i, j, k};
      s[ax] = gh.src[f];
      const int64_t sc = at(g, s[0], s[1], s[2]);
      const double us = u_old[sc] + (e ? (double)e[sc] : 0.0);
      v = m == G_NEG ? -us : us;
    }
  } else {
    v = u_old[c];
  }
  u_new[c] = v;
  const double a = alpha_of<double>(kap, rho ? rho[c] : 1.0, v);
  a64[c] = a;
  a32[c] = (float)a;
  if (ub32) {
    // d(alpha)/du of kappa0 |rho|^a |u|^b / (rho cp) is b * alpha / u
    ub32[c] = (float)v;
    const double d = (v != 0.0 && isfinite(v)) ? kap.t_exp * a / v : 0.0;
    da32[c] = (float)(isfinite(d) ? d : 0.0);
  }
}

// Plain flux q' of the face a+1/2 of cell (i, j, k) along `axis`, zero outside
// expand_lower(domain, axis) (fluxes_cpu.jl:40-41).
__device__ __forceinline__ double qp_face(const Geo& g, int harmonic, const double* __restrict__ u,
                                          const double* __restrict__ a, int axis, int i, int j, int k) {
  const int idx[3] = {i, j, k};
  for (int d = 0; d < g.ndim; ++d) {
    const int lo = g.lo[d] - (d == axis ? 1 : 0);
    if (idx[d] < lo || idx[d] > g.hi[d]) return 0.0;
  }
  const int64_t c = at(g, i, j, k);
  const int64_t st = axis == 0 ? 1 : (axis == 1 ? g.pitch : g.plane);
  const double abar = face_mean<double>(harmonic, a[c], a[c + st]);
  return -abar * (u[c + st] - u[c]);
}

__global__ __launch_bounds__(256) void k_dc_residual(Geo g, int harmonic, double dt, const double* __restrict__ u,
                                                     const double* __restrict__ a, const double* __restrict__ uprev,
                                                     DcCoefs cf, const double* __restrict__ S,
                                                     const double* __restrict__ tzA, const double* __restrict__ tzB,
                                                     float* __restrict__ r32, double* __restrict__ psum,
                                                     double* __restrict__ pmax) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  double sq = 0.0, mx = 0.0;
  if (t < ntot) {
    int i, j, k;
    unflat3(g, t, i, j, k);
    const int64_t c = at(g, i, j, k);
    bool dom = i >= g.lo[0] && i <= g.hi[0] && j >= g.lo[1] && j <= g.hi[1];
    if (g.ndim == 3) dom = dom && k >= g.lo[2] && k <= g.hi[2];
    if (!dom) {
      r32[c] = 0.0f;
    } else {
      double D;
      if (g.ndim == 2) {
        const double xc = qp_face(g, harmonic, u, a, 0, i, j, 0), xw = qp_face(g, harmonic, u, a, 0, i - 1, j, 0);
        const double xn = qp_face(g, harmonic, u, a, 0, i, j + 1, 0);
        const double xnw = qp_face(g, harmonic, u, a, 0, i - 1, j + 1, 0);
        const double xs = qp_face(g, harmonic, u, a, 0, i, j - 1, 0);
        const double xsw = qp_face(g, harmonic, u, a, 0, i - 1, j - 1, 0);
        const double yc = qp_face(g, harmonic, u, a, 1, i, j, 0), ys = qp_face(g, harmonic, u, a, 1, i, j - 1, 0);
        const double ye = qp_face(g, harmonic, u, a, 1, i + 1, j, 0);
        const double yes = qp_face(g, harmonic, u, a, 1, i + 1, j - 1, 0);
        const double yw = qp_face(g, harmonic, u, a, 1, i - 1, j, 0);
        const double yws = qp_face(g, harmonic, u, a, 1, i - 1, j - 1, 0);
        const double t1 = cf.c[0][c] * (xc - xw);
        const double t2 = cf.c[1][c] * (yc - ys);
        const double t3 = cf.c[2][c] * ((xn + xnw) - (xs + xsw));
        const double t4 = cf.c[2][c] * ((ye + yes) - (yw + yws));
        const double t5 = cf.c[3][c] * (0.5 * (xc + xw));
        const double t6 = cf.c[4][c] * (0.5 * (yc + ys));
        D = t1 + t2 + t3 + t4 + t5 + t6;
      } else {
        auto X = [&](int di, int dj, int dk) { return qp_face(g, harmonic, u, a, 0, i + di, j + dj, k + dk); };
        auto Y = [&](int di, int dj, int dk) { return qp_face(g, harmonic, u, a, 1, i + di, j + dj, k + dk); };
        auto Z = [&](int di, int dj, int dk) { return qp_face(g, harmonic, u, a, 2, i + di, j + dj, k + dk); };
        const double xc = X(0, 0, 0), xw = X(-1, 0, 0), yc = Y(0, 0, 0), ys = Y(0, -1, 0);
        const double zc = Z(0, 0, 0), zs = Z(0, 0, -1);
        const double a1 = cf.c[0][c] * (xc - xw);
        const double a2 = cf.c[1][c] * (yc - ys);
        const double a3 = cf.c[2][c] * (zc - zs);
        const double qi_eta = cf.c[3][c] * ((X(0, 1, 0) + X(-1, 1, 0)) - (X(0, -1, 0) + X(-1, -1, 0)));
        const double qi_zeta = cf.c[4][c] * ((X(0, 0, 1) + X(-1, 0, 1)) - (X(0, 0, -1) + X(-1, 0, -1)));
        const double qj_xi = cf.c[3][c] * ((Y(1, 0, 0) + Y(1, -1, 0)) - (Y(-1, 0, 0) + Y(-1, -1, 0)));
        const double qj_zeta = cf.c[5][c] * ((Y(0, 0, 1) + Y(0, -1, 1)) - (Y(0, 0, -1) + Y(0, -1, -1)));
        const double qk_eta = cf.c[5][c] * ((Z(0, 1, 0) + Z(0, 1, -1)) - (Z(0, -1, 0) + Z(0, -1, -1)));
        const double qk_xi = cf.c[4][c] * ((Z(1, 0, 0) + Z(1, 0, -1)) - (Z(-1, 0, 0) + Z(-1, 0, -1)));
        const double m1 = cf.c[6][c] * (0.5 * (xc + xw));
        const double m2 = cf.c[7][c] * (0.5 * (yc + ys));
        const double m3 = cf.c[8][c] * (0.5 * (zc + zs));
        D = a1 + a2 + a3 + qi_eta + qi_zeta + qj_xi + qj_zeta + qk_eta + qk_xi + m1 + m2 + m3;
      }
      const double un = u[c], um = uprev[c];
      double du = un - um;
      bool approx = un == um;
      if (!approx && isfinite(un) && isfinite(um))
        approx = fabs(un - um) <= std::numeric_limits<double>::epsilon() * fmax(fabs(un), fabs(um));
      if (approx) du = copysign(0.0, du);
      double r = -du / dt - D + (S ? S[c] : 0.0);
      if (tzA) r = r + (tzA[c] + tzB[c] * a[c]);
      r32[c] = (float)r;
      sq = r * r;
      mx = fabs(r);
      if (r != r) mx = r;
    }
  }
  __shared__ double ss[256], sm[256];
  ss[threadIdx.x] = sq;
  sm[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      ss[threadIdx.x] += ss[threadIdx.x + o];
      sm[threadIdx.x] = nanmax(sm[threadIdx.x], sm[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    psum[blockIdx.x] = ss[0];
    pmax[blockIdx.x] = sm[0];
  }
}

// Inner-loop decision: stop the FP32 sweeps once the correction residual has
// dropped below eta * (the outer FP64 residual), or after `cap` sweeps.  The
// inner residual is measured in the outer residual's norm (norm_kind 0: RMS,
// 1: max / sqrt(N)), so that the two are compared like for like.
__global__ __launch_bounds__(1024) void k_inner_finalize(Ctrl* cin, const Ctrl* cout, const double* __restrict__ psum,
                                                         const double* __restrict__ pmax, int nparts, double nfull,
                                                         int norm_kind, double eta, int iter, int cap,
                                                         double floor_frac, double rtol, double atol,
                                                         int out_buf) {
  if (cin->done) return;
  constexpr int NT = 1024, U = 8;
  __shared__ double ss[NT], sm[NT];
  double as[U], am[U];
#pragma unroll
  for (int u = 0; u < U; ++u) as[u] = am[u] = 0.0;
  for (int base = 0; base < nparts; base += NT * U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int p = base + u * NT + threadIdx.x;
      if (p < nparts) {
        as[u] += psum[p];
        am[u] = nanmax(am[u], pmax[p]);
      }
    }
  }
  double s = 0.0, m = 0.0;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    s += as[u];
    m = nanmax(m, am[u]);
  }
  ss[threadIdx.x] = s;
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int o = NT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      ss[threadIdx.x] += ss[threadIdx.x + o];
      sm[threadIdx.x] = nanmax(sm[threadIdx.x], sm[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double L2 = norm_kind == 0 ? sqrt(ss[0] / nfull) : sm[0] / sqrt(nfull);
  const double prev = cin->last_l2;
  cin->last_l2 = L2;
  cin->nchecks += 1;
  cin->final_buf = out_buf;
  const double prev_best = cin->init;  // +inf after k_ctrl_reset
  const double best = fmin(prev_best, L2);
  cin->init = best;
  cin->grow = (cin->nchecks > 1 && L2 > prev) ? cin->grow + 1 : 0;
  // stop at eta times the outer residual, or once the outer tolerance is in
  // reach: a fraction of max(atol, rtol * initial residual)
  double target = eta * cout->last_l2;
  if (floor_frac > 0.0) target = fmax(target, floor_frac * fmax(atol, rtol * cout->init));
  // the sweeps are unstable for these pseudo-time parameters: past the first
  // 24 sweeps, a residual 100x above the smallest one so far, or growth at 3
  // checks in a row to 1.5x it (a slow instability: measured on a 4096^2 blob
  // problem, the residual fell to 6.7e-5 and then grew 1.1x per 12 sweeps).
  // Not measured against the starting residual: the pseudo-transient's first
  // sweeps overshoot it by up to ~1e3 (4096^2 Gaussian, kappa = 1: 355 ->
  // 1.5e5 after 24 sweeps) and then fall
  const bool diverged = iter >= 24 && (L2 > 1e2 * prev_best || (cin->grow >= 3 && L2 > 1.5 * best));
  // the sweeps have stalled: the residual has not halved in the last
  // DC_STALL_SWEEPS sweeps.  The FP32 rounding of the correction equation
  // sets a floor under its residual (on the 4096^2 Gaussian with kappa = 1,
  // 1.1e-3 of the outer residual, above eta); the correction reached so far
  // goes to the outer FP64 loop, which decides convergence on its own residual
  if (L2 <= 0.5 * cin->ref) {
    cin->ref = L2;
    cin->ref_iter = iter;
  }
  const bool stalled = iter - cin->ref_iter >= DC_STALL_SWEEPS;
  if (!(L2 <= 1e300) || diverged || stalled || L2 <= target || iter >= cap) {
    cin->done = 1;
    cin->final_iter = iter;
    cin->status = !(L2 == L2) ? 1 : (diverged || !(L2 <= 1e300) ? 3 : 0);
  }
}

// The denominator of the relative criterion from another control block's
// first check (its init), with k_finalize's exit test re-evaluated: the
// reference divides by the residual after its first pseudo-iteration
// (PseudoTransientSolver.jl:280-293).
__global__ void k_ctrl_set_init(Ctrl* ctrl, const Ctrl* src, double rtol, double atol) {
  if (ctrl->done) return;
  const double init = src->init;
  const double ab = ctrl->abs_;
  const double rel = ab / init;
  ctrl->init = init;
  ctrl->rel = rel;
  int st = -1;
  if (!isfinite(rel) || !isfinite(ab))
    st = 1;  // CDIFF_NONFINITE
  else if (rel <= rtol || ab <= atol)
    st = 0;
  if (st >= 0) {
    ctrl->status = st;
    ctrl->final_iter = 1;
    ctrl->done = 1;
  }
}

__global__ void k_fill32(int64_t n, float v, float* __restrict__ dst) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) dst[t] = v;
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ src, float* __restrict__ dst) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) dst[t] = (float)src[t];
}

// ---------------------------------------------------------------------------
// Coefficients of the FP32 correction sweep in coefficient form (2D).

// Divergence coefficients folded per cell (constant per mesh):
// D = m0 x_c + m1 x_w + m2 y_c + m3 y_s + m4 (cross terms).
__global__ void k_lin_metrics(Geo g, const double* __restrict__ g11, const double* __restrict__ g22,
                              const double* __restrict__ c12, const double* __restrict__ ac,
                              const double* __restrict__ bc, float* __restrict__ m0, float* __restrict__ m1,
                              float* __restrict__ m2, float* __restrict__ m3, float* __restrict__ m4) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1];
  if (t >= ntot) return;
  const int i = (int)(t % g.n[0]), j = (int)(t / g.n[0]);
  const int64_t c = at(g, i, j, 0);
  const bool dom = i >= g.lo[0] && i <= g.hi[0] && j >= g.lo[1] && j <= g.hi[1];
  if (!dom) {
    m0[c] = m1[c] = m2[c] = m3[c] = m4[c] = 0.f;
    return;
  }
  const double ha = 0.5 * ac[c], hb = 0.5 * bc[c];
  m0[c] = (float)(g11[c] + ha);
  m1[c] = (float)(ha - g11[c]);
  m2[c] = (float)(g22[c] + hb);
  m3[c] = (float)(hb - g22[c]);
  m4[c] = (float)c12[c];
}

namespace {

// factor s of the correction in the first ghost layer of face f: e_ghost = s e_source
__device__ __forceinline__ double ghost_factor(const Ghost& gh, int f) {
  const int m = gh.mode[f];
  return m == G_COPY ? 1.0 : (m == G_NEG ? -1.0 : 0.0);
}

}  // namespace

// Per face of expand_lower(domain, axis), from the FP64 iterate u and its
// alpha a (full arrays, ghosts by the BC rules): P, M, W of
//   q' = P e0 - M e1,  q <- q' + W (q - q')
// the linearisation of q'(u + e) - q'(u) in e, with the Newton term of the
// face mean when `newton`; per domain cell C = dtau / (1 + dtau/dt).
// Ghost cells of boundary faces are folded: e_ghost = s e_domain.
__global__ void k_lin_coefs(Geo g, Ghost gh, int harmonic, IterConst ic, double t_exp, int newton,
                            const double* __restrict__ rho, const double* __restrict__ u,
                            const double* __restrict__ a, LinOut out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (t >= ntot) return;
  int idx[3];
  unflat3(g, t, idx[0], idx[1], idx[2]);
  const int64_t c = at(g, idx[0], idx[1], idx[2]);
  bool dom = true, near = true;
  for (int d = 0; d < g.ndim; ++d) {
    dom = dom && idx[d] >= g.lo[d] && idx[d] <= g.hi[d];
    near = near && idx[d] >= g.lo[d] - 1 && idx[d] <= g.hi[d] + 1;
  }
  // the sweep's coefficients in FP32 (lin_coef_ops.h), as the fused outer passes form them
  const IterConst32 k32 = iter_const32(ic, out.dtau_scale, out.theta_scale);
  auto params = [&](int64_t cell, float& C, float& theta) {
    iter_params32(k32, (float)(rho ? rho[cell] : 1.0), (float)a[cell], C, theta);
  };
  float C0 = 0.f, th0 = 0.f;
  if (near) params(c, C0, th0);
  out.cc[c] = dom ? C0 : 0.f;
  for (int ax = 0; ax < g.ndim; ++ax) {
    float* const* F = ax == 0 ? out.fx : (ax == 1 ? out.fy : out.fz);
    bool on = true;
    for (int d = 0; d < g.ndim; ++d) {
      const int lo = g.lo[d] - (d == ax ? 1 : 0);
      on = on && idx[d] >= lo && idx[d] <= g.hi[d];
    }
    if (!on) {
      F[0][c] = F[1][c] = F[2][c] = 0.f;
      continue;
    }
    const int64_t c1 = c + (ax == 0 ? (int64_t)1 : (ax == 1 ? g.pitch : g.plane));
    float C1, th1;
    params(c1, C1, th1);
    const double a0 = a[c], a1 = a[c1];
    float P = (float)face_mean<double>(harmonic, a0, a1), M = P;
    if (newton) {
      float d0, d1;  // derivatives of the face mean in a0, a1
      mean_weights32(harmonic, (float)a0, (float)a1, d0, d1);
      const float du = (float)(u[c1] - u[c]);
      P = __builtin_fmaf(-d0 * dadu32((float)t_exp, (float)a0, (float)u[c]), du, P);
      M = __builtin_fmaf(d1 * dadu32((float)t_exp, (float)a1, (float)u[c1]), du, M);
    }
    // boundary faces: cell 0 below the domain, or cell 1 above it
    if (idx[ax] == g.lo[ax] - 1) {
      M = __builtin_fmaf(-(float)ghost_factor(gh, 2 * ax), P, M);
      P = 0.f;
    } else if (idx[ax] == g.hi[ax]) {
      P = __builtin_fmaf(-(float)ghost_factor(gh, 2 * ax + 1), M, P);
      M = 0.f;
    }
    F[0][c] = P;
    F[1][c] = M;
    F[2][c] = damping32(harmonic, th0, th1);
  }
}

// Divergence coefficients of the 3D correction sweep, folded per cell.
__global__ void k_lin_metrics3d(Geo g, DcCoefs cf, float* __restrict__ m0, float* __restrict__ m1,
                                float* __restrict__ m2, float* __restrict__ m3, float* __restrict__ m4,
                                float* __restrict__ m5, float* __restrict__ m6, float* __restrict__ m7,
                                float* __restrict__ m8) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (t >= ntot) return;
  int i, j, k;
  unflat3(g, t, i, j, k);
  const int64_t c = at(g, i, j, k);
  const bool dom = i >= g.lo[0] && i <= g.hi[0] && j >= g.lo[1] && j <= g.hi[1] && k >= g.lo[2] && k <= g.hi[2];
  if (!dom) {
    m0[c] = m1[c] = m2[c] = m3[c] = m4[c] = m5[c] = m6[c] = m7[c] = m8[c] = 0.f;
    return;
  }
  const double ha = 0.5 * cf.c[6][c], hb = 0.5 * cf.c[7][c], hg = 0.5 * cf.c[8][c];
  m0[c] = (float)(cf.c[0][c] + ha);
  m1[c] = (float)(ha - cf.c[0][c]);
  m2[c] = (float)(cf.c[1][c] + hb);
  m3[c] = (float)(hb - cf.c[1][c]);
  m4[c] = (float)(cf.c[2][c] + hg);
  m5[c] = (float)(hg - cf.c[2][c]);
  m6[c] = (float)cf.c[3][c];
  m7[c] = (float)cf.c[4][c];
  m8[c] = (float)cf.c[5][c];
}

}  // namespace cdiff
