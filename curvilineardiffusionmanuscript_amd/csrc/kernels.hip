// kernels.hip -- the per-step and per-mesh kernels around the fused sweeps:
// metric coefficients, conductivity, validity flags, halo fill, convergence
// control, the epilogue (positivity, cutoff, next_dt) and diagnostic fields.
#include <limits>

#include "cdiff_dev.h"
#include "kernels.h"

namespace cdiff {

__device__ __forceinline__ void unflat(const Geo& g, int64_t e, int& i, int& j, int& k) {
  i = (int)(e % g.n[0]);
  int64_t r = e / g.n[0];
  j = (int)(r % g.n[1]);
  k = (int)(r / g.n[1]);
}

__device__ __forceinline__ bool in_domain(const Geo& g, int i, int j, int k) {
  bool d = i >= g.lo[0] && i <= g.hi[0] && j >= g.lo[1] && j <= g.hi[1];
  if (g.ndim == 3) d = d && k >= g.lo[2] && k <= g.hi[2];
  return d;
}

// ---------------------------------------------------------------------------
// Divergence coefficients and the metric cache, once per mesh (FP64).
// g_aa = |grad xi_a|^2, c_ab = 0.25 * grad xi_a . grad xi_b (the 0.25 of
// flux_divergence.jl:25,31; the 3D form 0.5*g*(0.5 X - 0.5 Y) of :154-211 is
// the same number), cache_a = mesh_metric_cache.jl:87-101 / :209-249.
__global__ void k_metric_coefs(Geo g, MetricIn m, double* __restrict__ out0, double* __restrict__ out1,
                               double* __restrict__ out2, double* __restrict__ out3, double* __restrict__ out4,
                               double* __restrict__ out5, double* __restrict__ out6, double* __restrict__ out7,
                               double* __restrict__ out8, const double* __restrict__ rz) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (e >= ntot) return;
  int i, j, k;
  unflat(g, e, i, j, k);
  const int64_t c = at(g, i, j, k);
  double* outs[9] = {out0, out1, out2, out3, out4, out5, out6, out7, out8};
  const int nd = g.ndim;
  if (!in_domain(g, i, j, k)) {
    for (int t = 0; t < (nd == 2 ? 5 : 9); ++t) outs[t][c] = 0.0;
    return;
  }
  double mm[3][3];
  for (int a = 0; a < nd; ++a)
    for (int cc = 0; cc < nd; ++cc) mm[a][cc] = m.cell[a * nd + cc][c];
  const int nm = nd * nd;
  double cache[3];
  for (int a = 0; a < nd; ++a) {
    double acc = 0.0;
    bool first = true;
    for (int b = 0; b < nd; ++b) {
      const int64_t st = b == 0 ? 1 : (b == 1 ? g.pitch : g.plane);
      const double* Jb = m.edge[b * (nm + 1) + nm];
      const double Jp = Jb[c], Jm = Jb[c - st];
      for (int cc = 0; cc < nd; ++cc) {
        const double* E = m.edge[b * (nm + 1) + a * nd + cc];
        const double vp = E[c] / Jp;
        const double vm = E[c - st] / Jm;
        const double term = mm[b][cc] * (vp - vm);
        if (first) {
          acc = term;
          first = false;
        } else {
          acc = acc + term;
        }
      }
    }
    cache[a] = acc;
  }
  if (nd == 2) {
    const double xix = mm[0][0], xiy = mm[0][1], etx = mm[1][0], ety = mm[1][1];
    out0[c] = xix * xix + xiy * xiy;
    out1[c] = etx * etx + ety * ety;
    out2[c] = 0.25 * (etx * xix + ety * xiy);
    // axisymmetric (x = r): (1/r) dT/dr enters like the metric-Laplacian terms,
    // cache_a + d(xi_a)/dr / r (not part of the reference APT)
    out3[c] = rz ? cache[0] + xix / rz[c] : cache[0];
    out4[c] = rz ? cache[1] + etx / rz[c] : cache[1];
  } else {
    const double xix = mm[0][0], xiy = mm[0][1], xiz = mm[0][2];
    const double etx = mm[1][0], ety = mm[1][1], etz = mm[1][2];
    const double zex = mm[2][0], zey = mm[2][1], zez = mm[2][2];
    out0[c] = xix * xix + xiy * xiy + xiz * xiz;
    out1[c] = etx * etx + ety * ety + etz * etz;
    out2[c] = zex * zex + zey * zey + zez * zez;
    out3[c] = 0.25 * (etx * xix + ety * xiy + etz * xiz);  // xi-eta
    out4[c] = 0.25 * (zex * xix + zey * xiy + zez * xiz);  // xi-zeta
    out5[c] = 0.25 * (zex * etx + zey * ety + zez * etz);  // eta-zeta
    out6[c] = cache[0];
    out7[c] = cache[1];
    out8[c] = cache[2];
  }
}

// ---------------------------------------------------------------------------
// alpha over the full array.  With `dom_src` set, domain cells take u from
// dom_src and the others from halo_src (the state of the last iteration:
// pre-update domain values with the ghosts that iteration applied).
__global__ void k_alpha_full(Geo g, KappaLaw kap, const double* __restrict__ rho, const double* __restrict__ dom_src,
                             const double* __restrict__ halo_src, double* __restrict__ alpha) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (e >= ntot) return;
  int i, j, k;
  unflat(g, e, i, j, k);
  const int64_t c = at(g, i, j, k);
  const double u = (dom_src && in_domain(g, i, j, k)) ? dom_src[c] : halo_src[c];
  alpha[c] = alpha_of<double>(kap, rho ? rho[c] : 1.0, u);
}

// dtau_rho and theta_r_dtau over the full array from alpha.
__global__ void k_iter_params_full(Geo g, IterConst ic, const double* __restrict__ rho, const double* __restrict__ alpha,
                                   double* __restrict__ dtau, double* __restrict__ theta) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (e >= ntot) return;
  int i, j, k;
  unflat(g, e, i, j, k);
  const int64_t c = at(g, i, j, k);
  double d, t;
  iter_params<double>(ic, rho ? rho[c] : 1.0, alpha[c], d, t);
  if (dtau) dtau[c] = d;
  if (theta) theta[c] = t;
}

template <typename T>
__global__ void k_convert(Geo g, const double* __restrict__ src, T* __restrict__ dst) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= g.total) return;
  dst[e] = (T)src[e];
}
template __global__ void k_convert<float>(Geo, const double*, float*);

__global__ void k_copy(int64_t n, const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < n) dst[e] = src[e];
}

// domain cells from dom_src, every other cell from rest
__global__ void k_merge(Geo g, const double* __restrict__ dom_src, const double* __restrict__ rest,
                        double* __restrict__ out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (e >= ntot) return;
  int i, j, k;
  unflat(g, e, i, j, k);
  const int64_t c = at(g, i, j, k);
  out[c] = in_domain(g, i, j, k) ? dom_src[c] : rest[c];
}

__global__ void k_fill(int64_t n, double v, double* __restrict__ dst) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < n) dst[e] = v;
}

// ---------------------------------------------------------------------------
// validate_scalar (validity_checks.jl:1-42): flags of non-finite (or negative
// with positivity) values over the domain (bit 0) and the one-layer halo
// planes of expand(domain, 1) (bit 1 + 2*axis + side), OR-ed into ctrl.
__global__ void k_validate(Geo g, const double* __restrict__ a, int positivity, int shift, Ctrl* ctrl) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  int flags = 0;
  if (e < ntot) {
    int i, j, k;
    unflat(g, e, i, j, k);
    int idx[3] = {i, j, k};
    bool inside = true;  // within expand(domain, 1)
    for (int ax = 0; ax < g.ndim; ++ax) inside = inside && idx[ax] >= g.lo[ax] - 1 && idx[ax] <= g.hi[ax] + 1;
    if (inside) {
      const double v = a[at(g, i, j, k)];
      const bool bad = !isfinite(v) || (positivity && !(v >= 0.0));
      if (bad) {
        if (in_domain(g, i, j, k)) flags |= 1;
        for (int ax = 0; ax < g.ndim; ++ax) {
          if (idx[ax] == g.lo[ax] - 1) flags |= 1 << (1 + 2 * ax);
          if (idx[ax] == g.hi[ax] + 1) flags |= 1 << (2 + 2 * ax);
        }
      }
    }
  }
  // wave-level OR, one atomic per wave
  for (int o = WAVE / 2; o > 0; o >>= 1) flags |= __shfl_xor(flags, o, WAVE);
  if ((threadIdx.x & (WAVE - 1)) == 0 && flags) atomicOr(&ctrl->valid_flags, flags << shift);
}

// Step-start control block.
__global__ void k_ctrl_reset(Ctrl* ctrl, double rtol, double atol) {
  ctrl->done = 0;
  ctrl->status = 0;
  ctrl->final_iter = 0;
  ctrl->nchecks = 0;
  ctrl->rel = 2.0 * rtol;
  ctrl->abs_ = 2.0 * atol;
  ctrl->init = INFINITY;
  ctrl->last_l2 = 0.0;
  ctrl->valid_flags = 0;
  ctrl->grow = 0;
  ctrl->ref_iter = 0;
  ctrl->ref = INFINITY;
}

__global__ void k_ctrl_reset_follow(Ctrl* ctrl, const Ctrl* follow) {
  ctrl->done = (follow->done || follow->status != 0) ? 1 : 0;
  ctrl->status = 0;
  ctrl->final_iter = 0;
  ctrl->nchecks = 0;
  ctrl->rel = 0.0;
  ctrl->abs_ = 0.0;
  ctrl->init = INFINITY;
  ctrl->last_l2 = 0.0;
  ctrl->valid_flags = 0;
  ctrl->grow = 0;
  ctrl->ref_iter = 0;
  ctrl->ref = INFINITY;
}

__global__ void k_ctrl_validated(Ctrl* ctrl) {
  const int f = ctrl->valid_flags;
  if (f & 0xffff) {
    ctrl->done = 1;
    ctrl->status = 3;  // CDIFF_INVALID_ALPHA
  } else if (f >> 16) {
    ctrl->done = 1;
    ctrl->status = 4;  // CDIFF_INVALID_SOURCE
  }
}

// ---------------------------------------------------------------------------
// One error check (PseudoTransientSolver.jl:280-318): reduce the per-wave
// partials in a fixed order, update rel/abs, and decide the exit in the
// reference's order: non-finite, then iter > max_iter, then tolerance.
// `check` = 0 is the max_iter-only test of a non-check iteration.
__global__ __launch_bounds__(1024) void k_finalize(Ctrl* ctrl, const double* __restrict__ psum,
                                                   const double* __restrict__ pmax, int nparts, int iter, int max_iter,
                                                   double rtol, double atol, double nfull, int norm_kind, int check) {
  if (ctrl->done) return;
  constexpr int NT = 1024, U = 8;
  __shared__ double ss[NT], sm[NT];
  const int t = threadIdx.x;
  double s = 0.0, m = 0.0;
  if (check) {
    // U independent loads in flight per thread; fixed summation order
    double as[U], am[U];
#pragma unroll
    for (int u = 0; u < U; ++u) as[u] = am[u] = 0.0;
    for (int base = 0; base < nparts; base += NT * U) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int p = base + u * NT + t;
        if (p < nparts) {
          as[u] += psum[p];
          am[u] = nanmax(am[u], pmax[p]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      s += as[u];
      m = nanmax(m, am[u]);
    }
  }
  ss[t] = s;
  sm[t] = m;
  __syncthreads();
  for (int o = NT / 2; o > 0; o >>= 1) {
    if (t < o) {
      ss[t] += ss[t + o];
      sm[t] = nanmax(sm[t], sm[t + o]);
    }
    __syncthreads();
  }
  if (t != 0) return;
  double rel = ctrl->rel, ab = ctrl->abs_;
  if (check) {
    const double L2 = norm_kind == 0 ? sqrt(ss[0] / nfull) : sm[0] / sqrt(nfull);
    if (iter == 1) ctrl->init = L2;
    rel = L2 / ctrl->init;
    ab = L2;
    ctrl->rel = rel;
    ctrl->abs_ = ab;
    ctrl->last_l2 = L2;
    ctrl->nchecks += 1;
  }
  int st = -1;
  if (!isfinite(rel) || !isfinite(ab))
    st = 1;  // CDIFF_NONFINITE
  else if (iter > max_iter)
    st = 2;  // CDIFF_MAX_ITER
  else if (rel <= rtol || ab <= atol)
    st = 0;
  if (st >= 0) {
    ctrl->status = st;
    ctrl->final_iter = iter;
    ctrl->done = 1;
  }
}

// ---------------------------------------------------------------------------
// Ghost fill of a stored array (applybcs!, boundary_operators.jl:26-263) for
// the first halo layer, from the per-face rules; `outer` also sets the 2D
// Dirichlet layers beyond the first (boundary_operators.jl:144-150).
__global__ void k_halo_fill(Geo g, Ghost gh, double* __restrict__ u, int outer) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (e >= ntot) return;
  int i, j, k;
  unflat(g, e, i, j, k);
  int idx[3] = {i, j, k};
  int gax = -1, nout = 0;
  for (int ax = 0; ax < g.ndim; ++ax) {
    if (idx[ax] < g.lo[ax] || idx[ax] > g.hi[ax]) {
      gax = ax;
      nout++;
    }
  }
  if (nout != 1) return;  // domain cell or corner/edge region: never written
  const int ax = gax;
  const int side = idx[ax] < g.lo[ax] ? 0 : 1;
  const int f = 2 * ax + side;
  const int layer = side == 0 ? g.lo[ax] - idx[ax] : idx[ax] - g.hi[ax];
  const int m = gh.mode[f];
  if (layer > 1) {
    // only 2D Dirichlet writes layers beyond the first
    if (outer && g.ndim == 2 && m == G_CONST) u[at(g, i, j, k)] = gh.val[f];
    return;
  }
  if (outer) return;
  if (m == G_KEEP) return;
  if (m == G_CONST) {
    u[at(g, i, j, k)] = gh.val[f];
    return;
  }
  int s[3] = {i, j, k};
  s[ax] = gh.src[f];
  const double v = u[at(g, s[0], s[1], s[2])];
  u[at(g, i, j, k)] = m == G_NEG ? -v : v;
}

// ---------------------------------------------------------------------------
// Epilogue (PseudoTransientSolver.jl:323-346): T = cutoff(|u|) over the full
// array; domain values from `dom`, everything else from `rest`.
__global__ void k_epilogue_u(Geo g, const double* __restrict__ u, double* __restrict__ T, int positivity, int cutoff) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (e >= ntot) return;
  int i, j, k;
  unflat(g, e, i, j, k);
  const int64_t c = at(g, i, j, k);
  double a = u[c];
  if (positivity) a = fabs(a);
  if (cutoff) {
    const double v = 0.5 * (fabs(a) + a);
    a = isfinite(a) ? v : copysign(0.0, v);
  }
  T[c] = a;
}

// next_dt (timestep_control.jl:33-65): pass 0 = per-block max of u_prev
// (NaN-propagating, as Julia maximum); pass 1 = per-block max of the strong-
// zeroed relative change with u0 = umax * tol from the reduced pass-0 value.
__global__ __launch_bounds__(256) void k_next_dt_partial(Geo g, const double* __restrict__ unew,
                                                         const double* __restrict__ uold, const double* __restrict__ red,
                                                         double tol, int pass, double* __restrict__ part) {
  const int64_t ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  double m = -INFINITY;
  const double u0 = pass ? red[0] * tol : 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ntot; e += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    unflat(g, e, i, j, k);
    const int64_t c = at(g, i, j, k);
    double v;
    if (pass == 0) {
      v = uold[c];
    } else {
      const double a = unew[c];
      v = fabs(a - uold[c]) / (a + u0);
      v = strong_zero(v);
    }
    m = nanmax(m, v);
  }
  __shared__ double sm[256];
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] = nanmax(sm[threadIdx.x], sm[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ __launch_bounds__(256) void k_reduce_max(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double sm[256];
  double m = -INFINITY;
  for (int p = threadIdx.x; p < n; p += 256) m = nanmax(m, part[p]);
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] = nanmax(sm[threadIdx.x], sm[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sm[0];
}

__global__ void k_next_dt_final(const double* __restrict__ mx, double dt, double muc, double growth,
                                double* __restrict__ out) {
  const double dtn = dt * sqrt(fabs(muc / mx[0]));
  const double gdt = growth * dt;
  double r;
  if (dtn != dtn)
    r = dtn;
  else if (gdt != gdt)
    r = gdt;
  else
    r = dtn < gdt ? dtn : gdt;
  out[0] = r;
}

// ---------------------------------------------------------------------------
// Diagnostic fields (q' and res of the final state), unfused.
__global__ void k_plain_flux(Geo g, int axis, int harmonic, const double* __restrict__ u, const double* __restrict__ alpha,
                             double* __restrict__ qp) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (e >= ntot) return;
  int i, j, k;
  unflat(g, e, i, j, k);
  int idx[3] = {i, j, k};
  bool comp = true;
  for (int ax = 0; ax < g.ndim; ++ax) {
    const int lo = g.lo[ax] - (ax == axis ? 1 : 0);
    comp = comp && idx[ax] >= lo && idx[ax] <= g.hi[ax];
  }
  const int64_t c = at(g, i, j, k);
  if (!comp) {
    qp[c] = 0.0;
    return;
  }
  const int64_t st = axis == 0 ? 1 : (axis == 1 ? g.pitch : g.plane);
  const double abar = face_mean<double>(harmonic, alpha[c], alpha[c + st]);
  qp[c] = -abar * (u[c + st] - u[c]);
}

// res = -mask*(u - u_prev)/dt - div(q') + S (+ f) on the domain, 0 elsewhere
// (residuals.jl:24-47 with flux_divergence.jl:7-42 / :119-239).
__global__ void k_residual_field(Geo g, double dt, const double* __restrict__ u, const double* __restrict__ uprev,
                                 const double* __restrict__ q0, const double* __restrict__ q1,
                                 const double* __restrict__ q2, Coefs cf, const double* __restrict__ S,
                                 const double* __restrict__ tzA, const double* __restrict__ tzB,
                                 const double* __restrict__ alpha, double* __restrict__ res) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (e >= ntot) return;
  int i, j, k;
  unflat(g, e, i, j, k);
  const int64_t c = at(g, i, j, k);
  if (!in_domain(g, i, j, k)) {
    res[c] = 0.0;
    return;
  }
  const int64_t sI = 1, sJ = g.pitch, sK = g.plane;
  double D;
  if (g.ndim == 2) {
    const double t1 = cf.c[0][c] * (q0[c] - q0[c - sI]);
    const double t2 = cf.c[1][c] * (q1[c] - q1[c - sJ]);
    const double t3 = cf.c[2][c] * ((q0[c + sJ] + q0[c - sI + sJ]) - (q0[c - sJ] + q0[c - sI - sJ]));
    const double t4 = cf.c[2][c] * ((q1[c + sI] + q1[c + sI - sJ]) - (q1[c - sI] + q1[c - sI - sJ]));
    const double t5 = cf.c[3][c] * (0.5 * (q0[c] + q0[c - sI]));
    const double t6 = cf.c[4][c] * (0.5 * (q1[c] + q1[c - sJ]));
    D = t1 + t2 + t3 + t4 + t5 + t6;
  } else {
    const double a1 = cf.c[0][c] * (q0[c] - q0[c - sI]);
    const double a2 = cf.c[1][c] * (q1[c] - q1[c - sJ]);
    const double a3 = cf.c[2][c] * (q2[c] - q2[c - sK]);
    const double xe = cf.c[3][c], xz = cf.c[4][c], ez = cf.c[5][c];
    const double qi_eta = xe * ((q0[c + sJ] + q0[c - sI + sJ]) - (q0[c - sJ] + q0[c - sI - sJ]));
    const double qi_zeta = xz * ((q0[c + sK] + q0[c - sI + sK]) - (q0[c - sK] + q0[c - sI - sK]));
    const double qj_xi = xe * ((q1[c + sI] + q1[c + sI - sJ]) - (q1[c - sI] + q1[c - sI - sJ]));
    const double qj_zeta = ez * ((q1[c + sK] + q1[c - sJ + sK]) - (q1[c - sK] + q1[c - sJ - sK]));
    const double qk_eta = ez * ((q2[c + sJ] + q2[c + sJ - sK]) - (q2[c - sJ] + q2[c - sJ - sK]));
    const double qk_xi = xz * ((q2[c + sI] + q2[c + sI - sK]) - (q2[c - sI] + q2[c - sI - sK]));
    const double m1 = cf.c[6][c] * (0.5 * (q0[c] + q0[c - sI]));
    const double m2 = cf.c[7][c] * (0.5 * (q1[c] + q1[c - sJ]));
    const double m3 = cf.c[8][c] * (0.5 * (q2[c] + q2[c - sK]));
    D = a1 + a2 + a3 + qi_eta + qi_zeta + qj_xi + qj_zeta + qk_eta + qk_xi + m1 + m2 + m3;
  }
  const double un = u[c], um = uprev[c];
  double du = un - um;
  bool approx = un == um;
  if (!approx && isfinite(un) && isfinite(um))
    approx = fabs(un - um) <= std::numeric_limits<double>::epsilon() * fmax(fabs(un), fabs(um));
  if (approx) du = copysign(0.0, du);
  double r = -du / dt - D + (S ? S[c] : 0.0);
  if (tzA) r = r + (tzA[c] + tzB[c] * alpha[c]);
  res[c] = r;
}

// Cell Jacobian of a 2D mesh from the cell metrics: 1 / (xi_x eta_y - xi_y eta_x).
__global__ void k_cell_jacobian2d(Geo g, const double* __restrict__ xx, const double* __restrict__ xy,
                                  const double* __restrict__ ex, const double* __restrict__ ey,
                                  double* __restrict__ jac) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1];
  if (e >= ntot) return;
  const int i = (int)(e % g.n[0]), j = (int)(e / g.n[0]);
  const int64_t c = at(g, i, j, 0);
  const double det = xx[c] * ey[c] - xy[c] * ex[c];
  jac[c] = det != 0.0 ? 1.0 / det : 0.0;
}

}  // namespace cdiff
