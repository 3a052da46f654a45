// sweep2d.hip -- one fused APT pseudo-iteration on a 2D curvilinear mesh.
//
// One launch performs, for every domain cell, what the reference does in five
// passes (src/pseudo_transient/PseudoTransientSolver.jl:254-281):
//   applybcs!(u)                  boundary_operators.jl:26-263   (ghosts synthesised on load)
//   update_conductivity!          pseudo_transient/conductivity.jl:1-13   (alpha from u)
//   update_iteration_params!      iteration_parameters.jl:2-34   (Re, dtau_rho, theta_r_dtau)
//   compute_flux!                 flux/fluxes.jl:5-26, fluxes_cpu.jl:33-115 (damped q, plain q')
//   compute_update!               update.jl:2-21 with flux_divergence.jl:7-42
// and, on error-check iterations, update_residual! (residuals.jl:24-47) with a
// per-wave partial of sum(r^2) and max|r| for L2_norm (residuals.jl:1-22).
//
// Decomposition.  A 64-lane wave owns a strip of 62 columns (lane l <-> column
// x0-1+l; lanes 0 and 63 are the strip's halo columns) and marches up a chunk
// of rows.  At row j it loads row j+1 once, builds the new damped fluxes of
// row j+1 (xi faces) and row j (eta faces), exchanges neighbour columns with
// wave shuffles, and updates u(i, j).  Every cell of u and every face of q is
// read from HBM once per iteration (plus the two halo rows of the chunk), and
// u and q are written once: the algorithmic 12 values per cell.
//
// u and q are ping-ponged (src -> dst) because neighbouring waves read the
// old values of a strip's halo columns and rows.  Faces outside
// expand_lower(domain, axis) are never computed by the reference and stay 0
// (fluxes_cpu.jl:40-41): here they are masked to 0.
//
// The arithmetic follows the reference operation by operation (same operand
// order, no contraction: the FP64 build uses -ffp-contract=off), so the FP64
// sweep reproduces the oracle to the last bit wherever pow() is not involved.
#include <limits>

#include "cdiff_dev.h"
#include "sweep_args.h"

namespace cdiff {

template <typename T>
struct RowVals {
  T u, a, th, dt;
};

// Value of u at (i, r) after applybcs!, for i in [lo-1, hi+1], r in [jlo-1, jhi+1].
template <typename T>
__device__ __forceinline__ T ghost_u(const Sweep2DArgs<T>& A, int i, int r) {
  const Geo& g = A.g;
  const bool gi = (i < g.lo[0]) | (i > g.hi[0]);
  const bool gj = (r < g.lo[1]) | (r > g.hi[1]);
  if (gi == gj) return A.u_src[at(g, i, r, 0)];  // domain cell, or a corner (never written)
  int f, si, sr;
  if (gi) {
    f = (i < g.lo[0]) ? 0 : 1;
    si = A.gh.src[f];
    sr = r;
  } else {
    f = (r < g.lo[1]) ? 2 : 3;
    si = i;
    sr = A.gh.src[f];
  }
  const int m = A.gh.mode[f];
  if (m == G_CONST) return (T)A.gh.val[f];
  if (m == G_KEEP) return A.u_src[at(g, i, r, 0)];
  T v = A.u_src[at(g, si, sr, 0)];
  return m == G_NEG ? -v : v;
}

template <typename T>
__device__ __forceinline__ RowVals<T> load_row(const Sweep2DArgs<T>& A, int i, int r, bool active) {
  RowVals<T> v;
  v.u = T(0);
  v.a = T(0);
  v.th = T(0);
  v.dt = T(0);
  if (!active) return v;
  const int64_t c = at(A.g, i, r, 0);
  v.u = ghost_u(A, i, r);
  if (A.frozen_alpha) {
    v.a = A.frozen_alpha[c];
    if (A.alpha_only) {
      const T rho = A.rho ? A.rho[c] : T(1);
      iter_params<T>(A.ic, rho, v.a, v.dt, v.th);
    } else {
      v.th = A.frozen_theta[c];
      v.dt = A.frozen_dtau[c];
    }
  } else {
    const T rho = A.rho ? A.rho[c] : T(1);
    v.a = alpha_of<T>(A.kap, rho, v.u);
    iter_params<T>(A.ic, rho, v.a, v.dt, v.th);
  }
  return v;
}

// flux_kernel! / fluxprime_kernel! (flux/fluxes.jl:5-26) for the face between
// cells "c" (lower) and "p" (upper).
template <typename T>
__device__ __forceinline__ void face_flux(int harmonic, T q_old, const RowVals<T>& p, const RowVals<T>& c,
                                          T& q_new, T& q_plain) {
  const T abar = face_mean<T>(harmonic, c.a, p.a);
  const T tbar = face_mean<T>(harmonic, c.th, p.th);
  const T du = p.u - c.u;
  const T qp = -abar * du;
  q_new = (q_old * tbar + qp) / (T(1) + tbar);
  q_plain = qp;
}

// flux_divergence (flux_divergence.jl:16-40) from the 12 face values around a cell.
template <typename T>
__device__ __forceinline__ T divergence(T g11, T g22, T c12, T ac, T bc, T x_c, T x_w, T x_n, T x_nw, T x_s,
                                        T x_sw, T y_c, T y_s, T y_e, T y_es, T y_w, T y_ws) {
  const T t1 = g11 * (x_c - x_w);
  const T t2 = g22 * (y_c - y_s);
  const T t3 = c12 * ((x_n + x_nw) - (x_s + x_sw));
  const T t4 = c12 * ((y_e + y_es) - (y_w + y_ws));
  const T t5 = ac * (T(0.5) * (x_c + x_w));
  const T t6 = bc * (T(0.5) * (y_c + y_s));
  return t1 + t2 + t3 + t4 + t5 + t6;
}

// Julia isapprox(x, y; rtol) with atol = 0.
template <typename T>
__device__ __forceinline__ bool jl_isapprox(T x, T y, T rtol) {
  if (x == y) return true;
  if (!(isfinite(x) && isfinite(y))) return false;
  const T m = fabs(x) > fabs(y) ? fabs(x) : fabs(y);
  return fabs(x - y) <= rtol * m;
}

// Raw loads of one pipeline stage: row r of the iterate (u with its ghost
// rule, rho), the old fluxes of the faces built at this stage (xi faces of
// row r, eta faces between rows r-1 and r), and the cell data of row r-1.
template <typename T>
struct Pre {
  T u, rho, qx, qy;
  T up, g11, g22, c12, ac, bc, S, tA, tB;
  T fa, fth, fdt;
};

template <typename T>
__device__ __forceinline__ Pre<T> prefetch(const Sweep2DArgs<T>& A, int i, int r, bool active, bool xi_face,
                                           bool eta_face, bool dom_col) {
  const Geo& g = A.g;
  Pre<T> p;
  p.u = p.rho = p.qx = p.qy = T(0);
  p.up = p.g11 = p.g22 = p.c12 = p.ac = p.bc = p.S = p.tA = p.tB = T(0);
  p.fa = p.fth = p.fdt = T(0);
  if (!active) return p;
  const int64_t c = at(g, i, r, 0);
  p.u = ghost_u(A, i, r);
  if (A.frozen_alpha) {
    p.fa = A.frozen_alpha[c];
    if (!A.alpha_only) {
      p.fth = A.frozen_theta[c];
      p.fdt = A.frozen_dtau[c];
    }
  }
  if (A.rho) p.rho = A.rho[c];
  if (xi_face && r <= g.hi[1]) p.qx = A.qx_src[c];
  if (eta_face) p.qy = A.qy_src[at(g, i, r - 1, 0)];
  if (dom_col && r - 1 <= g.hi[1]) {
    const int64_t cm = at(g, i, r - 1, 0);
    p.g11 = A.g11[cm];
    p.g22 = A.g22[cm];
    p.c12 = A.c12[cm];
    p.ac = A.ac[cm];
    p.bc = A.bc[cm];
    if (A.base) {
      p.up = A.base[cm];
    } else {
      p.up = A.u_prev[cm];
      if (A.S) p.S = A.S[cm];
      if (A.tzA) {
        p.tA = A.tzA[cm];
        p.tB = A.tzB[cm];
      }
    }
  }
  return p;
}

template <typename T>
__device__ __forceinline__ RowVals<T> make_row(const Sweep2DArgs<T>& A, const Pre<T>& p, int i, int r, bool active) {
  RowVals<T> v;
  v.u = p.u;
  v.a = v.th = v.dt = T(0);
  if (!active) return v;
  const T rho = A.rho ? p.rho : T(1);
  if (A.frozen_alpha) {
    v.a = p.fa;
    if (A.alpha_only) {
      iter_params<T>(A.ic, rho, v.a, v.dt, v.th);
    } else {
      v.th = p.fth;
      v.dt = p.fdt;
    }
  } else {
    v.a = alpha_of<T>(A.kap, rho, v.u);
    iter_params<T>(A.ic, rho, v.a, v.dt, v.th);
  }
  return v;
}

template <typename T, bool CHECK>
__global__ __launch_bounds__(64) void sweep2d_kernel(Sweep2DArgs<T> A) {
  if (A.ctrl->done) return;
  const int lane = threadIdx.x;
  const Geo& g = A.g;
  // XCD-aware tile order: blocks b, b+8, b+16, ... land on one XCD; give them
  // consecutive tiles (vertical neighbours in a strip) so shared halo rows hit
  // that XCD's L2.
  const int nb = gridDim.x;
  const int per = nb >> 3;
  const int b = blockIdx.x;
  const int tile = (b & 7) * per + (b >> 3);
  if (tile >= A.ntiles) {
    if (CHECK && lane == 0) {
      A.part_sum[b] = 0.0;
      A.part_max[b] = 0.0;
    }
    return;
  }
  const int strip = tile / A.chunks;
  const int chunk = tile - strip * A.chunks;
  const int x0 = g.lo[0] + strip * OWN;
  const int i = x0 - 1 + lane;
  const int y0 = g.lo[1] + chunk * A.ch;
  const int y1 = min(y0 + A.ch - 1, g.hi[1]);

  const bool active = i <= g.hi[0] + 1;
  const bool dom_col = (lane >= 1) & (lane <= OWN) & (i <= g.hi[0]);
  const bool xi_face = (i >= g.lo[0] - 1) & (i <= g.hi[0]);  // column of a computed xi face
  const bool eta_face = (i >= g.lo[0]) & (i <= g.hi[0]);      // column of a computed eta face
  const bool writes_x = xi_face & (dom_col | (i == g.lo[0] - 1));
  const bool ghost_col = active & ((i == g.lo[0] - 1) | (i == g.hi[0] + 1));
  const bool ghost_col_write = ghost_col && A.gh.mode[i < g.lo[0] ? 0 : 1] != G_KEEP;
  const int harm = A.harmonic;
  const T dt = (T)A.ic.dt;
  const T eps = std::numeric_limits<T>::epsilon();

  T* __restrict__ udst = A.u_dst;
  T* __restrict__ qxd = A.qx_dst;
  T* __restrict__ qyd = A.qy_dst;

  // ---- prologue: rows y0-1 and y0
  RowVals<T> rs = load_row(A, i, y0 - 1, active);
  RowVals<T> rc = load_row(A, i, y0, active);
  RowVals<T> rsR, rcR;
  rsR.u = from_right(rs.u);
  rsR.a = from_right(rs.a);
  rsR.th = from_right(rs.th);
  rcR.u = from_right(rc.u);
  rcR.a = from_right(rc.a);
  rcR.th = from_right(rc.th);

  T x_s = T(0), px_s = T(0), x_c = T(0), px_c = T(0), y_s = T(0), py_s = T(0);
  {
    T qn, qp;
    if (xi_face && (y0 - 1 >= g.lo[1])) {
      face_flux<T>(harm, A.qx_src[at(g, i, y0 - 1, 0)], rsR, rs, qn, qp);
      x_s = qn;
      px_s = qp;
    }
    if (xi_face) {
      face_flux<T>(harm, A.qx_src[at(g, i, y0, 0)], rcR, rc, qn, qp);
      x_c = qn;
      px_c = qp;
    }
    if (eta_face) {
      face_flux<T>(harm, A.qy_src[at(g, i, y0 - 1, 0)], rc, rs, qn, qp);
      y_s = qn;
      py_s = qp;
    }
  }
  if (y0 == g.lo[1] && dom_col) {
    // eta faces of row jlo-1 and the ghost row below the domain
    qyd[at(g, i, y0 - 1, 0)] = y_s;
    if (A.gh.mode[2] != G_KEEP) udst[at(g, i, y0 - 1, 0)] = rs.u;
  }
  T x_sw = from_left(x_s), x_w = from_left(x_c);
  T y_ws = from_left(y_s), y_es = from_right(y_s);
  T px_sw = T(0), px_w = T(0), py_ws = T(0), py_es = T(0);
  if (CHECK) {
    px_sw = from_left(px_s);
    px_w = from_left(px_c);
    py_ws = from_left(py_s);
    py_es = from_right(py_s);
  }

  double acc = 0.0, amax = 0.0;

  // Software pipeline: the loads of row j+2 (and of the cell data of row j+1)
  // are issued before the arithmetic of row j+1 / cell row j, so memory
  // latency overlaps the FP64 work.
  Pre<T> pn = prefetch(A, i, y0 + 1, active, xi_face, eta_face, dom_col);
  for (int j = y0; j <= y1; ++j) {
    const int r = j + 1;
    const Pre<T> pc = pn;
    if (j + 1 <= y1) pn = prefetch(A, i, j + 2, active, xi_face, eta_face, dom_col);
    const RowVals<T> rn = make_row(A, pc, i, r, active);
    RowVals<T> rnR;
    rnR.u = from_right(rn.u);
    rnR.a = from_right(rn.a);
    rnR.th = from_right(rn.th);
    T x_n = T(0), px_n = T(0), y_c = T(0), py_c = T(0);
    {
      T qn, qp;
      if (xi_face && r <= g.hi[1]) {
        face_flux<T>(harm, pc.qx, rnR, rn, qn, qp);
        x_n = qn;
        px_n = qp;
      }
      if (eta_face) {
        face_flux<T>(harm, pc.qy, rn, rc, qn, qp);
        y_c = qn;
        py_c = qp;
      }
    }
    const T x_nw = from_left(x_n);
    const T y_w = from_left(y_c);
    const T y_e = from_right(y_c);
    T px_nw = T(0), py_w = T(0), py_e = T(0);
    if (CHECK) {
      px_nw = from_left(px_n);
      py_w = from_left(py_c);
      py_e = from_right(py_c);
    }

    if (dom_col) {
      const int64_t c = at(g, i, j, 0);
      const T D = divergence<T>(pc.g11, pc.g22, pc.c12, pc.ac, pc.bc, x_c, x_w, x_n, x_nw, x_s, x_sw, y_c, y_s, y_e,
                                y_es, y_w, y_ws);
      T un, base = T(0), up = T(0), S = T(0), f = T(0);
      if (A.base) {
        base = pc.up;
        un = (rc.u + rc.dt * (base - D)) / (T(1) + rc.dt / dt);
      } else {
        up = pc.up;
        S = pc.S;
        if (A.tzA) f = pc.tA + pc.tB * rc.a;
        T rhs = up / dt - D + S;
        if (A.tzA) rhs = rhs + f;
        un = (rc.u + rc.dt * rhs) / (T(1) + rc.dt / dt);
      }
      udst[c] = un;
      qyd[c] = y_c;
      if (CHECK) {
        const T Dp = divergence<T>(pc.g11, pc.g22, pc.c12, pc.ac, pc.bc, px_c, px_w, px_n, px_nw, px_s, px_sw, py_c,
                                   py_s, py_e, py_es, py_w, py_ws);
        T res;
        if (A.base) {
          res = base - un / dt - Dp;
        } else {
          T du = un - up;
          if (jl_isapprox<T>(un, up, eps)) du = copysign(T(0), du);
          res = -du / dt - Dp + S;
          if (A.tzA) res = res + f;
        }
        const double rd = (double)res;
        acc += rd * rd;
        amax = nanmax(amax, fabs(rd));
      }
    }
    if (writes_x) qxd[at(g, i, j, 0)] = x_c;
    if (ghost_col_write) udst[at(g, i, j, 0)] = rc.u;
    if (r == g.hi[1] + 1 && dom_col && A.gh.mode[3] != G_KEEP) udst[at(g, i, r, 0)] = rn.u;

    // rotate the window
    x_s = x_c;
    x_c = x_n;
    x_sw = x_w;
    x_w = x_nw;
    y_s = y_c;
    y_ws = y_w;
    y_es = y_e;
    if (CHECK) {
      px_s = px_c;
      px_c = px_n;
      px_sw = px_w;
      px_w = px_nw;
      py_s = py_c;
      py_ws = py_w;
      py_es = py_e;
    }
    rc = rn;
  }
  if (CHECK) {
    acc = wave_sum(acc);
    amax = wave_max(amax);
    if (lane == 0) {
      A.part_sum[b] = acc;
      A.part_max[b] = amax;
    }
  }
}

template <typename T>
void launch_sweep2d(const Sweep2DArgs<T>& A, bool check, int nblocks, hipStream_t s) {
  if (check)
    hipLaunchKernelGGL((sweep2d_kernel<T, true>), dim3(nblocks), dim3(WAVE), 0, s, A);
  else
    hipLaunchKernelGGL((sweep2d_kernel<T, false>), dim3(nblocks), dim3(WAVE), 0, s, A);
}

template void launch_sweep2d<double>(const Sweep2DArgs<double>&, bool, int, hipStream_t);
template void launch_sweep2d<float>(const Sweep2DArgs<float>&, bool, int, hipStream_t);

}  // namespace cdiff
