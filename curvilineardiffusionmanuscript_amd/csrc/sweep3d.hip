// sweep3d.hip -- one fused APT pseudo-iteration on a 3D curvilinear mesh.
//
// The 3D counterpart of sweep2d.hip: alpha, the iteration parameters, the
// damped fluxes on the i+1/2, j+1/2, k+1/2 faces, the 3D divergence with its
// six cross terms and three metric-Laplacian terms
// (src/pseudo_transient/flux_divergence.jl:119-239), the update
// (update.jl:2-21) and, on error-check iterations, the residual with q'
// (residuals.jl:24-47) and its norm partials -- in one pass.
//
// Decomposition.  A block is BJ waves: lane l <-> column i = x0-1+l, wave w
// <-> row j = y0-1+w.  Lanes 1..62 of waves 1..BJ-2 own cells; lane 0/63 and
// waves 0/BJ-1 are the tile's halo.  The block marches up a chunk of k
// planes.  At plane k it loads plane k+1 once, builds the new xi/eta fluxes of
// plane k+1 and the zeta fluxes between planes k and k+1, exchanges
// neighbour columns with wave shuffles and neighbour rows through LDS, and
// updates u(i, j, k).
//
// The 3D cross terms 0.5*g*(0.5*(a+b) - 0.5*(c+d)) of the reference are
// evaluated as (0.25*g)*((a+b) - (c+d)): halving is exact in binary floating
// point, so the two forms give the same number (outside the subnormal range).
#include <limits>

#include "cdiff_dev.h"
#include "sweep3d_ops.h"
#include "sweep_args.h"

namespace cdiff {

namespace {

template <typename T>
struct Cell3 {
  T u, a, th, dt;
};

template <typename T>
__device__ __forceinline__ T ghost_u3(const Sweep3DArgs<T>& A, int i, int j, int k) {
  const Geo& g = A.g;
  const int idx[3] = {i, j, k};
  int nout = 0, ax = 0;
  for (int d = 0; d < 3; ++d)
    if (idx[d] < g.lo[d] || idx[d] > g.hi[d]) {
      ++nout;
      ax = d;
    }
  if (nout != 1) return A.u_src[at(g, i, j, k)];
  const int f = 2 * ax + (idx[ax] < g.lo[ax] ? 0 : 1);
  const int m = A.gh.mode[f];
  if (m == G_CONST) return (T)A.gh.val[f];
  if (m == G_KEEP) return A.u_src[at(g, i, j, k)];
  int s[3] = {i, j, k};
  s[ax] = A.gh.src[f];
  const T v = A.u_src[at(g, s[0], s[1], s[2])];
  return m == G_NEG ? -v : v;
}

template <typename T>
__device__ __forceinline__ Cell3<T> load3(const Sweep3DArgs<T>& A, int i, int j, int k, bool active) {
  Cell3<T> v;
  v.u = v.a = v.th = v.dt = T(0);
  if (!active) return v;
  const int64_t c = at(A.g, i, j, k);
  v.u = ghost_u3(A, i, j, k);
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

}  // namespace

template <typename T, bool CHECK, int BJ>
__global__ __launch_bounds__(64 * BJ) void sweep3d_kernel(Sweep3DArgs<T> A) {
  if (A.ctrl->done) return;
  constexpr int NQ = CHECK ? 12 : 6;
  __shared__ T sU[BJ][WAVE], sA[BJ][WAVE], sT[BJ][WAVE];
  __shared__ T sQ[NQ][BJ][WAVE];  // X YK YN YS ZK ZS (+ the same for q')
  __shared__ double sRed[2][BJ];
  const int l = threadIdx.x, w = threadIdx.y;
  const Geo& g = A.g;
  const int nb = gridDim.x, per = nb >> 3, b = blockIdx.x;
  const int tile = (b & 7) * per + (b >> 3);
  if (tile >= A.ntiles) {
    if (CHECK && l == 0 && w == 0) {
      A.part_sum[b] = 0.0;
      A.part_max[b] = 0.0;
    }
    return;
  }
  const int kc = tile % A.chunks;
  const int rest = tile / A.chunks;
  const int si = rest % A.stripsI, sj = rest / A.stripsI;
  const int x0 = g.lo[0] + si * OWN, y0 = g.lo[1] + sj * (BJ - 2);
  const int i = x0 - 1 + l, j = y0 - 1 + w;
  const int z0 = g.lo[2] + kc * A.ch;
  const int z1 = min(z0 + A.ch - 1, g.hi[2]);

  const bool active = (i <= g.hi[0] + 1) & (j <= g.hi[1] + 1);
  const bool lane_own = (l >= 1) & (l <= OWN);
  const bool row_own = (w >= 1) & (w <= BJ - 2);
  const bool own = lane_own & row_own & (i <= g.hi[0]) & (j <= g.hi[1]);
  const bool in_i = (i >= g.lo[0]) & (i <= g.hi[0]);
  const bool in_j = (j >= g.lo[1]) & (j <= g.hi[1]);
  const bool xface = (i >= g.lo[0] - 1) & (i <= g.hi[0]) & in_j;
  const bool yface = in_i & (j >= g.lo[1] - 1) & (j <= g.hi[1]);
  const bool zface = in_i & in_j;
  const bool write_x = xface & row_own & (lane_own | (i == g.lo[0] - 1)) & (j <= g.hi[1]);
  const bool write_y = yface & lane_own & (row_own | (j == g.lo[1] - 1)) & (i <= g.hi[0]);
  // ghost u writes on the i and j faces of the domain
  const bool gi = active & row_own & in_j & ((i == g.lo[0] - 1) | (i == g.hi[0] + 1));
  const bool gi_write = gi && A.gh.mode[i < g.lo[0] ? 0 : 1] != G_KEEP;
  const bool gj = active & lane_own & in_i & ((j == g.lo[1] - 1) | (j == g.hi[1] + 1));
  const bool gj_write = gj && A.gh.mode[j < g.lo[1] ? 2 : 3] != G_KEEP;
  const int harm = A.harmonic;
  const T dt = (T)A.ic.dt;

  T cf[9];
  // ---- prologue: planes z0-1 and z0
  Cell3<T> rs = load3(A, i, j, z0 - 1, active);
  Cell3<T> rc = load3(A, i, j, z0, active);
  T xs = T(0), xk = T(0), ys = T(0), yk = T(0), zs = T(0);
  T pxs = T(0), pxk = T(0), pys = T(0), pyk = T(0), pzs = T(0);
  {
    const T rsRu = from_right(rs.u), rsRa = from_right(rs.a), rsRt = from_right(rs.th);
    const T rcRu = from_right(rc.u), rcRa = from_right(rc.a), rcRt = from_right(rc.th);
    sU[w][l] = rs.u;
    sA[w][l] = rs.a;
    sT[w][l] = rs.th;
    __syncthreads();
    const int wu = w + 1 < BJ ? w + 1 : w;
    const T rsUu = sU[wu][l], rsUa = sA[wu][l], rsUt = sT[wu][l];
    __syncthreads();
    sU[w][l] = rc.u;
    sA[w][l] = rc.a;
    sT[w][l] = rc.th;
    __syncthreads();
    const T rcUu = sU[wu][l], rcUa = sA[wu][l], rcUt = sT[wu][l];
    T qn, qp;
    if (xface && z0 - 1 >= g.lo[2]) {
      flux3<T>(harm, A.qx_src[at(g, i, j, z0 - 1)], rsRu, rsRa, rsRt, rs.u, rs.a, rs.th, qn, qp);
      xs = qn;
      pxs = qp;
    }
    if (xface) {
      flux3<T>(harm, A.qx_src[at(g, i, j, z0)], rcRu, rcRa, rcRt, rc.u, rc.a, rc.th, qn, qp);
      xk = qn;
      pxk = qp;
    }
    if (yface && z0 - 1 >= g.lo[2]) {
      flux3<T>(harm, A.qy_src[at(g, i, j, z0 - 1)], rsUu, rsUa, rsUt, rs.u, rs.a, rs.th, qn, qp);
      ys = qn;
      pys = qp;
    }
    if (yface) {
      flux3<T>(harm, A.qy_src[at(g, i, j, z0)], rcUu, rcUa, rcUt, rc.u, rc.a, rc.th, qn, qp);
      yk = qn;
      pyk = qp;
    }
    if (zface) {
      flux3<T>(harm, A.qz_src[at(g, i, j, z0 - 1)], rc.u, rc.a, rc.th, rs.u, rs.a, rs.th, qn, qp);
      zs = qn;
      pzs = qp;
    }
    if (z0 == g.lo[2] && own) {
      A.qz_dst[at(g, i, j, z0 - 1)] = zs;
      if (A.gh.mode[4] != G_KEEP) A.u_dst[at(g, i, j, z0 - 1)] = rs.u;
    }
    __syncthreads();
  }

  double acc = 0.0, amax = 0.0;
  for (int k = z0; k <= z1; ++k) {
    const int kn = k + 1;
    Cell3<T> rn = load3(A, i, j, kn, active);
    sU[w][l] = rn.u;
    sA[w][l] = rn.a;
    sT[w][l] = rn.th;
    const T rnRu = from_right(rn.u), rnRa = from_right(rn.a), rnRt = from_right(rn.th);
    __syncthreads();
    const int wu = w + 1 < BJ ? w + 1 : w;
    const T rnUu = sU[wu][l], rnUa = sA[wu][l], rnUt = sT[wu][l];
    T xn = T(0), yn = T(0), zk = T(0), pxn = T(0), pyn = T(0), pzk = T(0);
    {
      T qn, qp;
      if (xface && kn <= g.hi[2]) {
        flux3<T>(harm, A.qx_src[at(g, i, j, kn)], rnRu, rnRa, rnRt, rn.u, rn.a, rn.th, qn, qp);
        xn = qn;
        pxn = qp;
      }
      if (yface && kn <= g.hi[2]) {
        flux3<T>(harm, A.qy_src[at(g, i, j, kn)], rnUu, rnUa, rnUt, rn.u, rn.a, rn.th, qn, qp);
        yn = qn;
        pyn = qp;
      }
      if (zface) {
        flux3<T>(harm, A.qz_src[at(g, i, j, k)], rn.u, rn.a, rn.th, rc.u, rc.a, rc.th, qn, qp);
        zk = qn;
        pzk = qp;
      }
    }
    sQ[0][w][l] = xk;
    sQ[1][w][l] = yk;
    sQ[2][w][l] = yn;
    sQ[3][w][l] = ys;
    sQ[4][w][l] = zk;
    sQ[5][w][l] = zs;
    if (CHECK) {
      sQ[6][w][l] = pxk;
      sQ[7][w][l] = pyk;
      sQ[8][w][l] = pyn;
      sQ[9][w][l] = pys;
      sQ[10][w][l] = pzk;
      sQ[11][w][l] = pzs;
    }
    // neighbour columns (all lanes take part in the shuffles)
    Around<T> q;
    q.xc = xk;
    q.xW = from_left(xk);
    q.xKp = xn;
    q.xKpW = from_left(xn);
    q.xKm = xs;
    q.xKmW = from_left(xs);
    q.yc = yk;
    q.yE = from_right(yk);
    q.yW = from_left(yk);
    q.yKp = yn;
    q.yKm = ys;
    q.zc = zk;
    q.zKm = zs;
    q.zE = from_right(zk);
    q.zEKm = from_right(zs);
    q.zW = from_left(zk);
    q.zWKm = from_left(zs);
    Around<T> p;
    if (CHECK) {
      p.xc = pxk;
      p.xW = from_left(pxk);
      p.xKp = pxn;
      p.xKpW = from_left(pxn);
      p.xKm = pxs;
      p.xKmW = from_left(pxs);
      p.yc = pyk;
      p.yE = from_right(pyk);
      p.yW = from_left(pyk);
      p.yKp = pyn;
      p.yKm = pys;
      p.zc = pzk;
      p.zKm = pzs;
      p.zE = from_right(pzk);
      p.zEKm = from_right(pzs);
      p.zW = from_left(pzk);
      p.zWKm = from_left(pzs);
    }
    __syncthreads();
    if (own) {
      const int wp = w + 1, wm = w - 1;
      q.xJp = sQ[0][wp][l];
      q.xJpW = sQ[0][wp][l - 1];
      q.xJm = sQ[0][wm][l];
      q.xJmW = sQ[0][wm][l - 1];
      q.yJm = sQ[1][wm][l];
      q.yEJm = sQ[1][wm][l + 1];
      q.yWJm = sQ[1][wm][l - 1];
      q.yKpJm = sQ[2][wm][l];
      q.yKmJm = sQ[3][wm][l];
      q.zJp = sQ[4][wp][l];
      q.zJpKm = sQ[5][wp][l];
      q.zJm = sQ[4][wm][l];
      q.zJmKm = sQ[5][wm][l];
      const int64_t c = at(g, i, j, k);
      for (int t = 0; t < 9; ++t) cf[t] = A.coef[t][c];
      const T D = div3<T>(cf, q);
      T un, base = T(0), up = T(0), S = T(0), f = T(0);
      if (A.base) {
        base = A.base[c];
        un = (rc.u + rc.dt * (base - D)) / (T(1) + rc.dt / dt);
      } else {
        up = A.u_prev[c];
        S = A.S ? A.S[c] : T(0);
        if (A.tzA) f = A.tzA[c] + A.tzB[c] * rc.a;
        T rhs = up / dt - D + S;
        if (A.tzA) rhs = rhs + f;
        un = (rc.u + rc.dt * rhs) / (T(1) + rc.dt / dt);
      }
      A.u_dst[c] = un;
      A.qy_dst[c] = yk;
      A.qz_dst[c] = zk;
      if (CHECK) {
        p.xJp = sQ[6][wp][l];
        p.xJpW = sQ[6][wp][l - 1];
        p.xJm = sQ[6][wm][l];
        p.xJmW = sQ[6][wm][l - 1];
        p.yJm = sQ[7][wm][l];
        p.yEJm = sQ[7][wm][l + 1];
        p.yWJm = sQ[7][wm][l - 1];
        p.yKpJm = sQ[8][wm][l];
        p.yKmJm = sQ[9][wm][l];
        p.zJp = sQ[10][wp][l];
        p.zJpKm = sQ[11][wp][l];
        p.zJm = sQ[10][wm][l];
        p.zJmKm = sQ[11][wm][l];
        const T Dp = div3<T>(cf, p);
        T res;
        if (A.base) {
          res = base - un / dt - Dp;
        } else {
          T du = un - up;
          if (isapprox_eps<T>(un, up)) du = copysign(T(0), du);
          res = -du / dt - Dp + S;
          if (A.tzA) res = res + f;
        }
        const double rd = (double)res;
        acc += rd * rd;
        amax = nanmax(amax, fabs(rd));
      }
    }
    if (write_x) A.qx_dst[at(g, i, j, k)] = xk;
    if (write_y && !own) A.qy_dst[at(g, i, j, k)] = yk;
    if (gi_write || gj_write) A.u_dst[at(g, i, j, k)] = rc.u;
    if (kn == g.hi[2] + 1 && own && A.gh.mode[5] != G_KEEP) A.u_dst[at(g, i, j, kn)] = rn.u;
    xs = xk;
    xk = xn;
    ys = yk;
    yk = yn;
    zs = zk;
    if (CHECK) {
      pxs = pxk;
      pxk = pxn;
      pys = pyk;
      pyk = pyn;
      pzs = pzk;
    }
    rc = rn;
    __syncthreads();
  }
  if (CHECK) {
    acc = wave_sum(acc);
    amax = wave_max(amax);
    if (l == 0) {
      sRed[0][w] = acc;
      sRed[1][w] = amax;
    }
    __syncthreads();
    if (l == 0 && w == 0) {
      double s = 0.0, m = 0.0;
      for (int t = 0; t < BJ; ++t) {
        s += sRed[0][t];
        m = nanmax(m, sRed[1][t]);
      }
      A.part_sum[b] = s;
      A.part_max[b] = m;
    }
  }
}

template <typename T>
void launch_sweep3d(const Sweep3DArgs<T>& A, bool check, int nblocks, hipStream_t s) {
  const dim3 blk(WAVE, SWEEP3D_BJ);
  if (check)
    hipLaunchKernelGGL((sweep3d_kernel<T, true, SWEEP3D_BJ>), dim3(nblocks), blk, 0, s, A);
  else
    hipLaunchKernelGGL((sweep3d_kernel<T, false, SWEEP3D_BJ>), dim3(nblocks), blk, 0, s, A);
}

template void launch_sweep3d<double>(const Sweep3DArgs<double>&, bool, int, hipStream_t);
template void launch_sweep3d<float>(const Sweep3DArgs<float>&, bool, int, hipStream_t);

}  // namespace cdiff
