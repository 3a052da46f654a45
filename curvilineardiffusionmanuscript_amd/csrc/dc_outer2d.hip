// dc_outer2d.hip -- one pass per outer iteration of the 2D defect-correction
// step: apply the correction, evaluate the FP64 residual, and build the
// coefficients of the next FP32 correction sweep.
//
// Per 64 x 8 tile (plus a one-cell ring, in LDS):
//   cells   u = u_old + e (domain), the BC rule for first-layer ghosts
//           (local rules only: no periodic faces), alpha(u), theta, dtau,
//           d(alpha)/du; writes u and alpha of the tile and its ghost strips
//   faces   q' = -abar (u1 - u0) on expand_lower(domain, axis) (FP64), and for
//           the tile's faces the sweep coefficients P, M, W (k_lin_coefs)
//   update  the residual of the reference discrete equations
//           (residuals.jl:24-47) per tile cell, in FP64, stored in FP32 as the
//           right-hand side of the correction equation; block partials of
//           sum(r^2) and max|r| for the FP64 norm
// Arithmetic per cell and per face is that of k_dc_update, k_dc_residual and
// k_lin_coefs; each face is evaluated once instead of by every cell using it.
#include <limits>

#include "cdiff_dev.h"
#include "dc.h"

namespace cdiff {

namespace {

constexpr int OTX = 64, OTY = DC_OUTER2D_TY;
constexpr int OCX = OTX + 2, OCY = OTY + 2, ONC = OCX * OCY;
constexpr int OFXN = OTX + 1, OFYN = OTX + 2;
constexpr int ONFX = OFXN * OCY, ONFY = OFYN * (OTY + 1);
constexpr int ONT = OTX * OTY;

__device__ __forceinline__ double gfac(const Ghost& gh, int f) {
  const int m = gh.mode[f];
  return m == G_COPY ? 1.0 : (m == G_NEG ? -1.0 : 0.0);
}

__device__ __forceinline__ double dadu(double t_exp, double a, double u) {
  const double d = (u != 0.0 && isfinite(u)) ? t_exp * a / u : 0.0;
  return isfinite(d) ? d : 0.0;
}

}  // namespace

__global__ __launch_bounds__(ONT) void k_dc_outer2d(DcOuter2D A) {
  __shared__ double sU[ONC], sA[ONC], sT[ONC], sDa[ONC];
  __shared__ double sX[ONFX], sY[ONFY];
  __shared__ double sRed[2][ONT / WAVE];
  const Geo& g = A.g;
  const int tid = threadIdx.x;
  const int nb = gridDim.x, per = nb >> 3, b = blockIdx.x;
  const int tile = (b & 7) * per + (b >> 3);
  if (tile >= A.ntiles) {
    if (tid == 0) {
      A.psum[b] = 0.0;
      A.pmax[b] = 0.0;
    }
    return;
  }
  const int bx = tile / A.chunks, by = tile - bx * A.chunks;
  const int x0 = g.lo[0] + bx * OTX, y0 = g.lo[1] + by * OTY;
  const int harm = A.harmonic;
  const double dt = A.ic.dt;

  auto new_u = [&](int i, int r) {
    const int64_t c = at(g, i, r, 0);
    return A.u_old[c] + (A.e ? (double)A.e[c] : 0.0);
  };

  // ---- cells of the ringed tile
#pragma unroll
  for (int p = 0; p < (ONC + ONT - 1) / ONT; ++p) {
    const int t = tid + p * ONT;
    if (t >= ONC) break;
    const int cx = t % OCX, cy = t / OCX;
    const int i = x0 - 1 + cx, r = y0 - 1 + cy;
    double u = 0.0, a = 0.0, th = 0.0, da = 0.0;
    if (i >= g.lo[0] - 1 && i <= g.hi[0] + 1 && r >= g.lo[1] - 1 && r <= g.hi[1] + 1) {
      const int64_t c = at(g, i, r, 0);
      const bool gi = (i < g.lo[0]) | (i > g.hi[0]);
      const bool gj = (r < g.lo[1]) | (r > g.hi[1]);
      if (!gi && !gj) {
        u = new_u(i, r);
      } else if (gi != gj) {
        const int f = gi ? (i < g.lo[0] ? 0 : 1) : (r < g.lo[1] ? 2 : 3);
        const int m = A.gh.mode[f];
        if (m == G_CONST) {
          u = A.gh.val[f];
        } else {
          const double us = gi ? new_u(A.gh.src[f], r) : new_u(i, A.gh.src[f]);
          u = m == G_NEG ? -us : us;
        }
      } else {
        u = A.u_old[c];
      }
      const double rho = A.rho ? A.rho[c] : 1.0;
      a = alpha_of<double>(A.kap, rho, u);
      double dtau;
      iter_params<double>(A.ic, rho, a, dtau, th);
      dtau *= A.dtau_scale;
      th *= A.theta_scale;
      da = dadu(A.t_exp, a, u);
      const bool own_r = cy >= 1 && cy <= OTY && r <= g.hi[1];
      const bool own_i = cx >= 1 && cx <= OTX && i <= g.hi[0];
      if (!gi && !gj && own_r && own_i) {
        A.u_new[c] = u;
        A.a64[c] = a;
        A.cc[c] = (float)(dtau / (1.0 + dtau / dt));
      } else if ((gi && !gj && own_r) || (gj && !gi && own_i)) {
        A.u_new[c] = u;
        A.a64[c] = a;
      }
    }
    sU[t] = u;
    sA[t] = a;
    sT[t] = th;
    sDa[t] = da;
  }
  __syncthreads();

  // ---- faces: q' for the residual; P, M, W of the tile's faces
  auto coefs = [&](int c0, int c1, double a0, double a1, double abar, bool lo_face, bool hi_face, int flo, int fhi,
                   float* const* F, int64_t gc) {
    const double tbar = face_mean<double>(harm, sT[c0], sT[c1]);
    double P = abar, M = abar;
    if (A.newton) {
      const double s = a0 + a1;
      double d0 = 0.5, d1 = 0.5;
      if (harm) {
        d0 = (s != 0.0) ? 2.0 * a1 * a1 / (s * s) : 0.0;
        d1 = (s != 0.0) ? 2.0 * a0 * a0 / (s * s) : 0.0;
      }
      const double du = sU[c1] - sU[c0];
      P -= d0 * sDa[c0] * du;
      M += d1 * sDa[c1] * du;
    }
    if (lo_face) {
      M -= gfac(A.gh, flo) * P;
      P = 0.0;
    } else if (hi_face) {
      P -= gfac(A.gh, fhi) * M;
      M = 0.0;
    }
    F[0][gc] = (float)P;
    F[1][gc] = (float)M;
    F[2][gc] = (float)(tbar / (1.0 + tbar));
  };
#pragma unroll
  for (int p = 0; p < (ONFX + ONFY + ONT - 1) / ONT; ++p) {
    const int t = tid + p * ONT;
    if (t < ONFX) {
      const int fx = t % OFXN, cy = t / OFXN;
      const int i = x0 - 1 + fx, r = y0 - 1 + cy;
      double q = 0.0;
      if (i >= g.lo[0] - 1 && i <= g.hi[0] && r >= g.lo[1] && r <= g.hi[1]) {
        const int c0 = cy * OCX + fx, c1 = c0 + 1;
        const double a0 = sA[c0], a1 = sA[c1], abar = face_mean<double>(harm, a0, a1);
        q = -abar * (sU[c1] - sU[c0]);
        if (cy >= 1 && cy <= OTY && (fx >= 1 || i == g.lo[0] - 1))
          coefs(c0, c1, a0, a1, abar, i == g.lo[0] - 1, i == g.hi[0], 0, 1, A.fx, at(g, i, r, 0));
      }
      sX[t] = q;
    } else if (t < ONFX + ONFY) {
      const int u = t - ONFX;
      const int fx = u % OFYN, fy = u / OFYN;
      const int i = x0 - 1 + fx, r = y0 - 1 + fy;
      double q = 0.0;
      if (i >= g.lo[0] && i <= g.hi[0] && r >= g.lo[1] - 1 && r <= g.hi[1]) {
        const int c0 = fy * OCX + fx, c1 = c0 + OCX;
        const double a0 = sA[c0], a1 = sA[c1], abar = face_mean<double>(harm, a0, a1);
        q = -abar * (sU[c1] - sU[c0]);
        if (fx >= 1 && fx <= OTX && (fy >= 1 || r == g.lo[1] - 1))
          coefs(c0, c1, a0, a1, abar, r == g.lo[1] - 1, r == g.hi[1], 2, 3, A.fy, at(g, i, r, 0));
      }
      sY[u] = q;
    }
  }
  __syncthreads();

  // ---- residual of the tile's cells
  const int tx = tid % OTX, tyy = tid / OTX;
  const int i = x0 + tx, j = y0 + tyy;
  double sq = 0.0, mx = 0.0;
  if (i <= g.hi[0] && j <= g.hi[1]) {
    const int64_t c = at(g, i, j, 0);
    const int xo = (tyy + 1) * OFXN + (tx + 1);
    const int yo = (tyy + 1) * OFYN + (tx + 1);
    const double xc = sX[xo], xw = sX[xo - 1], xn = sX[xo + OFXN], xnw = sX[xo + OFXN - 1];
    const double xs = sX[xo - OFXN], xsw = sX[xo - OFXN - 1];
    const double yc = sY[yo], ys = sY[yo - OFYN], ye = sY[yo + 1], yes = sY[yo - OFYN + 1];
    const double yw = sY[yo - 1], yws = sY[yo - OFYN - 1];
    const double t1 = A.c[0][c] * (xc - xw);
    const double t2 = A.c[1][c] * (yc - ys);
    const double t3 = A.c[2][c] * ((xn + xnw) - (xs + xsw));
    const double t4 = A.c[2][c] * ((ye + yes) - (yw + yws));
    const double t5 = A.c[3][c] * (0.5 * (xc + xw));
    const double t6 = A.c[4][c] * (0.5 * (yc + ys));
    const double D = t1 + t2 + t3 + t4 + t5 + t6;
    const double un = sU[(tyy + 1) * OCX + tx + 1], um = A.uprev[c];
    double du = un - um;
    bool approx = un == um;
    if (!approx && isfinite(un) && isfinite(um))
      approx = fabs(un - um) <= std::numeric_limits<double>::epsilon() * fmax(fabs(un), fabs(um));
    if (approx) du = copysign(0.0, du);
    double r = -du / dt - D + (A.S ? A.S[c] : 0.0);
    if (A.tzA) r = r + (A.tzA[c] + A.tzB[c] * sA[(tyy + 1) * OCX + tx + 1]);
    A.r32[c] = (float)r;
    sq = r * r;
    mx = fabs(r);
    if (r != r) mx = r;
  }
  sq = wave_sum(sq);
  mx = wave_max(mx);
  if ((tid & (WAVE - 1)) == 0) {
    sRed[0][tid / WAVE] = sq;
    sRed[1][tid / WAVE] = mx;
  }
  __syncthreads();
  if (tid == 0) {
    double s = 0.0, m = 0.0;
    for (int w = 0; w < ONT / WAVE; ++w) {
      s += sRed[0][w];
      m = nanmax(m, sRed[1][w]);
    }
    A.psum[b] = s;
    A.pmax[b] = m;
  }
}

}  // namespace cdiff
