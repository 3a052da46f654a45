// dc_outer2d.hip -- one pass per outer iteration of the 2D defect-correction
// step: apply the correction, evaluate the FP64 residual, and build the
// coefficients of the next FP32 correction sweep.
//
// Per 64 x 8 tile (plus a one-cell ring, in LDS):
//   cells   u = u_old + e (domain), the BC rule for first-layer ghosts
//           (local rules only: no periodic faces), alpha(u) in FP64; theta,
//           the update coefficient C and d(alpha)/du in FP32; writes u of the
//           tile and its ghost strips
//   faces   q' = -abar (u1 - u0) on expand_lower(domain, axis) (FP64), and for
//           the tile's faces the sweep coefficients P, M, W (FP32)
//   update  the residual of the reference discrete equations
//           (residuals.jl:24-47) per tile cell, in FP64, stored in FP32 as the
//           right-hand side of the correction equation; block partials of
//           sum(r^2) and max|r| for the FP64 norm
// The FP64 arithmetic (u, alpha, abar, q', the residual and its norm) is that
// of k_dc_update and k_dc_residual; each face is evaluated once instead of by
// every cell using it.  alpha is not stored: the step's epilogue forms it
// once from the final iterate.
#include <limits>

#include "cdiff_dev.h"
#include "dc.h"
#include "lin_coef_ops.h"

namespace cdiff {

namespace {

constexpr int OTX = 64, OTY = DC_OUTER2D_TY;
constexpr int OCX = OTX + 2, OCY = OTY + 2, ONC = OCX * OCY;
constexpr int OFXN = OTX + 1, OFYN = OTX + 2;
constexpr int ONFX = OFXN * OCY, ONFY = OFYN * (OTY + 1);
constexpr int ONT = OTX * OTY;

__device__ __forceinline__ float gfac(const Ghost& gh, int f) {
  const int m = gh.mode[f];
  return m == G_COPY ? 1.0f : (m == G_NEG ? -1.0f : 0.0f);
}

}  // namespace

// UNIT: no density array and cp = 1 (alpha without its division by rho cp)
template <bool UNIT>
__global__ __launch_bounds__(ONT) __attribute__((amdgpu_waves_per_eu(8))) void k_dc_outer2d(DcOuter2D A) {
  __shared__ double sU[ONC], sA[ONC];
  __shared__ float sT[ONC], sDa[ONC];
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
  const IterConst32 k32 = iter_const32(A.ic, A.dtau_scale, A.theta_scale);
  const float t_exp = (float)A.t_exp;

  // inputs of the residual of this thread's cell, loaded before the ring
  // work so that their latency overlaps it
  const int tx = tid % OTX, tyy = tid / OTX;
  const int i = x0 + tx, j = y0 + tyy;
  const bool mine = i <= g.hi[0] && j <= g.hi[1];
  const int64_t cm = mine ? at(g, i, j, 0) : 0;
  double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0, m4 = 0.0, um = 0.0;
  if (mine) {
    m0 = A.c[0][cm];
    m1 = A.c[1][cm];
    m2 = A.c[2][cm];
    m3 = A.c[3][cm];
    m4 = A.c[4][cm];
    um = A.uprev[cm];
  }

  const float* Ae = A.e_sel ? ((A.e_sel->final_buf & 1) ? A.e_buf1 : A.e_buf0) : A.e;
  auto new_u = [&](int i, int r) {
    const int64_t c = at(g, i, r, 0);
    return A.u_old[c] + (Ae ? (double)Ae[c] : 0.0);
  };

  // ---- cells of the ringed tile
#pragma unroll
  for (int p = 0; p < (ONC + ONT - 1) / ONT; ++p) {
    const int t = tid + p * ONT;
    if (t >= ONC) break;
    const int cx = t % OCX, cy = t / OCX;
    const int i = x0 - 1 + cx, r = y0 - 1 + cy;
    double u = 0.0, a = 0.0;
    float th = 0.0f, da = 0.0f;
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
      a = alpha_of<double, UNIT>(A.kap, rho, u);
      float C;
      iter_params32(k32, (float)rho, (float)a, C, th);
      da = dadu32(t_exp, (float)a, (float)u);
      const bool own_r = cy >= 1 && cy <= OTY && r <= g.hi[1];
      const bool own_i = cx >= 1 && cx <= OTX && i <= g.hi[0];
      if (!gi && !gj && own_r && own_i) {
        A.u_new[c] = u;
        A.cc[c] = C;
      } else if ((gi && !gj && own_r) || (gj && !gi && own_i)) {
        A.u_new[c] = u;
      }
    }
    sU[t] = u;
    sA[t] = a;
    sT[t] = th;
    sDa[t] = da;
  }
  __syncthreads();

  // ---- faces: q' for the residual; P, M, W of the tile's faces
  auto coefs = [&](int c0, int c1, double abar, bool lo_face, bool hi_face, int flo, int fhi, float* const* F,
                   int64_t gc) {
    float P = (float)abar, M = P;
    if (A.newton) {
      float d0, d1;
      mean_weights32(harm, (float)sA[c0], (float)sA[c1], d0, d1);
      const float du = (float)(sU[c1] - sU[c0]);
      P = __builtin_fmaf(-d0 * sDa[c0], du, P);
      M = __builtin_fmaf(d1 * sDa[c1], du, M);
    }
    if (lo_face) {
      M = __builtin_fmaf(-gfac(A.gh, flo), P, M);
      P = 0.0f;
    } else if (hi_face) {
      P = __builtin_fmaf(-gfac(A.gh, fhi), M, P);
      M = 0.0f;
    }
    F[0][gc] = P;
    F[1][gc] = M;
    F[2][gc] = damping32(harm, sT[c0], sT[c1]);
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
          coefs(c0, c1, abar, i == g.lo[0] - 1, i == g.hi[0], 0, 1, A.fx, at(g, i, r, 0));
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
          coefs(c0, c1, abar, r == g.lo[1] - 1, r == g.hi[1], 2, 3, A.fy, at(g, i, r, 0));
      }
      sY[u] = q;
    }
  }
  __syncthreads();

  // ---- residual of the tile's cells
  double sq = 0.0, mx = 0.0;
  if (mine) {
    const int xo = (tyy + 1) * OFXN + (tx + 1);
    const int yo = (tyy + 1) * OFYN + (tx + 1);
    const double xc = sX[xo], xw = sX[xo - 1], xn = sX[xo + OFXN], xnw = sX[xo + OFXN - 1];
    const double xs = sX[xo - OFXN], xsw = sX[xo - OFXN - 1];
    const double yc = sY[yo], ys = sY[yo - OFYN], ye = sY[yo + 1], yes = sY[yo - OFYN + 1];
    const double yw = sY[yo - 1], yws = sY[yo - OFYN - 1];
    const double t1 = m0 * (xc - xw);
    const double t2 = m1 * (yc - ys);
    const double t3 = m2 * ((xn + xnw) - (xs + xsw));
    const double t4 = m2 * ((ye + yes) - (yw + yws));
    const double t5 = m3 * (0.5 * (xc + xw));
    const double t6 = m4 * (0.5 * (yc + ys));
    const double D = t1 + t2 + t3 + t4 + t5 + t6;
    const double un = sU[(tyy + 1) * OCX + tx + 1];
    double du = un - um;
    bool approx = un == um;
    if (!approx && isfinite(un) && isfinite(um))
      approx = fabs(un - um) <= std::numeric_limits<double>::epsilon() * fmax(fabs(un), fabs(um));
    if (approx) du = copysign(0.0, du);
    double r = -du / dt - D + (A.S ? A.S[cm] : 0.0);
    if (A.tzA) r = r + (A.tzA[cm] + A.tzB[cm] * sA[(tyy + 1) * OCX + tx + 1]);
    A.r32[cm] = (float)r;
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

void launch_dc_outer2d(const DcOuter2D& A, int nblk, bool unit, hipStream_t stream) {
  if (unit)
    hipLaunchKernelGGL(k_dc_outer2d<true>, dim3(nblk), dim3(ONT), 0, stream, A);
  else
    hipLaunchKernelGGL(k_dc_outer2d<false>, dim3(nblk), dim3(ONT), 0, stream, A);
}

}  // namespace cdiff
