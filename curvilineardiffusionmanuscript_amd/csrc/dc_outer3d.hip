// dc_outer3d.hip -- one pass per outer iteration of the 3D defect-correction
// step: apply the correction, evaluate the FP64 residual, and build the
// coefficients of the next FP32 correction sweep (sweep3d_lin.hip).
//
// A 512-thread block owns a 64 x 8 column of cells and marches up a chunk of
// planes z0..z1.  Plane step p, for p = z0-1 .. z1+1:
//   cells    of the ringed 66 x 10 plane p: u = u_old + e (domain), the BC
//            rule for first-layer ghosts (local rules only: no periodic
//            faces), alpha(u) in FP64; theta, the update coefficient C and
//            d(alpha)/du in FP32 (lin_coef_ops.h) (LDS, planes p-1 and p);
//            writes u and C of the column (alpha is not stored: the step's
//            epilogue forms it from the final iterate)
//   faces    q' = -abar (u1 - u0) on expand_lower(domain, axis) (FP64, LDS)
//            of the xi and eta faces of plane p and of the zeta faces between
//            planes p-1 and p; the sweep coefficients P, M, W of the column's
//            faces in FP32, as k_lin_coefs forms them
//   residual of plane p-1 per column cell, in FP64: the reference residual
//            (residuals.jl:24-47) with the 3D divergence of
//            flux_divergence.jl:119-239, summed in the order of k_dc_residual;
//            the face planes are double-buffered in LDS (plane p's faces in
//            buffer p & 1), so the residual reads the faces of planes p-1 and
//            p there and carries only the face sums of plane p-2.
// The cell inputs of the next plane step and the residual inputs of the
// next residual are loaded one plane step ahead.
#include <limits>

#include "column3d.h"
#include "dc.h"
#include "lin_coef_ops.h"

namespace cdiff {

namespace {

using OC = Col3<DC_OUTER3D_LY>;
constexpr int OX = OC::X, OY = OC::Y, ORX = OC::RX, ORY = OC::RY, OR = OC::R, OT = OC::T;
constexpr int FXN = OX + 1;  // xi faces per ring row: between region columns fx and fx+1
constexpr int FYN = OY + 1;  // eta face rows: between region rows fy and fy+1
constexpr int NFX = FXN * ORY, NFY = ORX * FYN;
constexpr int NFACE = NFX + NFY + OR;

__device__ __forceinline__ float gfac3(const Ghost& gh, int f) {
  const int m = gh.mode[f];
  return m == G_COPY ? 1.0f : (m == G_NEG ? -1.0f : 0.0f);
}

// what a region cell needs from memory
struct CellIn {
  double v;   // u_old of the cell, or of the BC source cell of a face ghost
  double rho;
  float e;    // correction of the same cell
  int kind;   // 0 none, 1 domain cell or copied face ghost, -1 negated face ghost, 2 fixed value, 3 u_old
};

}  // namespace

// UNIT: no density array and cp = 1 (alpha without its division by rho cp)
template <bool UNIT>
__global__ __launch_bounds__(OT) void k_dc_outer3d(DcOuter3D A) {
  __shared__ double sU[2][OR], sA[2][OR];
  __shared__ float sT[2][OR], sDa[2][OR];
  __shared__ double sX[2][NFX], sY[2][NFY], sZ[2][OR];
  __shared__ double sRed[2][OT / WAVE];
  const Geo& g = A.g;
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  Column3D col3;
  if (!column_of<DC_OUTER3D_LY>(g, A.ntiles, A.chunks, A.stripsI, A.ch, col3)) {
    if (tid == 0) {
      A.psum[b] = 0.0;
      A.pmax[b] = 0.0;
    }
    return;
  }
  const int x0 = col3.x0, y0 = col3.y0, z0 = col3.z0, z1 = col3.z1;
  const int gx = x0 - 1, gy = y0 - 1;
  const int harm = A.harmonic;
  const double dt = A.ic.dt;
  const IterConst32 k32 = iter_const32(A.ic, A.dtau_scale, A.theta_scale);
  const float t_exp = (float)A.t_exp;
  const float* Ae = A.e_sel ? ((A.e_sel->final_buf & 1) ? A.e_buf1 : A.e_buf0) : A.e;

  int rlx[2], rly[2];
  bool rv[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) rv[s] = region_cell<DC_OUTER3D_LY>(s, tid, rlx[s], rly[s]);

  // ---- loads of the cells of plane p
  auto fetch_cells = [&](int p, CellIn* C) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      C[s].v = 0.0;
      C[s].rho = 1.0;
      C[s].e = 0.f;
      C[s].kind = 0;
      if (!rv[s]) continue;
      const int i = gx + rlx[s], j = gy + rly[s];
      const bool valid = i >= g.lo[0] - 1 && i <= g.hi[0] + 1 && j >= g.lo[1] - 1 && j <= g.hi[1] + 1 &&
                         p >= g.lo[2] - 1 && p <= g.hi[2] + 1;
      if (!valid) continue;
      const bool oi = i < g.lo[0] || i > g.hi[0], oj = j < g.lo[1] || j > g.hi[1], ok = p < g.lo[2] || p > g.hi[2];
      const int nout = (int)oi + (int)oj + (int)ok;
      const int64_t c = at(g, i, j, p);
      if (A.rho) C[s].rho = A.rho[c];
      if (nout == 0) {
        C[s].kind = 1;
        C[s].v = A.u_old[c];
        if (Ae) C[s].e = Ae[c];
      } else if (nout == 1) {
        const int f = oi ? (i < g.lo[0] ? 0 : 1) : (oj ? (j < g.lo[1] ? 2 : 3) : (p < g.lo[2] ? 4 : 5));
        const int mode = pick6(A.gh.mode, f);
        if (mode == G_CONST) {
          C[s].kind = 2;
          C[s].v = pick6(A.gh.val, f);
        } else {
          const int src = pick6(A.gh.src, f);
          const int64_t sc = at(g, oi ? src : i, oj ? src : j, ok ? src : p);
          C[s].kind = mode == G_NEG ? -1 : 1;
          C[s].v = A.u_old[sc];
          if (Ae) C[s].e = Ae[sc];
        }
      } else {
        C[s].kind = 3;
        C[s].v = A.u_old[c];
      }
    }
  };

  // ---- cells of plane p from their loads: LDS, and the column's writes
  double u_own = 0.0, a_own = 0.0;  // column cell of plane p (slot 0)
  auto cells = [&](int p, const CellIn* C) {
    const int buf = p & 1;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (!rv[s]) continue;
      const int r = rly[s] * ORX + rlx[s];
      double u = 0.0, a = 0.0;
      float th = 0.0f, da = 0.0f;
      if (C[s].kind != 0) {
        if (C[s].kind == 1 || C[s].kind == -1) {
          const double us = C[s].v + (double)C[s].e;
          u = C[s].kind == -1 ? -us : us;
        } else {
          u = C[s].v;
        }
        a = alpha_of<double, UNIT>(A.kap, C[s].rho, u);
        float Cc;
        iter_params32(k32, (float)C[s].rho, (float)a, Cc, th);
        da = dadu32(t_exp, (float)a, (float)u);
        const int i = gx + rlx[s], j = gy + rly[s];
        const bool gi = i < g.lo[0] || i > g.hi[0], gj = j < g.lo[1] || j > g.hi[1];
        const bool gk = p < g.lo[2] || p > g.hi[2];
        const bool own_i = rlx[s] >= 1 && rlx[s] <= OX, own_j = rly[s] >= 1 && rly[s] <= OY;
        const bool own_k = p >= z0 && p <= z1;
        const int64_t c = at(g, i, j, p);
        if (!gi && !gj && !gk) {
          if (own_i && own_j && own_k) {
            A.u_new[c] = u;
            A.cc[c] = Cc;
          }
        } else if ((int)gi + (int)gj + (int)gk == 1) {
          // face ghosts: the block owning the adjacent column cells writes them
          if ((gi && own_j && own_k) || (gj && own_i && own_k) || (gk && own_i && own_j)) A.u_new[c] = u;
        }
      }
      sU[buf][r] = u;
      sA[buf][r] = a;
      sT[buf][r] = th;
      sDa[buf][r] = da;
      if (s == 0) {
        u_own = u;
        a_own = a;
      }
    }
  };

  // ---- faces of plane p (zeta faces between p-1 and p when `zf`)
  auto coefs = [&](int b0, int r0, int b1, int r1, int flo, int fhi, bool lo_face, bool hi_face, float* F0, float* F1,
                   float* F2, int64_t gc) {
    const double a0 = sA[b0][r0], a1 = sA[b1][r1];
    float P = (float)face_mean<double>(harm, a0, a1), M = P;
    if (A.newton) {
      float d0, d1;
      mean_weights32(harm, (float)a0, (float)a1, d0, d1);
      const float du = (float)(sU[b1][r1] - sU[b0][r0]);
      P = __builtin_fmaf(-d0 * sDa[b0][r0], du, P);
      M = __builtin_fmaf(d1 * sDa[b1][r1], du, M);
    }
    if (lo_face) {
      M = __builtin_fmaf(-gfac3(A.gh, flo), P, M);
      P = 0.0f;
    } else if (hi_face) {
      P = __builtin_fmaf(-gfac3(A.gh, fhi), M, P);
      M = 0.0f;
    }
    F0[gc] = P;
    F1[gc] = M;
    F2[gc] = damping32(harm, sT[b0][r0], sT[b1][r1]);
  };
  auto faces = [&](int p, bool zf) {
    const int bc = p & 1, bm = (p + 1) & 1;
    const bool pdom = p >= g.lo[2] && p <= g.hi[2];
    const bool own_p = p >= z0 && p <= z1;
#pragma unroll 1
    for (int q = 0; q < (NFACE + OT - 1) / OT; ++q) {
      const int t = tid + q * OT;
      if (t < NFX) {
        const int fx = t % FXN, cy = t / FXN;
        const int i = gx + fx, j = gy + cy;
        double v = 0.0;
        if (pdom && i >= g.lo[0] - 1 && i <= g.hi[0] && j >= g.lo[1] && j <= g.hi[1]) {
          const int r = cy * ORX + fx;
          v = -face_mean<double>(harm, sA[bc][r], sA[bc][r + 1]) * (sU[bc][r + 1] - sU[bc][r]);
          if (own_p && cy >= 1 && cy <= OY && (fx >= 1 || i == g.lo[0] - 1))
            coefs(bc, r, bc, r + 1, 0, 1, i == g.lo[0] - 1, i == g.hi[0], A.f + 0 * A.fstride, A.f + 1 * A.fstride, A.f + 2 * A.fstride, at(g, i, j, p));
        }
        sX[bc][t] = v;
      } else if (t < NFX + NFY) {
        const int u = t - NFX;
        const int fx = u % ORX, fy = u / ORX;
        const int i = gx + fx, j = gy + fy;
        double v = 0.0;
        if (pdom && i >= g.lo[0] && i <= g.hi[0] && j >= g.lo[1] - 1 && j <= g.hi[1]) {
          const int r = fy * ORX + fx;
          v = -face_mean<double>(harm, sA[bc][r], sA[bc][r + ORX]) * (sU[bc][r + ORX] - sU[bc][r]);
          if (own_p && fx >= 1 && fx <= OX && (fy >= 1 || j == g.lo[1] - 1))
            coefs(bc, r, bc, r + ORX, 2, 3, j == g.lo[1] - 1, j == g.hi[1], A.f + 3 * A.fstride, A.f + 4 * A.fstride, A.f + 5 * A.fstride, at(g, i, j, p));
        }
        sY[bc][u] = v;
      } else if (t < NFACE) {
        const int r = t - NFX - NFY;
        const int lx = r % ORX, ly = r / ORX;
        const int i = gx + lx, j = gy + ly, k = p - 1;  // face between planes k and k+1, stored at k
        double v = 0.0;
        if (zf && i >= g.lo[0] && i <= g.hi[0] && j >= g.lo[1] && j <= g.hi[1] && k >= g.lo[2] - 1 &&
            k <= g.hi[2]) {
          v = -face_mean<double>(harm, sA[bm][r], sA[bc][r]) * (sU[bc][r] - sU[bm][r]);
          const bool own_k = (k >= z0 && k <= z1) || k == g.lo[2] - 1;
          if (own_k && lx >= 1 && lx <= OX && ly >= 1 && ly <= OY)
            coefs(bm, r, bc, r, 4, 5, k == g.lo[2] - 1, k == g.hi[2], A.f + 6 * A.fstride, A.f + 7 * A.fstride, A.f + 8 * A.fstride, at(g, i, j, k));
        }
        sZ[bc][r] = v;
      }
    }
  };

  // ---- per column cell: residual inputs of plane p
  const int tx = tid % OX, ty = tid / OX;
  const int ci = x0 + tx, cj = y0 + ty;
  const bool col = ci <= g.hi[0] && cj <= g.hi[1];
  struct ResIn {
    double c[9], uprev, src;
  };
  auto fetch_res = [&](int p, ResIn& R) {
    if (!col || p < z0 || p > z1) return;
    const int64_t c = at(g, ci, cj, p);
#pragma unroll
    for (int q = 0; q < 9; ++q) R.c[q] = A.c[q * A.cstride + c];
    R.uprev = A.uprev[c];
    R.src = A.S ? A.S[c] : 0.0;
  };

  // faces around the column cell (LDS, buffer of the plane)
  struct Plane {
    double xc, xw, xeta, yc, ys, yxi;  // xi / eta faces of the plane
  };
  struct ZSt {
    double c, e, w, n, s;  // zeta faces stored at a plane: (0,0), (+1,0), (-1,0), (0,+1), (0,-1)
  };
  const int x_r = (ty + 1) * FXN + tx + 1;  // X(0,0) of the column cell
  const int y_r = (ty + 1) * ORX + tx + 1;  // Y(0,0) and Z(0,0)
  auto read_plane = [&](int fb) {
    const double* X = sX[fb];
    const double* Y = sY[fb];
    Plane P;
    P.xc = X[x_r];
    P.xw = X[x_r - 1];
    P.xeta = (X[x_r + FXN] + X[x_r + FXN - 1]) - (X[x_r - FXN] + X[x_r - FXN - 1]);
    P.yc = Y[y_r];
    P.ys = Y[y_r - ORX];
    P.yxi = (Y[y_r + 1] + Y[y_r - ORX + 1]) - (Y[y_r - 1] + Y[y_r - ORX - 1]);
    return P;
  };
  // X(0) + X(-1) and Y(0) + Y(0,-1) of a plane
  auto xsum_of = [&](int fb) { return sX[fb][x_r] + sX[fb][x_r - 1]; };
  auto ysum_of = [&](int fb) { return sY[fb][y_r] + sY[fb][y_r - ORX]; };
  auto read_z = [&](int fb) {
    const double* Z = sZ[fb];
    ZSt q;
    q.c = Z[y_r];
    q.e = Z[y_r + 1];
    q.w = Z[y_r - 1];
    q.n = Z[y_r + ORX];
    q.s = Z[y_r - ORX];
    return q;
  };

  double sq = 0.0, mx = 0.0;
  // residual of plane k (faces in buffer k & 1, those of plane k+1 and the
  // zeta faces stored at k in buffer (k+1) & 1, the zeta faces stored at k-1
  // in buffer k & 1) with the face sums xm, ym of plane k-1
  auto residual = [&](int k, const ResIn& R, double xm, double ym, double un, double an) {
    if (!col) return;
    const int kb = k & 1, nb = (k + 1) & 1;
    const Plane pc = read_plane(kb);
    const double pn_xsum = xsum_of(nb), pn_ysum = ysum_of(nb);
    const ZSt zm = read_z(kb), zc = read_z(nb);
    const double a1 = R.c[0] * (pc.xc - pc.xw);
    const double a2 = R.c[1] * (pc.yc - pc.ys);
    const double a3 = R.c[2] * (zc.c - zm.c);
    const double qi_eta = R.c[3] * pc.xeta;
    const double qi_zeta = R.c[4] * (pn_xsum - xm);
    const double qj_xi = R.c[3] * pc.yxi;
    const double qj_zeta = R.c[5] * (pn_ysum - ym);
    const double qk_eta = R.c[5] * ((zc.n + zm.n) - (zc.s + zm.s));
    const double qk_xi = R.c[4] * ((zc.e + zm.e) - (zc.w + zm.w));
    const double m1 = R.c[6] * (0.5 * (pc.xc + pc.xw));
    const double m2 = R.c[7] * (0.5 * (pc.yc + pc.ys));
    const double m3 = R.c[8] * (0.5 * (zc.c + zm.c));
    const double D = a1 + a2 + a3 + qi_eta + qi_zeta + qj_xi + qj_zeta + qk_eta + qk_xi + m1 + m2 + m3;
    const double um = R.uprev;
    double du = un - um;
    bool approx = un == um;
    if (!approx && isfinite(un) && isfinite(um))
      approx = fabs(un - um) <= std::numeric_limits<double>::epsilon() * fmax(fabs(un), fabs(um));
    if (approx) du = copysign(0.0, du);
    double r = -du / dt - D + R.src;
    const int64_t c = at(g, ci, cj, k);
    if (A.tzA) r = r + (A.tzA[c] + A.tzB[c] * an);
    A.r32[c] = (float)r;
    sq += r * r;
    const double ar = fabs(r);
    mx = (r != r) ? r : nanmax(mx, ar);
  };

  // ---- prologue: plane step z0-1 (no zeta faces), then z0
  CellIn C[2];
  double xm, ym;                // face sums of the plane below the next residual's
  double u_c = 0.0, a_c = 0.0;  // column cell of the plane whose residual is next
  fetch_cells(z0 - 1, C);
  cells(z0 - 1, C);
  fetch_cells(z0, C);
  __syncthreads();
  faces(z0 - 1, false);
  __syncthreads();
  xm = xsum_of((z0 - 1) & 1);
  ym = ysum_of((z0 - 1) & 1);
  cells(z0, C);
  u_c = u_own;
  a_c = a_own;
  fetch_cells(z0 + 1, C);
  // the residual inputs of a plane are loaded one plane step before the
  // residual that uses them
  ResIn Rn{};
  fetch_res(z0, Rn);
  __syncthreads();
  faces(z0, true);
  __syncthreads();
  // ---- plane steps z0+1 .. z1+1: residual of plane p-1
  for (int p = z0 + 1; p <= z1 + 1; ++p) {
    const ResIn Rc = Rn;
    cells(p, C);
    if (p <= z1) {
      fetch_cells(p + 1, C);
      fetch_res(p, Rn);
    }
    __syncthreads();
    faces(p, true);
    __syncthreads();
    residual(p - 1, Rc, xm, ym, u_c, a_c);
    xm = xsum_of((p - 1) & 1);
    ym = ysum_of((p - 1) & 1);
    u_c = u_own;
    a_c = a_own;
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
    for (int w = 0; w < OT / WAVE; ++w) {
      s += sRed[0][w];
      m = nanmax(m, sRed[1][w]);
    }
    A.psum[b] = s;
    A.pmax[b] = m;
  }
}

void launch_dc_outer3d(const DcOuter3D& A, int nblk, bool unit, hipStream_t stream) {
  if (unit)
    hipLaunchKernelGGL(k_dc_outer3d<true>, dim3(nblk), dim3(OT), 0, stream, A);
  else
    hipLaunchKernelGGL(k_dc_outer3d<false>, dim3(nblk), dim3(OT), 0, stream, A);
}

}  // namespace cdiff
