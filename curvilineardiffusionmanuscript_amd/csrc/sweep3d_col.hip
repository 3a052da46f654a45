// sweep3d_col.hip -- one fused APT pseudo-iteration on a 3D curvilinear mesh
// over 64 x 8 columns (column3d.h).
//
// The same arithmetic as sweep3d.hip (alpha, the iteration parameters, the
// damped fluxes on the xi, eta and zeta faces, the 3D divergence of
// flux_divergence.jl:119-239, the update of update.jl:2-21 and, on check
// iterations, the residual with q' of residuals.jl:24-47), organised so that
// every cell and face is evaluated once per block and the global loads run
// ahead of the arithmetic.  Plane step p, for p = z0-1 .. z1+1:
//   cells   of the ringed 66 x 10 plane p: u (the BC rule for face ghosts),
//           alpha, theta, dtau -> LDS (planes p-1 and p)
//   faces   q (and q' on check iterations) of the xi and eta faces of plane p
//           and of the zeta faces between planes p-1 and p -> LDS (two
//           planes); q of the column's faces -> memory
//   update  of plane p-1 per column cell; the zeta-neighbour face sums of
//           plane p-2 are carried in registers from the previous update
// The loads of plane p+1's cells and faces are issued during step p, the
// loads of a plane's update one step before it.
#include <limits>

#include "column3d.h"
#include "sweep3d_ops.h"
#include "sweep_args.h"

namespace cdiff {

namespace {

constexpr int CFX = L3X + 1;           // xi faces per ring row (region columns 0..64)
constexpr int CFY = L3Y + 1;           // eta face rows (region rows 0..8)
constexpr int CNX = CFX * R3Y;         // 650
constexpr int CNY = R3X * CFY;         // 594
constexpr int CNF = CNX + CNY + R3;    // faces per plane step
constexpr int CQ = (CNF + L3T - 1) / L3T;

// memory inputs of one region cell at one plane
template <typename T>
struct ColCell {
  T u, rho, fa, ft, fd;
  int kind;  // 0 outside, 1 value, -1 negated value
};

// memory inputs of one column cell's update
template <typename T>
struct ColUpd {
  T cf[9], up, S, fa, fb;  // coefficients, u_prev (or the correction's right-hand side), source, forcing A, B
};

}  // namespace

// PLAIN: no density, frozen parameters, correction right-hand side or forcing
// (the arguments that would hold them are not read)
template <typename T, bool CHECK, bool PLAIN>
__global__ __launch_bounds__(L3T) void sweep3d_col_kernel(Sweep3DArgs<T> A) {
  if (A.ctrl->done) return;
  constexpr int NP = CHECK ? 2 : 1;  // q, and q' on check iterations
  __shared__ T sU[2][R3], sA[2][R3], sT[2][R3];
  __shared__ T sX[NP][2][CNX], sY[NP][2][CNY], sZ[NP][2][R3];
  __shared__ double sRed[2][L3T / WAVE];
  const Geo& g = A.g;
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  Column3D col3;
  if (!column_of(g, A.ntiles, A.chunks, A.stripsI, A.ch, col3)) {
    if (CHECK && tid == 0) {
      A.part_sum[b] = 0.0;
      A.part_max[b] = 0.0;
    }
    return;
  }
  const int x0 = col3.x0, y0 = col3.y0, z0 = col3.z0, z1 = col3.z1;
  const int gx = x0 - 1, gy = y0 - 1;
  const int harm = A.harmonic;
  const T dt = (T)A.ic.dt;

  int rlx[2], rly[2];
  bool rv[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) rv[s] = region_cell(s, tid, rlx[s], rly[s]);
  const int64_t plane = g.plane;
  const int klo = g.lo[2], khi = g.hi[2];

  // ---- per region slot, what does not depend on the plane: the column's
  //      offset, and how its u is formed on domain planes (the BC rule of an
  //      i/j face ghost) -- ghost_u3 with the face known in advance
  enum { C_NONE = 0, C_DOM, C_COPY, C_NEG, C_CONST, C_CELL };
  // the BC table: selected by the face of a ghost column (once per thread,
  // before the plane loop)
  const Ghost gh = A.gh;
  int64_t coff[2], soff[2];
  int ck[2];
  bool wghost[2];  // writes its ghost u on the column's planes
  T cval[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int i = gx + rlx[s], j = gy + rly[s];
    coff[s] = g.off + (int64_t)i + g.pitch * (int64_t)j;
    soff[s] = coff[s];
    ck[s] = C_NONE;
    wghost[s] = false;
    cval[s] = T(0);
    if (!rv[s] || i < g.lo[0] - 1 || i > g.hi[0] + 1 || j < g.lo[1] - 1 || j > g.hi[1] + 1) continue;
    const bool oi = i < g.lo[0] || i > g.hi[0], oj = j < g.lo[1] || j > g.hi[1];
    if (!oi && !oj) {
      ck[s] = C_DOM;
    } else if (oi && oj) {
      ck[s] = C_CELL;  // edge column: never a face ghost
    } else {
      const int f = oi ? (i < g.lo[0] ? 0 : 1) : (j < g.lo[1] ? 2 : 3);
      const int m = pick6(gh.mode, f);
      if (m == G_CONST) {
        ck[s] = C_CONST;
        cval[s] = (T)pick6(gh.val, f);
      } else if (m == G_KEEP) {
        ck[s] = C_CELL;
      } else {
        const int src = pick6(gh.src, f);
        ck[s] = m == G_NEG ? C_NEG : C_COPY;
        soff[s] = g.off + (int64_t)(oi ? src : i) + g.pitch * (int64_t)(oj ? src : j);
      }
      const bool own_i = rlx[s] >= 1 && rlx[s] <= L3X, own_j = rly[s] >= 1 && rly[s] <= L3Y;
      wghost[s] = m != G_KEEP && ((oi && own_j) || (oj && own_i));
    }
  }
  // zeta faces of the domain: their BC rule
  const int zm_lo = gh.mode[4], zm_hi = gh.mode[5];
  const int zs_lo = gh.src[4], zs_hi = gh.src[5];
  const T zv_lo = (T)gh.val[4], zv_hi = (T)gh.val[5];

  // ---- cells: loads of plane p, then alpha / theta / dtau into LDS
  auto fetch_cells = [&](int p, ColCell<T>* C) {
    const bool pdom = p >= klo && p <= khi;
    const bool plo = p == klo - 1;
    const int zm = plo ? zm_lo : zm_hi, zsrc = plo ? zs_lo : zs_hi;
    const T zv = plo ? zv_lo : zv_hi;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      C[s].u = T(0);
      C[s].rho = T(1);
      C[s].fa = C[s].ft = C[s].fd = T(0);
      C[s].kind = 0;
      if (ck[s] == C_NONE || p < klo - 1 || p > khi + 1) continue;
      const int64_t c = coff[s] + plane * p;
      C[s].kind = 1;
      if (pdom) {
        if (ck[s] == C_CONST) {
          C[s].u = cval[s];
        } else {
          C[s].u = A.u_src[(ck[s] == C_COPY || ck[s] == C_NEG) ? soff[s] + plane * p : c];
          if (ck[s] == C_NEG) C[s].kind = -1;
        }
      } else if (ck[s] == C_DOM) {
        // face ghost of a zeta face
        if (zm == G_CONST) {
          C[s].u = zv;
        } else {
          C[s].u = A.u_src[(zm == G_KEEP) ? c : coff[s] + plane * zsrc];
          if (zm == G_NEG) C[s].kind = -1;
        }
      } else {
        C[s].u = A.u_src[c];  // edge or corner
      }
      if constexpr (!PLAIN) {
        if (A.rho) C[s].rho = A.rho[c];
        if (A.frozen_alpha) {
          C[s].fa = A.frozen_alpha[c];
          if (!A.alpha_only) {
            C[s].ft = A.frozen_theta[c];
            C[s].fd = A.frozen_dtau[c];
          }
        }
      }
    }
  };
  T u_own = T(0), a_own = T(0), d_own = T(0);  // column cell of the last plane from `cells`
  auto cells = [&](int p, const ColCell<T>* C) {
    const int buf = p & 1;
    const bool own_k = p >= z0 && p <= z1;
    const bool zghost = (p == klo - 1 && zm_lo != G_KEEP) || (p == khi + 1 && zm_hi != G_KEEP);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (!rv[s]) continue;
      const int r = rly[s] * R3X + rlx[s];
      T u = T(0), a = T(0), th = T(0), dtau = T(0);
      if (C[s].kind != 0) {
        u = C[s].kind < 0 ? -C[s].u : C[s].u;
        if (!PLAIN && A.frozen_alpha) {
          a = C[s].fa;
          if (A.alpha_only) {
            iter_params<T>(A.ic, C[s].rho, a, dtau, th);
          } else {
            th = C[s].ft;
            dtau = C[s].fd;
          }
        } else {
          const T rho = PLAIN ? T(1) : C[s].rho;
          a = alpha_of<T>(A.kap, rho, u);
          iter_params<T>(A.ic, rho, a, dtau, th);
        }
        // ghost values of this iterate on the column's face ghosts
        const bool wz = s == 0 && zghost && ck[s] == C_DOM;
        if ((wghost[s] && own_k) || wz) A.u_dst[coff[s] + plane * p] = u;
      }
      sU[buf][r] = u;
      sA[buf][r] = a;
      sT[buf][r] = th;
      if (s == 0) {
        u_own = u;
        a_own = a;
        d_own = dtau;
      }
    }
  };

  // ---- faces, CQ per thread: xi (t < CNX), eta (t < CNX + CNY), zeta; per
  //      face what does not depend on the plane: its kind, lower region cell,
  //      column offset, whether it is on in i/j, whether this block writes it
  int finfo[CQ];  // bits 0-9 region cell, 10-11 kind (3: none), 12 on in i/j, 13 written here
  int64_t foff[CQ];
#pragma unroll
  for (int q = 0; q < CQ; ++q) {
    const int t = tid + q * L3T;
    int kind = 3, r = 0, i = gx, j = gy;
    if (t < CNX) {
      kind = 0;
      r = (t / CFX) * R3X + t % CFX;
    } else if (t < CNX + CNY) {
      kind = 1;
      r = t - CNX;
    } else if (t < CNF) {
      kind = 2;
      r = t - CNX - CNY;
    }
    const int lx = r % R3X, ly = r / R3X;
    i = gx + lx;
    j = gy + ly;
    bool on = false, wr = false;
    if (kind == 0) {
      on = i >= g.lo[0] - 1 && i <= g.hi[0] && j >= g.lo[1] && j <= g.hi[1];
      wr = ly >= 1 && ly <= L3Y && (lx >= 1 || i == g.lo[0] - 1);
    } else if (kind == 1) {
      on = i >= g.lo[0] && i <= g.hi[0] && j >= g.lo[1] - 1 && j <= g.hi[1];
      wr = lx >= 1 && lx <= L3X && (ly >= 1 || j == g.lo[1] - 1);
    } else if (kind == 2) {
      on = i >= g.lo[0] && i <= g.hi[0] && j >= g.lo[1] && j <= g.hi[1];
      wr = lx >= 1 && lx <= L3X && ly >= 1 && ly <= L3Y;
    }
    finfo[q] = r | (kind << 10) | ((int)on << 12) | ((int)wr << 13);
    foff[q] = g.off + (int64_t)i + g.pitch * (int64_t)j;
  }
  auto fetch_faces = [&](int p, T* Q, bool zf) {
    const bool pdom = p >= klo && p <= khi;
    const bool zon = zf && p - 1 >= klo - 1 && p - 1 <= khi;
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      Q[q] = T(0);
      const int kind = (finfo[q] >> 10) & 3;
      if (!((finfo[q] >> 12) & 1)) continue;
      if (kind == 0 && pdom) Q[q] = A.qx_src[foff[q] + plane * p];
      else if (kind == 1 && pdom) Q[q] = A.qy_src[foff[q] + plane * p];
      else if (kind == 2 && zon) Q[q] = A.qz_src[foff[q] + plane * (p - 1)];
    }
  };
  auto faces = [&](int p, const T* Q, bool zf) {
    const int bc = p & 1, bm = (p + 1) & 1;
    const bool pdom = p >= klo && p <= khi;
    const bool own_p = p >= z0 && p <= z1;
    const int k = p - 1;  // zeta faces between planes k and k+1, stored at k
    const bool zon = zf && k >= klo - 1 && k <= khi;
    const bool own_z = (k >= z0 && k <= z1) || k == klo - 1;
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      const int kind = (finfo[q] >> 10) & 3;
      const int r = finfo[q] & 1023;
      const bool on = (finfo[q] >> 12) & 1, wr = (finfo[q] >> 13) & 1;
      T qn = T(0), qp = T(0);
      if (kind == 0) {
        if (on && pdom) {
          flux3<T>(harm, Q[q], sU[bc][r + 1], sA[bc][r + 1], sT[bc][r + 1], sU[bc][r], sA[bc][r], sT[bc][r], qn, qp);
          if (wr && own_p) A.qx_dst[foff[q] + plane * p] = qn;
        }
        const int idx = (r / R3X) * CFX + r % R3X;
        sX[0][bc][idx] = qn;
        if constexpr (CHECK) sX[NP - 1][bc][idx] = qp;
      } else if (kind == 1) {
        if (on && pdom) {
          flux3<T>(harm, Q[q], sU[bc][r + R3X], sA[bc][r + R3X], sT[bc][r + R3X], sU[bc][r], sA[bc][r], sT[bc][r],
                   qn, qp);
          if (wr && own_p) A.qy_dst[foff[q] + plane * p] = qn;
        }
        sY[0][bc][r] = qn;
        if constexpr (CHECK) sY[NP - 1][bc][r] = qp;
      } else if (kind == 2 && zf) {
        if (on && zon) {
          flux3<T>(harm, Q[q], sU[bc][r], sA[bc][r], sT[bc][r], sU[bm][r], sA[bm][r], sT[bm][r], qn, qp);
          if (wr && own_z) A.qz_dst[foff[q] + plane * k] = qn;
        }
        sZ[0][bm][r] = qn;
        if constexpr (CHECK) sZ[NP - 1][bm][r] = qp;
      }
    }
  };

  // ---- update of a column cell
  const int tx = tid % L3X, ty = tid / L3X;
  const int ci = x0 + tx, cj = y0 + ty;
  const bool col = ci <= g.hi[0] && cj <= g.hi[1];
  const int64_t ccol = g.off + (int64_t)ci + g.pitch * (int64_t)cj;
  const int x_r = (ty + 1) * CFX + tx + 1;  // its xi face in the face planes
  const int y_r = (ty + 1) * R3X + tx + 1;  // its eta and zeta faces
  auto fetch_upd = [&](int k) {
    ColUpd<T> U;
    const bool on = col && k >= z0 && k <= z1;
    const int64_t c = on ? ccol + plane * k : g.off;
    const T* cf0 = A.coef[0];
#pragma unroll
    for (int q = 0; q < 9; ++q) U.cf[q] = on ? cf0[c + A.coef_stride * q] : T(0);
    if constexpr (PLAIN) {
      U.up = on ? A.u_prev[c] : T(0);
      U.S = (on && A.S) ? A.S[c] : T(0);
      U.fa = U.fb = T(0);
    } else {
      U.up = on ? (A.base ? A.base[c] : A.u_prev[c]) : T(0);
      U.S = (on && !A.base && A.S) ? A.S[c] : T(0);
      U.fa = (on && !A.base && A.tzA) ? A.tzA[c] : T(0);
      U.fb = (on && !A.base && A.tzA) ? A.tzB[c] : T(0);
    }
    return U;
  };
  // faces around the column cell from the face planes k (buffer k&1) and k+1
  // and the zeta faces stored at k-1 and k
  auto around = [&](int P, int k, Around<T>& q) {
    const T* X = sX[P][k & 1];
    const T* Y = sY[P][k & 1];
    const T* Zc = sZ[P][k & 1];
    const T* Zm = sZ[P][(k + 1) & 1];
    q.xc = X[x_r];
    q.xW = X[x_r - 1];
    q.xJp = X[x_r + CFX];
    q.xJpW = X[x_r + CFX - 1];
    q.xJm = X[x_r - CFX];
    q.xJmW = X[x_r - CFX - 1];
    q.yc = Y[y_r];
    q.yJm = Y[y_r - R3X];
    q.yE = Y[y_r + 1];
    q.yEJm = Y[y_r - R3X + 1];
    q.yW = Y[y_r - 1];
    q.yWJm = Y[y_r - R3X - 1];
    q.zc = Zc[y_r];
    q.zKm = Zm[y_r];
    q.zE = Zc[y_r + 1];
    q.zEKm = Zm[y_r + 1];
    q.zW = Zc[y_r - 1];
    q.zWKm = Zm[y_r - 1];
    q.zJp = Zc[y_r + R3X];
    q.zJpKm = Zm[y_r + R3X];
    q.zJm = Zc[y_r - R3X];
    q.zJmKm = Zm[y_r - R3X];
  };
  // zeta-neighbour sums of the plane below the next update: X(0)+X(-1), Y(0)+Y(0,-1)
  T xs0 = T(0), xs1 = T(0), ys0 = T(0), ys1 = T(0);
  double acc = 0.0, amax = 0.0;
  auto update = [&](int k, const ColUpd<T>& U, T uc, T ac, T dc) {
    Around<T> q;
    around(0, k, q);
    const T* Xn = sX[0][(k + 1) & 1];
    const T* Yn = sY[0][(k + 1) & 1];
    const T xKpS = Xn[x_r] + Xn[x_r - 1], yKpS = Yn[y_r] + Yn[y_r - R3X];
    if (col) {
      const int64_t c = ccol + plane * k;
      const T D = div3_sums<T>(U.cf, q, xKpS, xs0, yKpS, ys0);
      T un;
      if (!PLAIN && A.base) {
        un = (uc + dc * (U.up - D)) / (T(1) + dc / dt);
      } else {
        T rhs = U.up / dt - D + U.S;
        if (!PLAIN && A.tzA) rhs = rhs + (U.fa + U.fb * ac);
        un = (uc + dc * rhs) / (T(1) + dc / dt);
      }
      A.u_dst[c] = un;
      if constexpr (CHECK) {
        Around<T> pq;
        around(NP - 1, k, pq);
        const T* Pn = sX[NP - 1][(k + 1) & 1];
        const T* Qn = sY[NP - 1][(k + 1) & 1];
        const T Dp = div3_sums<T>(U.cf, pq, Pn[x_r] + Pn[x_r - 1], xs1, Qn[y_r] + Qn[y_r - R3X], ys1);
        T res;
        if (!PLAIN && A.base) {
          res = U.up - un / dt - Dp;
        } else {
          T du = un - U.up;
          if (isapprox_eps<T>(un, U.up)) du = copysign(T(0), du);
          res = -du / dt - Dp + U.S;
          if (!PLAIN && A.tzA) res = res + (U.fa + U.fb * ac);
        }
        const double rd = (double)res;
        acc += rd * rd;
        amax = nanmax(amax, fabs(rd));
      }
    }
    xs0 = q.xc + q.xW;
    ys0 = q.yc + q.yJm;
    if constexpr (CHECK) {
      xs1 = sX[NP - 1][k & 1][x_r] + sX[NP - 1][k & 1][x_r - 1];
      ys1 = sY[NP - 1][k & 1][y_r] + sY[NP - 1][k & 1][y_r - R3X];
    }
  };

  // ---- prologue: plane steps z0-1 (no zeta faces) and z0
  ColCell<T> C[2];
  T Q[CQ];
  fetch_cells(z0 - 1, C);
  fetch_faces(z0 - 1, Q, false);
  cells(z0 - 1, C);
  fetch_cells(z0, C);
  __syncthreads();
  faces(z0 - 1, Q, false);
  fetch_faces(z0, Q, true);
  __syncthreads();
  xs0 = sX[0][(z0 - 1) & 1][x_r] + sX[0][(z0 - 1) & 1][x_r - 1];
  ys0 = sY[0][(z0 - 1) & 1][y_r] + sY[0][(z0 - 1) & 1][y_r - R3X];
  if constexpr (CHECK) {
    xs1 = sX[NP - 1][(z0 - 1) & 1][x_r] + sX[NP - 1][(z0 - 1) & 1][x_r - 1];
    ys1 = sY[NP - 1][(z0 - 1) & 1][y_r] + sY[NP - 1][(z0 - 1) & 1][y_r - R3X];
  }
  cells(z0, C);
  T uc = u_own, ac = a_own, dc = d_own;
  fetch_cells(z0 + 1, C);
  ColUpd<T> U = fetch_upd(z0);
  __syncthreads();
  faces(z0, Q, true);
  fetch_faces(z0 + 1, Q, true);

  // ---- plane steps z0+1 .. z1+1: update of plane p-1
  for (int p = z0 + 1; p <= z1 + 1; ++p) {
    cells(p, C);
    if (p <= z1) fetch_cells(p + 1, C);
    __syncthreads();  // cells of p written; faces and updates of the previous step done
    faces(p, Q, true);
    if (p <= z1) fetch_faces(p + 1, Q, true);
    __syncthreads();  // faces of p written
    update(p - 1, U, uc, ac, dc);
    uc = u_own;
    ac = a_own;
    dc = d_own;
    if (p <= z1) U = fetch_upd(p);
  }

  if (CHECK) {
    acc = wave_sum(acc);
    amax = wave_max(amax);
    if ((tid & (WAVE - 1)) == 0) {
      sRed[0][tid / WAVE] = acc;
      sRed[1][tid / WAVE] = amax;
    }
    __syncthreads();
    if (tid == 0) {
      double s = 0.0, m = 0.0;
      for (int w = 0; w < L3T / WAVE; ++w) {
        s += sRed[0][w];
        m = nanmax(m, sRed[1][w]);
      }
      A.part_sum[b] = s;
      A.part_max[b] = m;
    }
  }
}

template <typename T>
void launch_sweep3d_col(const Sweep3DArgs<T>& A, bool check, int nblocks, hipStream_t s) {
  const bool plain = !A.rho && !A.frozen_alpha && !A.base && !A.tzA;
  const dim3 grid(nblocks), blk(L3T);
  if (plain) {
    if (check)
      hipLaunchKernelGGL((sweep3d_col_kernel<T, true, true>), grid, blk, 0, s, A);
    else
      hipLaunchKernelGGL((sweep3d_col_kernel<T, false, true>), grid, blk, 0, s, A);
  } else {
    if (check)
      hipLaunchKernelGGL((sweep3d_col_kernel<T, true, false>), grid, blk, 0, s, A);
    else
      hipLaunchKernelGGL((sweep3d_col_kernel<T, false, false>), grid, blk, 0, s, A);
  }
}

template void launch_sweep3d_col<double>(const Sweep3DArgs<double>&, bool, int, hipStream_t);
template void launch_sweep3d_col<float>(const Sweep3DArgs<float>&, bool, int, hipStream_t);

}  // namespace cdiff
