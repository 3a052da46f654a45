// sweep2d_pair.hip -- two fused 2D APT pseudo-iterations per launch.
//
// A single iteration reads ~9 and writes 3 arrays per cell; on MI355X that
// traffic, not the arithmetic, bounds it.  This kernel performs iterations
// t and t+1 in one pass (temporal blocking): the block loads its tile with a
// two-cell ring once, runs iteration t on the tile plus a one-cell ring in LDS,
// then iteration t+1 on the tile, and writes only the final u and q.  HBM
// traffic per iteration is roughly halved; the arithmetic grows by the ring
// (~12% at 64 x 8 tiles).
//
// Per-cell arithmetic is the reference's, operation by operation, exactly as
// in sweep2d_tile.hip, so a pair is bit-identical to two single launches.  The
// host launches pairs only where iteration t is not an error-check iteration
// (the exit decision can only come after t+1), and only when every ghost rule
// is local (periodic faces read the far side of the domain, which a tile does
// not hold after iteration t).
//
// Regions, in tile-local coordinates with origin (x0-2, y0-2):
//   C2 = [0, TX+3] x [0, TY+3]   cells loaded (two-cell ring)
//   C1 = [1, TX+2] x [1, TY+2]   cells updated by iteration t (one-cell ring)
//   O  = [2, TX+1] x [2, TY+1]   cells updated by iteration t+1 (owned)
// Faces are stored at the index of their lower cell in C2-shaped arrays.
#include <limits>

#include "cdiff_dev.h"
#include "sweep_args.h"

namespace cdiff {

namespace {

constexpr int PTX = 64;
constexpr int LX = PTX + 4;  // C2 row length

template <typename T>
__device__ __forceinline__ T dmean2(int harmonic, T a, T b, T da0, T da1) {
  if (!harmonic) return T(0.5) * (da0 + da1);
  const T s = a + b;
  if (!(s != T(0))) return T(0);
  return (T(2) * b * b * da0 + T(2) * a * a * da1) / (s * s);
}

template <typename T>
__device__ __forceinline__ bool approx_eps2(T x, T y) {
  if (x == y) return true;
  if (!(isfinite(x) && isfinite(y))) return false;
  const T m = fabs(x) > fabs(y) ? fabs(x) : fabs(y);
  return fabs(x - y) <= std::numeric_limits<T>::epsilon() * m;
}

// Ghost rule of the first halo layer (no periodic rules: the host excludes
// them).  `dom` is the value of the source domain cell, `stored` the cell's own
// stored value.
__device__ __forceinline__ int ghost_face(const Geo& g, int i, int r) {
  if (i < g.lo[0]) return 0;
  if (i > g.hi[0]) return 1;
  if (r < g.lo[1]) return 2;
  return 3;
}

// Cells of C1 handled by a thread: slot 0 is the thread's owned cell
// (tx+2, ty+2); slot 1 is cell `tid` of the C1 ring (the 2(TX+2) + 2 TY cells
// of C1 outside the owned tile), for tid below the ring size.
template <int TY>
__device__ __forceinline__ bool c1_cell(int slot, int tid, int& lx, int& ly) {
  if (slot == 0) {
    lx = tid % PTX + 2;
    ly = tid / PTX + 2;
    return true;
  }
  constexpr int RW = PTX + 2;
  int k = tid;
  if (k < RW) {
    lx = 1 + k;
    ly = 1;
    return true;
  }
  k -= RW;
  if (k < RW) {
    lx = 1 + k;
    ly = TY + 2;
    return true;
  }
  k -= RW;
  if (k < TY) {
    lx = 1;
    ly = 2 + k;
    return true;
  }
  k -= TY;
  if (k < TY) {
    lx = PTX + 2;
    ly = 2 + k;
    return true;
  }
  return false;
}

}  // namespace

template <typename T, bool CHECK, int TY>
__global__ __launch_bounds__(PTX* TY) void sweep2d_pair_kernel(Sweep2DArgs<T> A) {
  // the iteration constants: kernel arguments, or device memory in a
  // captured graph (read field by field: a select of the two blocks would
  // hold them in vector registers and cost a wavefront per SIMD)
  IterConst ic = A.ic;
  if (A.ic_dev) {
    ic.L = __builtin_nontemporal_load(&A.ic_dev->L);
    ic.Vpdtau = __builtin_nontemporal_load(&A.ic_dev->Vpdtau);
    ic.dt = __builtin_nontemporal_load(&A.ic_dev->dt);
  }
  constexpr int LY = TY + 4;
  constexpr int NC2 = LX * LY;
  constexpr int NTHR = PTX * TY;
  constexpr bool F32 = sizeof(T) == 4;
  if (A.ctrl->done) return;
  // LDS (aliases noted): u_old -> b xi faces, alpha_old -> alpha^a,
  // theta_old -> theta^a, dtau_old -> b eta faces
  __shared__ T sU0[NC2], sA0[NC2], sT0[NC2], sD0[NC2];
  __shared__ T sXa[NC2], sYa[NC2], sU1[NC2], sD1[NC2];
  __shared__ T sPX[CHECK ? NC2 : 1], sPY[CHECK ? NC2 : 1];
  __shared__ T sUb[F32 ? NC2 : 1], sDa[F32 ? NC2 : 1];
  __shared__ double sRed[2][NTHR / WAVE];
  T* const sXb = sU0;
  T* const sYb = sD0;
  T* const sA1 = sA0;
  T* const sT1 = sT0;

  const Geo& g = A.g;
  const int tid = threadIdx.x;
  const int nb = gridDim.x, per = nb >> 3, b = blockIdx.x;
  const int tile = (b & 7) * per + (b >> 3);
  if (tile >= A.ntiles) {
    if (CHECK && tid == 0) {
      A.part_sum[b] = 0.0;
      A.part_max[b] = 0.0;
    }
    return;
  }
  const int bx = tile / A.chunks, by = tile - bx * A.chunks;
  const int x0 = g.lo[0] + bx * PTX, y0 = g.lo[1] + by * TY;
  const int gx = x0 - 2, gy = y0 - 2;  // global index of local (0, 0)
  const int harm = A.harmonic;
  const T dt = (T)ic.dt;
  const bool newton = F32 && A.ubar != nullptr;

  // cell (i, r) takes part iff it is a domain cell or a first-layer ghost
  auto live = [&](int i, int r) {
    return i >= g.lo[0] - 1 && i <= g.hi[0] + 1 && r >= g.lo[1] - 1 && r <= g.hi[1] + 1;
  };
  auto in_dom = [&](int i, int r) { return i >= g.lo[0] && i <= g.hi[0] && r >= g.lo[1] && r <= g.hi[1]; };
  auto xface_on = [&](int i, int r) { return i >= g.lo[0] - 1 && i <= g.hi[0] && r >= g.lo[1] && r <= g.hi[1]; };
  auto yface_on = [&](int i, int r) { return i >= g.lo[0] && i <= g.hi[0] && r >= g.lo[1] - 1 && r <= g.hi[1]; };

  // ---- all global loads up front
  constexpr int P1 = (NC2 + NTHR - 1) / NTHR;
  T lu[P1], lr[P1], lfa[P1], lfth[P1], lfdt[P1], lub[P1], lda[P1];
#pragma unroll
  for (int p = 0; p < P1; ++p) {
    const int t = tid + p * NTHR;
    lu[p] = lr[p] = lfa[p] = lfth[p] = lfdt[p] = lub[p] = lda[p] = T(0);
    if (t < NC2) {
      const int i = gx + t % LX, r = gy + t / LX;
      if (live(i, r)) {
        const int64_t c = at(g, i, r, 0);
        T v;
        const bool gi = (i < g.lo[0]) | (i > g.hi[0]);
        const bool gj = (r < g.lo[1]) | (r > g.hi[1]);
        if (gi != gj) {
          const int f = ghost_face(g, i, r);
          const int m = A.gh.mode[f];
          if (m == G_CONST) {
            v = (T)A.gh.val[f];
          } else if (m == G_KEEP) {
            v = A.u_src[c];
          } else {
            const int si = gi ? A.gh.src[f] : i, sr = gi ? r : A.gh.src[f];
            v = A.u_src[at(g, si, sr, 0)];
            if (m == G_NEG) v = -v;
          }
        } else {
          v = A.u_src[c];
        }
        lu[p] = v;
        if (A.rho) lr[p] = A.rho[c];
        if (A.frozen_alpha) {
          lfa[p] = A.frozen_alpha[c];
          if (!A.alpha_only) {
            lfth[p] = A.frozen_theta[c];
            lfdt[p] = A.frozen_dtau[c];
          }
        }
        if (newton) {
          lub[p] = A.ubar[c];
          lda[p] = A.dalpha[c];
        }
      }
    }
  }
  // old fluxes of the iteration-t faces (xi and eta, C2-shaped)
  constexpr int P2 = (2 * NC2 + NTHR - 1) / NTHR;
  T lq[P2];
#pragma unroll
  for (int p = 0; p < P2; ++p) {
    const int t = tid + p * NTHR;
    lq[p] = T(0);
    if (t < 2 * NC2) {
      const int f = t < NC2 ? t : t - NC2;
      const int i = gx + f % LX, r = gy + f / LX;
      if (t < NC2) {
        if (xface_on(i, r) && f % LX <= PTX + 2) lq[p] = A.qx_src[at(g, i, r, 0)];
      } else {
        if (yface_on(i, r) && f / LX <= TY + 2) lq[p] = A.qy_src[at(g, i, r, 0)];
      }
    }
  }
  // cell data of the iteration-t update region C1 (slot 0: the owned cell)
  constexpr int P3 = 2;
  T lg11[P3], lg22[P3], lc12[P3], lac[P3], lbc[P3], lup[P3], lS[P3], ltA[P3], ltB[P3];
#pragma unroll
  for (int p = 0; p < P3; ++p) {
    lg11[p] = lg22[p] = lc12[p] = lac[p] = lbc[p] = lup[p] = lS[p] = ltA[p] = ltB[p] = T(0);
    int lx, ly;
    if (!c1_cell<TY>(p, tid, lx, ly)) continue;
    const int i = gx + lx, r = gy + ly;
    if (in_dom(i, r)) {
      const int64_t c = at(g, i, r, 0);
      lg11[p] = A.g11[c];
      lg22[p] = A.g22[c];
      lc12[p] = A.c12[c];
      lac[p] = A.ac[c];
      lbc[p] = A.bc[c];
      if (A.base) {
        lup[p] = A.base[c];
      } else {
        lup[p] = A.u_prev[c];
        if (A.S) lS[p] = A.S[c];
        if (A.tzA) {
          ltA[p] = A.tzA[c];
          ltB[p] = A.tzB[c];
        }
      }
    }
  }

  // ---- iteration t, cells: alpha, theta, dtau on C2
#pragma unroll
  for (int p = 0; p < P1; ++p) {
    const int t = tid + p * NTHR;
    if (t >= NC2) break;
    const int i = gx + t % LX, r = gy + t / LX;
    T a = T(0), th = T(0), dtau = T(0);
    if (live(i, r)) {
      const T rho = A.rho ? lr[p] : T(1);
      if (A.frozen_alpha) {
        a = lfa[p];
        if (A.alpha_only) {
          iter_params<T>(ic, rho, a, dtau, th);
        } else {
          th = lfth[p];
          dtau = lfdt[p];
        }
      } else {
        a = alpha_of<T>(A.kap, rho, lu[p]);
        iter_params<T>(ic, rho, a, dtau, th);
      }
    }
    sU0[t] = lu[p];
    sA0[t] = a;
    sT0[t] = th;
    sD0[t] = dtau;
    if (newton) {
      sUb[t] = lub[p];
      sDa[t] = lda[p];
    }
  }
  __syncthreads();

  // ---- iteration t, faces (damped flux only: t is never a check iteration)
#pragma unroll
  for (int p = 0; p < P2; ++p) {
    const int t = tid + p * NTHR;
    if (t >= 2 * NC2) break;
    const bool xi = t < NC2;
    const int f = xi ? t : t - NC2;
    const int lx = f % LX, ly = f / LX;
    const int i = gx + lx, r = gy + ly;
    T qn = T(0);
    const bool on = xi ? (xface_on(i, r) && lx <= PTX + 2) : (yface_on(i, r) && ly <= TY + 2);
    if (on) {
      const int c0 = f, c1 = xi ? f + 1 : f + LX;
      const T abar = face_mean<T>(harm, sA0[c0], sA0[c1]);
      const T tbar = face_mean<T>(harm, sT0[c0], sT0[c1]);
      T qp = -abar * (sU0[c1] - sU0[c0]);
      if (newton) qp -= dmean2<T>(harm, sA0[c0], sA0[c1], sDa[c0] * sU0[c0], sDa[c1] * sU0[c1]) * (sUb[c1] - sUb[c0]);
      qn = (lq[p] * tbar + qp) / (T(1) + tbar);
    }
    if (xi)
      sXa[f] = qn;
    else
      sYa[f] = qn;
  }
  __syncthreads();

  // ---- iteration t, update of the domain cells of C1
  auto divergence = [&](const T* X, const T* Y, int f, T g11, T g22, T c12, T ac, T bc) {
    const T x_c = X[f], x_w = X[f - 1], x_n = X[f + LX], x_nw = X[f + LX - 1];
    const T x_s = X[f - LX], x_sw = X[f - LX - 1];
    const T y_c = Y[f], y_s = Y[f - LX], y_e = Y[f + 1], y_es = Y[f - LX + 1];
    const T y_w = Y[f - 1], y_ws = Y[f - LX - 1];
    const T t1 = g11 * (x_c - x_w);
    const T t2 = g22 * (y_c - y_s);
    const T t3 = c12 * ((x_n + x_nw) - (x_s + x_sw));
    const T t4 = c12 * ((y_e + y_es) - (y_w + y_ws));
    const T t5 = ac * (T(0.5) * (x_c + x_w));
    const T t6 = bc * (T(0.5) * (y_c + y_s));
    return t1 + t2 + t3 + t4 + t5 + t6;
  };
#pragma unroll
  for (int p = 0; p < P3; ++p) {
    int lx, ly;
    if (!c1_cell<TY>(p, tid, lx, ly)) continue;
    const int f = ly * LX + lx;
    const int i = gx + lx, r = gy + ly;
    if (in_dom(i, r)) {
      const T D = divergence(sXa, sYa, f, lg11[p], lg22[p], lc12[p], lac[p], lbc[p]);
      const T uc = sU0[f], dtau = sD0[f];
      T un;
      if (A.base) {
        un = (uc + dtau * (lup[p] - D)) / (T(1) + dtau / dt);
      } else {
        T rhs = lup[p] / dt - D + lS[p];
        if (A.tzA) rhs = rhs + (ltA[p] + ltB[p] * sA0[f]);
        un = (uc + dtau * rhs) / (T(1) + dtau / dt);
      }
      sU1[f] = un;
    }
  }
  __syncthreads();
  // ghosts of C1 after iteration t: the BC rule applied to the new domain values
  // (local rules only), then alpha/theta/dtau of the iterate on C1
  T a1v[P3], t1v[P3], d1v[P3], u1v[P3];
#pragma unroll
  for (int p = 0; p < P3; ++p) {
    a1v[p] = t1v[p] = d1v[p] = u1v[p] = T(0);
    int lx, ly;
    if (!c1_cell<TY>(p, tid, lx, ly)) continue;
    const int f = ly * LX + lx;
    const int i = gx + lx, r = gy + ly;
    if (!live(i, r)) continue;
    T v;
    const bool gi = (i < g.lo[0]) | (i > g.hi[0]);
    const bool gj = (r < g.lo[1]) | (r > g.hi[1]);
    if (gi == gj) {
      v = gi ? sU0[f] : sU1[f];  // corner: stored value; domain: updated
    } else {
      const int fc = ghost_face(g, i, r);
      const int m = A.gh.mode[fc];
      if (m == G_CONST) {
        v = (T)A.gh.val[fc];
      } else if (m == G_KEEP) {
        v = sU0[f];
      } else {
        // the source is the adjacent domain cell (copy / negated copy)
        const int fs = gi ? (ly * LX + (A.gh.src[fc] - gx)) : ((A.gh.src[fc] - gy) * LX + lx);
        v = sU1[fs];
        if (m == G_NEG) v = -v;
      }
    }
    u1v[p] = v;
    T a = T(0), th = T(0), dtau = T(0);
    const T rho = A.rho ? A.rho[at(g, i, r, 0)] : T(1);
    if (A.frozen_alpha) {
      a = sA0[f];  // frozen: unchanged by the iteration
      if (A.alpha_only) {
        iter_params<T>(ic, rho, a, dtau, th);
      } else {
        th = sT0[f];
        dtau = sD0[f];
      }
    } else {
      a = alpha_of<T>(A.kap, rho, v);
      iter_params<T>(ic, rho, a, dtau, th);
    }
    a1v[p] = a;
    t1v[p] = th;
    d1v[p] = dtau;
  }
  __syncthreads();  // sA0/sT0 become the iterate's alpha/theta
#pragma unroll
  for (int p = 0; p < P3; ++p) {
    int lx, ly;
    if (!c1_cell<TY>(p, tid, lx, ly)) continue;
    const int f = ly * LX + lx;
    sU1[f] = u1v[p];
    sA1[f] = a1v[p];
    sT1[f] = t1v[p];
    sD1[f] = d1v[p];
  }
  __syncthreads();

  // ---- iteration t+1, faces: xi faces lx in [1, TX+1], ly in [1, TY+2];
  //      eta faces lx in [1, TX+2], ly in [1, TY+1]
  {
    constexpr int NBX = (PTX + 1) * (TY + 2), NBY = (PTX + 2) * (TY + 1);
    constexpr int P4 = (NBX + NBY + NTHR - 1) / NTHR;
#pragma unroll
    for (int p = 0; p < P4; ++p) {
      const int t = tid + p * NTHR;
      if (t >= NBX + NBY) break;
      const bool xi = t < NBX;
      int lx, ly;
      if (xi) {
        lx = 1 + t % (PTX + 1);
        ly = 1 + t / (PTX + 1);
      } else {
        const int u = t - NBX;
        lx = 1 + u % (PTX + 2);
        ly = 1 + u / (PTX + 2);
      }
      const int f = ly * LX + lx;
      const int i = gx + lx, r = gy + ly;
      T qn = T(0), qp = T(0);
      const bool on = xi ? xface_on(i, r) : yface_on(i, r);
      if (on) {
        const int c0 = f, c1 = xi ? f + 1 : f + LX;
        const T abar = face_mean<T>(harm, sA1[c0], sA1[c1]);
        const T tbar = face_mean<T>(harm, sT1[c0], sT1[c1]);
        qp = -abar * (sU1[c1] - sU1[c0]);
        if (newton)
          qp -= dmean2<T>(harm, sA1[c0], sA1[c1], sDa[c0] * sU1[c0], sDa[c1] * sU1[c1]) * (sUb[c1] - sUb[c0]);
        const T qo = xi ? sXa[f] : sYa[f];
        qn = (qo * tbar + qp) / (T(1) + tbar);
      }
      if (xi) {
        sXb[f] = qn;
        if (CHECK) sPX[f] = qp;
      } else {
        sYb[f] = qn;
        if (CHECK) sPY[f] = qp;
      }
    }
  }
  __syncthreads();

  // ---- iteration t+1, update of the owned cells; writes
  const int tx = tid % PTX, tyy = tid / PTX;
  const int lx = tx + 2, ly = tyy + 2;
  const int f = ly * LX + lx;
  const int i = x0 + tx, j = y0 + tyy;
  double sq = 0.0, mx = 0.0;
  const T g11 = lg11[0], g22 = lg22[0], c12 = lc12[0], ac = lac[0], bc = lbc[0], up = lup[0], S = lS[0];
  const T tA = ltA[0], tB = ltB[0];
  const bool owned = i <= g.hi[0] && j <= g.hi[1];
  if (owned) {
    const int64_t c = at(g, i, j, 0);
    const T D = divergence(sXb, sYb, f, g11, g22, c12, ac, bc);
    const T uc = sU1[f], dtau = sD1[f];
    T un, fz = T(0);
    if (A.base) {
      un = (uc + dtau * (up - D)) / (T(1) + dtau / dt);
    } else {
      if (A.tzA) fz = tA + tB * sA1[f];
      T rhs = up / dt - D + S;
      if (A.tzA) rhs = rhs + fz;
      un = (uc + dtau * rhs) / (T(1) + dtau / dt);
    }
    A.u_dst[c] = un;
    A.qx_dst[c] = sXb[f];
    A.qy_dst[c] = sYb[f];
    if (CHECK) {
      // the iterate before iteration t+1, for the epilogue and diagnostics
      if (A.u_pre) A.u_pre[c] = uc;
      const T Dp = divergence(sPX, sPY, f, g11, g22, c12, ac, bc);
      T res;
      if (A.base) {
        res = up - un / dt - Dp;
      } else {
        T du = un - up;
        if (approx_eps2<T>(un, up)) du = copysign(T(0), du);
        res = -du / dt - Dp + S;
        if (A.tzA) res = res + fz;
      }
      const double rd = (double)res;
      sq = rd * rd;
      mx = fabs(rd);
      if (rd != rd) mx = rd;
    }
  }
  // boundary faces without an owning cell, and the ghosts of iteration t+1
  // (the BC of the iterate after iteration t), for this tile's rows/columns
  if (i == g.lo[0] && j <= g.hi[1]) {
    A.qx_dst[at(g, i - 1, j, 0)] = sXb[f - 1];
    if (A.gh.mode[0] != G_KEEP) A.u_dst[at(g, i - 1, j, 0)] = sU1[f - 1];
  }
  if (i == g.hi[0] && j <= g.hi[1] && A.gh.mode[1] != G_KEEP) A.u_dst[at(g, i + 1, j, 0)] = sU1[f + 1];
  if (j == g.lo[1] && i <= g.hi[0]) {
    A.qy_dst[at(g, i, j - 1, 0)] = sYb[f - LX];
    if (A.gh.mode[2] != G_KEEP) A.u_dst[at(g, i, j - 1, 0)] = sU1[f - LX];
  }
  if (j == g.hi[1] && i <= g.hi[0] && A.gh.mode[3] != G_KEEP) A.u_dst[at(g, i, j + 1, 0)] = sU1[f + LX];
  if (CHECK) {
    sq = wave_sum(sq);
    mx = wave_max(mx);
    if ((tid & (WAVE - 1)) == 0) {
      sRed[0][tid / WAVE] = sq;
      sRed[1][tid / WAVE] = mx;
    }
    __syncthreads();
    if (tid == 0) {
      double s = 0.0, m = 0.0;
      for (int w = 0; w < NTHR / WAVE; ++w) {
        s += sRed[0][w];
        m = nanmax(m, sRed[1][w]);
      }
      A.part_sum[b] = s;
      A.part_max[b] = m;
    }
  }
}

template <typename T>
void launch_sweep2d_pair(const Sweep2DArgs<T>& A, bool check, int nblocks, hipStream_t s) {
  constexpr int TY = 8;
  if (check)
    hipLaunchKernelGGL((sweep2d_pair_kernel<T, true, TY>), dim3(nblocks), dim3(PTX * TY), 0, s, A);
  else
    hipLaunchKernelGGL((sweep2d_pair_kernel<T, false, TY>), dim3(nblocks), dim3(PTX * TY), 0, s, A);
}

template void launch_sweep2d_pair<double>(const Sweep2DArgs<double>&, bool, int, hipStream_t);
template void launch_sweep2d_pair<float>(const Sweep2DArgs<float>&, bool, int, hipStream_t);

}  // namespace cdiff
