// sweep2d_tile.hip -- the fused 2D APT pseudo-iteration as an LDS-tiled kernel.
//
// Same arithmetic as sweep2d.hip (operation by operation, so the FP64 result
// is bit-identical); different decomposition.  A 256-thread block owns a
// 64 x 4 tile of cells and works in three phases:
//   1. cells   the 66 x 6 tile plus its one-cell ring: u (ghosts synthesised
//              from the BC rules), alpha, theta, dtau -> LDS
//   2. faces   the 65 x 6 xi faces and 66 x 5 eta faces around the tile: new
//              damped flux q and plain flux q' from the LDS cells and the old
//              q -> LDS (zero outside expand_lower(domain, axis))
//   3. update  per owned cell: divergence from 12 LDS faces, u update,
//              residual on error-check iterations; write u, q
// All global loads of a phase are independent, so a block keeps ~30 loads per
// thread in flight; the ring re-reads of neighbouring tiles are served by L2
// (tiles are ordered so that vertical neighbours share an XCD).
#include <limits>

#include "cdiff_dev.h"
#include "sweep_args.h"

namespace cdiff {

namespace {

constexpr int TX = 64;
constexpr int CX = TX + 2;   // cells per ringed row
constexpr int FXN = TX + 1;  // xi faces per row (x0-1 .. x0+63)
constexpr int FYN = TX + 2;  // eta faces per row (x0-1 .. x0+64)

template <typename T>
__device__ __forceinline__ T ghost_u_t(const Sweep2DArgs<T>& A, int i, int r) {
  const Geo& g = A.g;
  const bool gi = (i < g.lo[0]) | (i > g.hi[0]);
  const bool gj = (r < g.lo[1]) | (r > g.hi[1]);
  if (gi == gj) return A.u_src[at(g, i, r, 0)];
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
  const T v = A.u_src[at(g, si, sr, 0)];
  return m == G_NEG ? -v : v;
}

// Change of the face mean for cell changes (da0, da1) of its two conductivities:
// harmonic 2ab/(a+b) -> (2b^2 da0 + 2a^2 da1)/(a+b)^2, arithmetic -> (da0+da1)/2.
template <typename T>
__device__ __forceinline__ T dmean(int harmonic, T a, T b, T da0, T da1) {
  if (!harmonic) return T(0.5) * (da0 + da1);
  const T s = a + b;
  if (!(s != T(0))) return T(0);
  return (T(2) * b * b * da0 + T(2) * a * a * da1) / (s * s);
}

template <typename T>
__device__ __forceinline__ bool approx_eps(T x, T y) {
  if (x == y) return true;
  if (!(isfinite(x) && isfinite(y))) return false;
  const T m = fabs(x) > fabs(y) ? fabs(x) : fabs(y);
  return fabs(x - y) <= std::numeric_limits<T>::epsilon() * m;
}

}  // namespace

template <typename T, bool CHECK, int TY>
__global__ __launch_bounds__(TX* TY) void sweep2d_tile_kernel(Sweep2DArgs<T> A) {
  // the iteration constants: kernel arguments, or device memory in a
  // captured graph (read field by field: a select of the two blocks would
  // hold them in vector registers and cost a wavefront per SIMD)
  IterConst ic = A.ic;
  if (A.ic_dev) {
    ic.L = __builtin_nontemporal_load(&A.ic_dev->L);
    ic.Vpdtau = __builtin_nontemporal_load(&A.ic_dev->Vpdtau);
    ic.dt = __builtin_nontemporal_load(&A.ic_dev->dt);
  }
  constexpr int CY = TY + 2;             // ringed rows
  constexpr int NCELL = CX * CY;         // cells with ring
  constexpr int NFX = FXN * CY;          // xi faces
  constexpr int NFY = FYN * (TY + 1);    // eta faces
  constexpr int NTHR = TX * TY;
  if (A.ctrl->done) return;
  __shared__ T sU[NCELL], sA[NCELL], sT[NCELL], sD[NCELL];
  constexpr int NNEWT = sizeof(T) == 4 ? NCELL : 1;  // the Newton correction runs in FP32 only
  __shared__ T sUb[NNEWT], sDa[NNEWT];
  __shared__ T sX[NFX], sY[NFY];
  __shared__ T sPX[CHECK ? NFX : 1], sPY[CHECK ? NFY : 1];
  __shared__ double sRed[2][NTHR / WAVE];
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
  // tiles run down the j direction first, so consecutive tiles are vertical neighbours
  const int ty_n = A.chunks;  // tiles along j
  const int bx = tile / ty_n, by = tile - bx * ty_n;
  const int x0 = g.lo[0] + bx * TX, y0 = g.lo[1] + by * TY;
  const int harm = A.harmonic;
  const T dt = (T)ic.dt;

  // ---- all global loads up front: one memory round trip per block
  constexpr int P1 = (NCELL + NTHR - 1) / NTHR;        // 2 cells per thread
  constexpr int P2 = (NFX + NFY + NTHR - 1) / NTHR;    // 3 faces per thread
  T lu[P1], lr[P1], lfa[P1], lfth[P1], lfdt[P1], lub[P1], lda[P1];
  const bool newton = sizeof(T) == 4 && A.ubar != nullptr;
#pragma unroll
  for (int p = 0; p < P1; ++p) {
    const int t = tid + p * NTHR;
    lu[p] = lr[p] = lfa[p] = lfth[p] = lfdt[p] = lub[p] = lda[p] = T(0);
    if (t < NCELL) {
      const int i = x0 - 1 + t % CX, r = y0 - 1 + t / CX;
      if (i <= g.hi[0] + 1 && r <= g.hi[1] + 1) {
        const int64_t c = at(g, i, r, 0);
        lu[p] = ghost_u_t(A, i, r);
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
  T lq[P2];
#pragma unroll
  for (int p = 0; p < P2; ++p) {
    const int t = tid + p * NTHR;
    lq[p] = T(0);
    if (t < NFX) {
      const int i = x0 - 1 + t % FXN, r = y0 - 1 + t / FXN;
      if (i >= g.lo[0] - 1 && i <= g.hi[0] && r >= g.lo[1] && r <= g.hi[1]) lq[p] = A.qx_src[at(g, i, r, 0)];
    } else if (t < NFX + NFY) {
      const int u = t - NFX;
      const int i = x0 - 1 + u % FYN, r = y0 - 1 + u / FYN;
      if (i >= g.lo[0] && i <= g.hi[0] && r >= g.lo[1] - 1 && r <= g.hi[1]) lq[p] = A.qy_src[at(g, i, r, 0)];
    }
  }
  const int tx = tid % TX, tyy = tid / TX;
  const int ci = x0 + tx, cj = y0 + tyy;
  const bool owned = ci <= g.hi[0] && cj <= g.hi[1];
  T g11 = T(0), g22 = T(0), c12 = T(0), ac = T(0), bc = T(0), up = T(0), S = T(0), tA = T(0), tB = T(0);
  int64_t cown = 0;
  if (owned) {
    cown = at(g, ci, cj, 0);
    g11 = A.g11[cown];
    g22 = A.g22[cown];
    c12 = A.c12[cown];
    ac = A.ac[cown];
    bc = A.bc[cown];
    if (A.base) {
      up = A.base[cown];
    } else {
      up = A.u_prev[cown];
      if (A.S) S = A.S[cown];
      if (A.tzA) {
        tA = A.tzA[cown];
        tB = A.tzB[cown];
      }
    }
  }

  // ---- phase 1: cells of the ringed tile
#pragma unroll
  for (int p = 0; p < P1; ++p) {
    const int t = tid + p * NTHR;
    if (t >= NCELL) break;
    const int cx = t % CX, cy = t / CX;
    const int i = x0 - 1 + cx, r = y0 - 1 + cy;
    T u = lu[p], a = T(0), th = T(0), dtau = T(0);
    if (i <= g.hi[0] + 1 && r <= g.hi[1] + 1) {
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
        a = alpha_of<T>(A.kap, rho, u);
        iter_params<T>(ic, rho, a, dtau, th);
      }
      // first-layer ghosts of this tile's owned rows/columns go to the new iterate
      const bool gi = (i == g.lo[0] - 1) | (i == g.hi[0] + 1);
      const bool gj = (r == g.lo[1] - 1) | (r == g.hi[1] + 1);
      const bool own_r = (cy >= 1) & (cy <= TY) & (r >= g.lo[1]) & (r <= g.hi[1]);
      const bool own_i = (cx >= 1) & (cx <= TX) & (i >= g.lo[0]) & (i <= g.hi[0]);
      if (gi && !gj && own_r && A.gh.mode[i < g.lo[0] ? 0 : 1] != G_KEEP) A.u_dst[at(g, i, r, 0)] = u;
      if (gj && !gi && own_i && A.gh.mode[r < g.lo[1] ? 2 : 3] != G_KEEP) A.u_dst[at(g, i, r, 0)] = u;
    }
    sU[t] = u;
    sA[t] = a;
    sT[t] = th;
    sD[t] = dtau;
    if (newton) {
      sUb[t] = lub[p];
      sDa[t] = lda[p];
    }
  }
  __syncthreads();

  // ---- phase 2: faces
#pragma unroll
  for (int p = 0; p < P2; ++p) {
    const int t = tid + p * NTHR;
    if (t < NFX) {
      const int fx = t % FXN, cy = t / FXN;  // face between tile cells (fx, cy) and (fx+1, cy)
      const int i = x0 - 1 + fx, r = y0 - 1 + cy;
      T qn = T(0), qp = T(0);
      if (i >= g.lo[0] - 1 && i <= g.hi[0] && r >= g.lo[1] && r <= g.hi[1]) {
        const int c0 = cy * CX + fx, c1 = c0 + 1;
        const T abar = face_mean<T>(harm, sA[c0], sA[c1]);
        const T tbar = face_mean<T>(harm, sT[c0], sT[c1]);
        const T du = sU[c1] - sU[c0];
        qp = -abar * du;
        if (newton) qp -= dmean<T>(harm, sA[c0], sA[c1], sDa[c0] * sU[c0], sDa[c1] * sU[c1]) * (sUb[c1] - sUb[c0]);
        qn = (lq[p] * tbar + qp) / (T(1) + tbar);
        // the xi face left of the domain has no owning cell: written here
        if (i == g.lo[0] - 1 && cy >= 1 && cy <= TY) A.qx_dst[at(g, i, r, 0)] = qn;
      }
      sX[t] = qn;
      if (CHECK) sPX[t] = qp;
    } else if (t < NFX + NFY) {
      const int u = t - NFX;
      const int fx = u % FYN, fy = u / FYN;  // face between tile cells (fx, fy) and (fx, fy+1)
      const int i = x0 - 1 + fx, r = y0 - 1 + fy;
      T qn = T(0), qp = T(0);
      if (i >= g.lo[0] && i <= g.hi[0] && r >= g.lo[1] - 1 && r <= g.hi[1]) {
        const int c0 = fy * CX + fx, c1 = c0 + CX;
        const T abar = face_mean<T>(harm, sA[c0], sA[c1]);
        const T tbar = face_mean<T>(harm, sT[c0], sT[c1]);
        const T du = sU[c1] - sU[c0];
        qp = -abar * du;
        if (newton) qp -= dmean<T>(harm, sA[c0], sA[c1], sDa[c0] * sU[c0], sDa[c1] * sU[c1]) * (sUb[c1] - sUb[c0]);
        qn = (lq[p] * tbar + qp) / (T(1) + tbar);
        // the eta face below the domain has no owning cell: written here
        if (r == g.lo[1] - 1 && fx >= 1 && fx <= TX) A.qy_dst[at(g, i, r, 0)] = qn;
      }
      sY[u] = qn;
      if (CHECK) sPY[u] = qp;
    }
  }
  __syncthreads();

  // ---- phase 3: update of the owned cells
  double sq = 0.0, mx = 0.0;
  if (owned) {
    const int64_t c = cown;
    const int cc = (tyy + 1) * CX + (tx + 1);
    // xi faces: own (tx+1, tyy+1); west (tx, tyy+1); north/south rows tyy+2 / tyy
    const int xo = (tyy + 1) * FXN + (tx + 1);
    // eta faces: own (tx+1, tyy+1); south (tx+1, tyy); east/west columns tx+2 / tx
    const int yo = (tyy + 1) * FYN + (tx + 1);
    auto div = [&](const T* X, const T* Y) {
      const T x_c = X[xo], x_w = X[xo - 1], x_n = X[xo + FXN], x_nw = X[xo + FXN - 1];
      const T x_s = X[xo - FXN], x_sw = X[xo - FXN - 1];
      const T y_c = Y[yo], y_s = Y[yo - FYN], y_e = Y[yo + 1], y_es = Y[yo - FYN + 1];
      const T y_w = Y[yo - 1], y_ws = Y[yo - FYN - 1];
      const T t1 = g11 * (x_c - x_w);
      const T t2 = g22 * (y_c - y_s);
      const T t3 = c12 * ((x_n + x_nw) - (x_s + x_sw));
      const T t4 = c12 * ((y_e + y_es) - (y_w + y_ws));
      const T t5 = ac * (T(0.5) * (x_c + x_w));
      const T t6 = bc * (T(0.5) * (y_c + y_s));
      return t1 + t2 + t3 + t4 + t5 + t6;
    };
    const T D = div(sX, sY);
    const T uc = sU[cc], dtau = sD[cc];
    T un, f = T(0);
    if (A.base) {
      un = (uc + dtau * (up - D)) / (T(1) + dtau / dt);
    } else {
      if (A.tzA) f = tA + tB * sA[cc];
      T rhs = up / dt - D + S;
      if (A.tzA) rhs = rhs + f;
      un = (uc + dtau * rhs) / (T(1) + dtau / dt);
    }
    A.u_dst[c] = un;
    A.qx_dst[c] = sX[xo];
    A.qy_dst[c] = sY[yo];
    if (CHECK) {
      const T Dp = div(sPX, sPY);
      T res;
      if (A.base) {
        res = up - un / dt - Dp;
      } else {
        T du = un - up;
        if (approx_eps<T>(un, up)) du = copysign(T(0), du);
        res = -du / dt - Dp + S;
        if (A.tzA) res = res + f;
      }
      const double rd = (double)res;
      sq = rd * rd;
      mx = fabs(rd);
      if (rd != rd) mx = rd;
    }
  }
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

template <typename T, int TY>
static void launch_ty(const Sweep2DArgs<T>& A, bool check, int nblocks, hipStream_t s) {
  if (check)
    hipLaunchKernelGGL((sweep2d_tile_kernel<T, true, TY>), dim3(nblocks), dim3(TX * TY), 0, s, A);
  else
    hipLaunchKernelGGL((sweep2d_tile_kernel<T, false, TY>), dim3(nblocks), dim3(TX * TY), 0, s, A);
}

template <typename T>
void launch_sweep2d_tile(const Sweep2DArgs<T>& A, bool check, int nblocks, hipStream_t s) {
  switch (A.ch) {  // rows per tile
    case 4: launch_ty<T, 4>(A, check, nblocks, s); break;
    case 16: launch_ty<T, 16>(A, check, nblocks, s); break;
    default: launch_ty<T, 8>(A, check, nblocks, s); break;
  }
}

template void launch_sweep2d_tile<double>(const Sweep2DArgs<double>&, bool, int, hipStream_t);
template void launch_sweep2d_tile<float>(const Sweep2DArgs<float>&, bool, int, hipStream_t);

}  // namespace cdiff
