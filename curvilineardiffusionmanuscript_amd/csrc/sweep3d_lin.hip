// sweep3d_lin.hip -- the FP32 correction sweep of the 3D defect-correction
// step in coefficient form (see sweep2d_lin.hip for the 2D version and the
// derivation):
//
//   face f (xi, eta, zeta) between cells 0 and 1:  q'_f = P_f e_0 - M_f e_1,
//                                                   q_f <- q'_f + W_f (q_f - q'_f)
//   cell: D = m0 x_c + m1 x_w + m2 y_c + m3 y_s + m4 z_c + m5 z_s
//           + cxe (X_eta + Y_xi) + cxz (X_zeta + Z_xi) + cez (Y_zeta + Z_eta)
//         e <- e + C (r - D - e/dt)
//
// with the cross-term groups of flux_divergence.jl:119-239:
//   X_eta  = (X(0,1,0) + X(-1,1,0)) - (X(0,-1,0) + X(-1,-1,0))
//   X_zeta = (X(0,0,1) + X(-1,0,1)) - (X(0,0,-1) + X(-1,0,-1))
//   Y_xi   = (Y(1,0,0) + Y(1,-1,0)) - (Y(-1,0,0) + Y(-1,-1,0))
//   Y_zeta = (Y(0,0,1) + Y(0,-1,1)) - (Y(0,0,-1) + Y(0,-1,-1))
//   Z_eta  = (Z(0,1,0) + Z(0,1,-1)) - (Z(0,-1,0) + Z(0,-1,-1))
//   Z_xi   = (Z(1,0,0) + Z(1,0,-1)) - (Z(-1,0,0) + Z(-1,0,-1))
//
// Decomposition: a 512-thread block owns a 64 x 8 column of cells and marches
// up a chunk of k planes.  Per plane it holds the ringed 66 x 10 plane of e and
// the xi/eta faces of three consecutive planes and the zeta faces of two in
// LDS: plane s is updated once the faces of plane s+1 (and the zeta faces
// between s and s+1) exist.  Every face is computed once per block; the ring
// of the plane is recomputed by the neighbouring blocks.
#include "cdiff_dev.h"
#include "column3d.h"
#include "sweep_args.h"

namespace cdiff {

template <bool CHECK>
__global__ __launch_bounds__(L3T) void sweep3d_lin_kernel(Lin3DArgs A) {
  if (A.ctrl->done) return;
  __shared__ float sE[2][R3], sX[3][R3], sY[3][R3], sZ[2][R3];
  __shared__ float sPX[CHECK ? 3 : 1][CHECK ? R3 : 1], sPY[CHECK ? 3 : 1][CHECK ? R3 : 1];
  __shared__ float sPZ[CHECK ? 2 : 1][CHECK ? R3 : 1];
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
  const float inv_dt = A.inv_dt;

  int rlx[2], rly[2];
  bool rv[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) rv[p] = region_cell(p, tid, rlx[p], rly[p]);
  auto in_dom = [&](int i, int j, int k) {
    return i >= g.lo[0] && i <= g.hi[0] && j >= g.lo[1] && j <= g.hi[1] && k >= g.lo[2] && k <= g.hi[2];
  };
  auto face_on = [&](int ax, int i, int j, int k) {
    const int lo0 = g.lo[0] - (ax == 0), lo1 = g.lo[1] - (ax == 1), lo2 = g.lo[2] - (ax == 2);
    return i >= lo0 && i <= g.hi[0] && j >= lo1 && j <= g.hi[1] && k >= lo2 && k <= g.hi[2];
  };

  // Everything a plane step needs from memory, fetched one step ahead so the
  // loads are in flight while the previous plane is updated: e of plane k,
  // the xi/eta faces of plane k and the zeta face below it (between planes
  // k-1 and k, stored at plane k-1): P, M, W and the old flux.
  struct Fetch {
    float e, f[3][4];
  };
  auto fetch = [&](int k, Fetch* F, bool zf = true) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      F[p].e = 0.f;
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) F[p].f[ax][0] = F[p].f[ax][1] = F[p].f[ax][2] = F[p].f[ax][3] = 0.f;
      if (!rv[p]) continue;
      const int i = gx + rlx[p], j = gy + rly[p];
      if (!A.zero_src && in_dom(i, j, k)) F[p].e = A.e_src[at(g, i, j, k)];
      if (rlx[p] < R3X - 1 && face_on(0, i, j, k)) {
        const int64_t c = at(g, i, j, k);
        F[p].f[0][0] = A.f[0 * A.stride + c];
        F[p].f[0][1] = A.f[1 * A.stride + c];
        F[p].f[0][2] = A.f[2 * A.stride + c];
        F[p].f[0][3] = A.zero_src ? 0.f : A.q_src[0][c];
      }
      if (rly[p] < R3Y - 1 && face_on(1, i, j, k)) {
        const int64_t c = at(g, i, j, k);
        F[p].f[1][0] = A.f[3 * A.stride + c];
        F[p].f[1][1] = A.f[4 * A.stride + c];
        F[p].f[1][2] = A.f[5 * A.stride + c];
        F[p].f[1][3] = A.zero_src ? 0.f : A.q_src[1][c];
      }
      if (zf && face_on(2, i, j, k - 1)) {
        const int64_t c = at(g, i, j, k - 1);
        F[p].f[2][0] = A.f[6 * A.stride + c];
        F[p].f[2][1] = A.f[7 * A.stride + c];
        F[p].f[2][2] = A.f[8 * A.stride + c];
        F[p].f[2][3] = A.zero_src ? 0.f : A.q_src[2][c];
      }
    }
  };
  // plane step k: e of plane k into LDS; then the xi/eta faces of plane k and
  // the zeta faces between planes k-1 and k (e of k-1 is in the other buffer)
  auto store_e = [&](int k, const Fetch* F) {
#pragma unroll
    for (int p = 0; p < 2; ++p)
      if (rv[p]) sE[k & 1][rly[p] * R3X + rlx[p]] = F[p].e;
  };
  auto faces = [&](int k, const Fetch* F, bool zf = true) {
    const float* E = sE[k & 1];
    const float* Em = sE[(k + 1) & 1];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      if (!rv[p]) continue;
      const int r = rly[p] * R3X + rlx[p];
      const float ec = E[r];
      const float px = F[p].f[0][0] * ec - F[p].f[0][1] * (rlx[p] < R3X - 1 ? E[r + 1] : 0.f);
      const float py = F[p].f[1][0] * ec - F[p].f[1][1] * (rly[p] < R3Y - 1 ? E[r + R3X] : 0.f);
      const int s = (k + 3) % 3, z = (k + 1) & 1;  // slots of plane k / of the zeta face k-1
      sX[s][r] = px + F[p].f[0][2] * (F[p].f[0][3] - px);
      sY[s][r] = py + F[p].f[1][2] * (F[p].f[1][3] - py);
      if (CHECK) {
        sPX[s][r] = px;
        sPY[s][r] = py;
      }
      if (zf) {
        const float pz = F[p].f[2][0] * Em[r] - F[p].f[2][1] * ec;
        sZ[z][r] = pz + F[p].f[2][2] * (F[p].f[2][3] - pz);
        if (CHECK) sPZ[z][r] = pz;
      }
    }
  };
  // per owned cell: the update data of plane k
  const int tx = tid % L3X, ty = tid / L3X;
  const int i = x0 + tx, j = y0 + ty;
  const bool col = i <= g.hi[0] && j <= g.hi[1];
  const int r0 = (ty + 1) * R3X + (tx + 1);
  struct Upd {
    float rb, cc, m[9];
  };
  auto fetch_upd = [&](int k, Upd& U) {
    if (!col || k > z1) return;
    const int64_t c = at(g, i, j, k);
    U.rb = A.base[c];
    U.cc = A.cc[c];
#pragma unroll
    for (int q = 0; q < 9; ++q) U.m[q] = A.m[q * A.stride + c];
  };

  // ---- prologue: plane steps z0-1, z0, z0+1 (the zeta faces below z0-1 are
  //      not needed)
  Fetch F[2];
  fetch(z0 - 1, F, false);
  store_e(z0 - 1, F);
  __syncthreads();
  faces(z0 - 1, F, false);
  fetch(z0, F);
  __syncthreads();
  store_e(z0, F);
  __syncthreads();
  faces(z0, F);
  fetch(z0 + 1, F);
  Upd U{}, Un{};
  fetch_upd(z0, U);
  __syncthreads();
  store_e(z0 + 1, F);
  __syncthreads();
  faces(z0 + 1, F);
  __syncthreads();

  double sq = 0.0, smax = 0.0;
  for (int k = z0; k <= z1; ++k) {
    // loads of the next plane step and of the next update, in flight during this update
    if (k < z1) {
      fetch(k + 2, F);
      fetch_upd(k + 1, Un);
    }
    // ---- update of plane k (faces of k-1..k+1, zeta faces k-1, k exist)
    const int sm = (k + 2) % 3, sc = (k + 3) % 3, sp = (k + 4) % 3;
    const int zc = (k + 2) & 1, zm = (k + 1) & 1;
    if (col) {
      const int64_t c = at(g, i, j, k);
      auto div = [&](const float* Xm, const float* Xc, const float* Xp, const float* Ym, const float* Yc,
                     const float* Yp, const float* Zm, const float* Zc) {
        const float xc = Xc[r0], xw = Xc[r0 - 1], yc = Yc[r0], ys = Yc[r0 - R3X], zcv = Zc[r0], zs = Zm[r0];
        const float x_eta = (Xc[r0 + R3X] + Xc[r0 + R3X - 1]) - (Xc[r0 - R3X] + Xc[r0 - R3X - 1]);
        const float x_zeta = (Xp[r0] + Xp[r0 - 1]) - (Xm[r0] + Xm[r0 - 1]);
        const float y_xi = (Yc[r0 + 1] + Yc[r0 - R3X + 1]) - (Yc[r0 - 1] + Yc[r0 - R3X - 1]);
        const float y_zeta = (Yp[r0] + Yp[r0 - R3X]) - (Ym[r0] + Ym[r0 - R3X]);
        const float z_eta = (Zc[r0 + R3X] + Zm[r0 + R3X]) - (Zc[r0 - R3X] + Zm[r0 - R3X]);
        const float z_xi = (Zc[r0 + 1] + Zm[r0 + 1]) - (Zc[r0 - 1] + Zm[r0 - 1]);
        return U.m[0] * xc + U.m[1] * xw + U.m[2] * yc + U.m[3] * ys + U.m[4] * zcv + U.m[5] * zs +
               U.m[6] * (x_eta + y_xi) + U.m[7] * (x_zeta + z_xi) + U.m[8] * (y_zeta + z_eta);
      };
      const float D = div(sX[sm], sX[sc], sX[sp], sY[sm], sY[sc], sY[sp], sZ[zm], sZ[zc]);
      const float e = sE[k & 1][r0];
      const float en = e + U.cc * (U.rb - D - e * inv_dt);
      A.e_dst[c] = en;
      A.q_dst[0][c] = sX[sc][r0];
      A.q_dst[1][c] = sY[sc][r0];
      A.q_dst[2][c] = sZ[zc][r0];
      if (CHECK) {
        const float Dp = div(sPX[sm], sPX[sc], sPX[sp], sPY[sm], sPY[sc], sPY[sp], sPZ[zm], sPZ[zc]);
        const double res = (double)(U.rb - en * inv_dt - Dp);
        sq += res * res;
        smax = nanmax(smax, fabs(res));
      }
      // boundary faces below the domain that no cell owns
      if (i == g.lo[0]) A.q_dst[0][at(g, i - 1, j, k)] = sX[sc][r0 - 1];
      if (j == g.lo[1]) A.q_dst[1][at(g, i, j - 1, k)] = sY[sc][r0 - R3X];
      if (k == g.lo[2]) A.q_dst[2][at(g, i, j, k - 1)] = sZ[zm][r0];
    }
    if (k == z1) break;
    U = Un;
    __syncthreads();  // plane k's faces and e are read
    // ---- plane step k+2
    store_e(k + 2, F);
    __syncthreads();
    faces(k + 2, F);
    __syncthreads();
  }
  if (CHECK) {
    sq = wave_sum(sq);
    smax = wave_max(smax);
    if ((tid & (WAVE - 1)) == 0) {
      sRed[0][tid / WAVE] = sq;
      sRed[1][tid / WAVE] = smax;
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

void launch_sweep3d_lin(const Lin3DArgs& A, bool check, int nblocks, hipStream_t s) {
  if (check)
    hipLaunchKernelGGL(sweep3d_lin_kernel<true>, dim3(nblocks), dim3(L3T), 0, s, A);
  else
    hipLaunchKernelGGL(sweep3d_lin_kernel<false>, dim3(nblocks), dim3(L3T), 0, s, A);
}

}  // namespace cdiff
