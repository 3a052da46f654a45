// sweep3d_lin2.hip -- two pseudo-iterations of the 3D FP32 correction sweep
// per launch (temporal blocking of sweep3d_lin.hip, whose header has the
// coefficient form and the cross-term groups).
//
// A block owns a 64 x 8 column of cells and marches up a chunk of k planes,
// one thread per cell of the column's two-cell ring (68 x 12 cells, 816
// threads).  Per plane step t the block
//   1. stores e0(t) (the launch's input correction) in LDS,
//   2. computes the iteration-1 faces of plane t on the whole ring (xi, eta)
//      and the zeta faces between t-1 and t,
//   3. updates plane t-1 to e1 on the one-cell ring (66 x 10),
//   4. computes the iteration-2 faces of plane t-1 on that ring from e1, with
//      the iteration-1 faces as the old fluxes, and the zeta faces between
//      t-2 and t-1,
//   5. updates plane t-2 of the column to e2 and writes e2 and its fluxes.
// The coefficients are read once for both iterations and e, q once in and
// once out; every value is computed by the same FP32 expressions, in the same
// order, as two launches of the one-iteration sweep, so the results are
// bit-identical to those.  Loads of the next plane (two planes without the
// convergence check) and of the update coefficients of plane t are in flight
// while steps 3-5 run; the face coefficients of a plane wait for their second
// use one step later in the thread's own LDS slots.
#include "cdiff_dev.h"
#include "sweep_args.h"

namespace cdiff {

namespace {

constexpr int T2X = 64, T2Y = 8;                        // owned column
constexpr int Q2X = T2X + 4, Q2Y = T2Y + 4, Q2 = Q2X * Q2Y;  // two-cell ring: 68 x 12
constexpr int T2N = (Q2 + WAVE - 1) / WAVE * WAVE;      // 832 threads

}  // namespace

template <bool CHECK>
__global__ __launch_bounds__(T2N) void sweep3d_lin2_kernel(Lin3DArgs A) {
  if (A.ctrl->done) return;
  __shared__ float sE0[2][Q2], sX1[3][Q2], sY1[3][Q2], sZ1[2][Q2];
  __shared__ float sE1[2][Q2], sX2[3][Q2], sY2[3][Q2], sZ2[2][Q2];
  __shared__ float sPX[CHECK ? 3 : 1][CHECK ? Q2 : 1], sPY[CHECK ? 3 : 1][CHECK ? Q2 : 1];
  __shared__ float sPZ[CHECK ? 2 : 1][CHECK ? Q2 : 1];
  __shared__ float sC[2][9][Q2];  // per-thread store of the face coefficients between plane steps
  __shared__ double sRed[2][T2N / WAVE];
  const Geo& g = A.g;
  const int tid = threadIdx.x;
  const int nb = gridDim.x, per = nb >> 3, b = blockIdx.x;
  const int tile = (b & 7) * per + (b >> 3);  // XCD-aware: consecutive tiles on one XCD
  if (tile >= A.ntiles) {
    if (CHECK && tid == 0) {
      A.part_sum[b] = 0.0;
      A.part_max[b] = 0.0;
    }
    return;
  }
  const int kc = tile % A.chunks, rest = tile / A.chunks;
  const int si = rest % A.stripsI, sj = rest / A.stripsI;
  const int x0 = g.lo[0] + si * T2X, y0 = g.lo[1] + sj * T2Y;
  const int z0 = g.lo[2] + kc * A.ch, z1 = min(z0 + A.ch - 1, g.hi[2]);
  const float inv_dt = A.inv_dt;

  // this thread's cell of the ring and its roles
  const bool valid = tid < Q2;
  const int lx = tid % Q2X, ly = tid / Q2X;
  const int i = x0 - 2 + lx, j = y0 - 2 + ly;
  const int r = tid;
  const bool inner = valid && lx >= 1 && lx <= Q2X - 2 && ly >= 1 && ly <= Q2Y - 2;
  const bool owned = lx >= 2 && lx <= Q2X - 3 && ly >= 2 && ly <= Q2Y - 3 && i <= g.hi[0] && j <= g.hi[1];
  const bool xf1 = valid && lx < Q2X - 1, yf1 = valid && ly < Q2Y - 1;  // iteration-1 faces
  const bool xf2 = inner && lx < Q2X - 2, yf2 = inner && ly < Q2Y - 2;  // iteration-2 faces
  const bool ij_dom = i >= g.lo[0] && i <= g.hi[0] && j >= g.lo[1] && j <= g.hi[1];
  auto in_dom = [&](int k) { return ij_dom && k >= g.lo[2] && k <= g.hi[2]; };
  auto face_on = [&](int ax, int k) {
    const int lo0 = g.lo[0] - (ax == 0), lo1 = g.lo[1] - (ax == 1), lo2 = g.lo[2] - (ax == 2);
    return i >= lo0 && i <= g.hi[0] && j >= lo1 && j <= g.hi[1] && k >= lo2 && k <= g.hi[2];
  };

  // e0 of plane k, the xi/eta faces of plane k and the zeta face below it
  // (between k-1 and k, stored at k-1): P, M, W and the old flux
  struct Fetch {
    float e, f[3][4];
  };
  auto fetch = [&](int k, Fetch& F, bool zf) {
    F.e = 0.f;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) F.f[ax][0] = F.f[ax][1] = F.f[ax][2] = F.f[ax][3] = 0.f;
    if (!valid) return;
    const int64_t c = at(g, i, j, k);
    if (!A.zero_src && in_dom(k)) F.e = A.e_src[c];
    if (xf1 && face_on(0, k)) {
      F.f[0][0] = A.f[0 * A.stride + c];
      F.f[0][1] = A.f[1 * A.stride + c];
      F.f[0][2] = A.f[2 * A.stride + c];
      F.f[0][3] = A.zero_src ? 0.f : A.q_src[0][c];
    }
    if (yf1 && face_on(1, k)) {
      F.f[1][0] = A.f[3 * A.stride + c];
      F.f[1][1] = A.f[4 * A.stride + c];
      F.f[1][2] = A.f[5 * A.stride + c];
      F.f[1][3] = A.zero_src ? 0.f : A.q_src[1][c];
    }
    if (zf && face_on(2, k - 1)) {
      const int64_t cz = c - g.plane;
      F.f[2][0] = A.f[6 * A.stride + cz];
      F.f[2][1] = A.f[7 * A.stride + cz];
      F.f[2][2] = A.f[8 * A.stride + cz];
      F.f[2][3] = A.zero_src ? 0.f : A.q_src[2][cz];
    }
  };
  struct Upd {
    float rb, cc, m[9];
  };
  auto fetch_upd = [&](int k, Upd& U) {
    if (!inner || !in_dom(k)) return;
    const int64_t c = at(g, i, j, k);
    U.rb = A.base[c];
    U.cc = A.cc[c];
#pragma unroll
    for (int q = 0; q < 9; ++q) U.m[q] = A.m[q * A.stride + c];
  };
  // divergence of the faces around cell r (planes m, c, p of the xi/eta
  // faces, zeta faces below (Zm) and above (Zc))
  auto div = [&](const Upd& U, const float* Xm, const float* Xc, const float* Xp, const float* Ym, const float* Yc,
                 const float* Yp, const float* Zm, const float* Zc) {
    const float xc = Xc[r], xw = Xc[r - 1], yc = Yc[r], ys = Yc[r - Q2X], zcv = Zc[r], zs = Zm[r];
    const float x_eta = (Xc[r + Q2X] + Xc[r + Q2X - 1]) - (Xc[r - Q2X] + Xc[r - Q2X - 1]);
    const float x_zeta = (Xp[r] + Xp[r - 1]) - (Xm[r] + Xm[r - 1]);
    const float y_xi = (Yc[r + 1] + Yc[r - Q2X + 1]) - (Yc[r - 1] + Yc[r - Q2X - 1]);
    const float y_zeta = (Yp[r] + Yp[r - Q2X]) - (Ym[r] + Ym[r - Q2X]);
    const float z_eta = (Zc[r + Q2X] + Zm[r + Q2X]) - (Zc[r - Q2X] + Zm[r - Q2X]);
    const float z_xi = (Zc[r + 1] + Zm[r + 1]) - (Zc[r - 1] + Zm[r - 1]);
    return U.m[0] * xc + U.m[1] * xw + U.m[2] * yc + U.m[3] * ys + U.m[4] * zcv + U.m[5] * zs +
           U.m[6] * (x_eta + y_xi) + U.m[7] * (x_zeta + z_xi) + U.m[8] * (y_zeta + z_eta);
  };

  // loads of the next planes: step t consumes the register set of plane t and
  // refills it with plane t+PD.  Without the check, two sets (PD = 2,
  // alternating with the step) put two planes' loads in flight behind the
  // update; the check variant's residual needs the registers of the second
  // set (PD = 1)
  constexpr int PD = CHECK ? 1 : 2;
  Fetch FS[PD];
  Upd U1{}, U2{}, Un{};
  const int t0 = z0 - 2, t1 = z1 + 2;
  fetch(t0, FS[0], false);
  if (PD == 2) fetch(t0 + 1, FS[PD - 1], true);
  double sq = 0.0, smax = 0.0;
  // The LDS rings hold plane q in slot (q - t0) mod 2 (e, the zeta faces
  // between q and q + 1, the face coefficients) or mod 3 (the xi and eta
  // faces).  The plane loop is unrolled over the period 6, so with
  // P = (t - t0) mod 6 every slot is a compile-time constant.
  auto step = [&](auto Pc, int t) {
    constexpr int P = decltype(Pc)::value;
    // slots of plane t - d
    constexpr auto s2 = [](int d) { return (P - d + 6) % 2; };
    constexpr auto s3 = [](int d) { return (P - d + 6) % 3; };
    Fetch& F = FS[P % PD];
    // ---- 1-2: e0(t) into LDS; iteration-1 faces of plane t (+ zeta faces t-1|t)
    if (valid) sE0[s2(0)][r] = F.e;
    __syncthreads();
    const bool zf1 = t >= z0 - 1;
    if (valid) {
      const float* E = sE0[s2(0)];
      const float* Em = sE0[s2(1)];
      const float ec = E[r];
      if (xf1) {
        const float px = F.f[0][0] * ec - F.f[0][1] * E[r + 1];
        sX1[s3(0)][r] = px + F.f[0][2] * (F.f[0][3] - px);
      }
      if (yf1) {
        const float py = F.f[1][0] * ec - F.f[1][1] * E[r + Q2X];
        sY1[s3(0)][r] = py + F.f[1][2] * (F.f[1][3] - py);
      }
      if (zf1) {
        const float pz = F.f[2][0] * Em[r] - F.f[2][1] * ec;
        sZ1[s2(1)][r] = pz + F.f[2][2] * (F.f[2][3] - pz);
      }
    }
    // the face coefficients of plane t, for its iteration-2 faces one step
    // later, in this thread's own LDS slots (no barrier needed)
    if (inner) {
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int q = 0; q < 3; ++q) sC[s2(0)][3 * ax + q][r] = F.f[ax][q];
    }
    // loads of the next plane step and of the update coefficients of plane t
    if (t + PD <= t1) fetch(t + PD, F, t + PD >= z0 - 1);  // F is free now
    if (t >= z0 - 1 && t <= z1 + 1) fetch_upd(t, Un);
    __syncthreads();

    // ---- 3: iteration-1 update of plane t-1 on the one-cell ring
    const int p1 = t - 1;
    if (p1 >= z0 - 1) {
      if (inner) {
        float e1 = 0.f;
        if (in_dom(p1)) {
          const float D = div(U1, sX1[s3(2)], sX1[s3(1)], sX1[s3(0)], sY1[s3(2)], sY1[s3(1)], sY1[s3(0)],
                              sZ1[s2(2)], sZ1[s2(1)]);
          const float e = sE0[s2(1)][r];
          e1 = e + U1.cc * (U1.rb - D - e * inv_dt);
        }
        sE1[s2(1)][r] = e1;
      }
      __syncthreads();

      // ---- 4: iteration-2 faces of plane t-1 (+ zeta faces t-2|t-1)
      if (inner) {
        const float(*Cf)[Q2] = sC[s2(1)];
        const float* E = sE1[s2(1)];
        const float* Em = sE1[s2(2)];
        const float ec = E[r];
        if (xf2) {
          const float px = Cf[0][r] * ec - Cf[1][r] * E[r + 1];
          sX2[s3(1)][r] = px + Cf[2][r] * (sX1[s3(1)][r] - px);
          if (CHECK) sPX[s3(1)][r] = px;
        }
        if (yf2) {
          const float py = Cf[3][r] * ec - Cf[4][r] * E[r + Q2X];
          sY2[s3(1)][r] = py + Cf[5][r] * (sY1[s3(1)][r] - py);
          if (CHECK) sPY[s3(1)][r] = py;
        }
        if (p1 >= z0) {
          // face p1-1 | p1
          const float pz = Cf[6][r] * Em[r] - Cf[7][r] * ec;
          sZ2[s2(2)][r] = pz + Cf[8][r] * (sZ1[s2(2)][r] - pz);
          if (CHECK) sPZ[s2(2)][r] = pz;
        }
      }
      __syncthreads();

      // ---- 5: iteration-2 update of plane t-2 of the column
      const int p2 = t - 2;
      if (p2 >= z0 && owned) {
        const int64_t c = at(g, i, j, p2);
        const float D = div(U2, sX2[s3(3)], sX2[s3(2)], sX2[s3(1)], sY2[s3(3)], sY2[s3(2)], sY2[s3(1)],
                            sZ2[s2(3)], sZ2[s2(2)]);
        const float e = sE1[s2(2)][r];
        const float en = e + U2.cc * (U2.rb - D - e * inv_dt);
        A.e_dst[c] = en;
        A.q_dst[0][c] = sX2[s3(2)][r];
        A.q_dst[1][c] = sY2[s3(2)][r];
        A.q_dst[2][c] = sZ2[s2(2)][r];
        if (CHECK) {
          const float Dp = div(U2, sPX[s3(3)], sPX[s3(2)], sPX[s3(1)], sPY[s3(3)], sPY[s3(2)], sPY[s3(1)],
                               sPZ[s2(3)], sPZ[s2(2)]);
          const double res = (double)(U2.rb - en * inv_dt - Dp);
          sq += res * res;
          smax = nanmax(smax, fabs(res));
        }
        // boundary faces below the domain that no cell owns
        if (i == g.lo[0]) A.q_dst[0][c - 1] = sX2[s3(2)][r - 1];
        if (j == g.lo[1]) A.q_dst[1][c - g.pitch] = sY2[s3(2)][r - Q2X];
        if (p2 == g.lo[2]) A.q_dst[2][c - g.plane] = sZ2[s2(3)][r];
      }
    }
    U2 = U1;
    U1 = Un;
  };
  int t = t0;
  while (march_phases<0, 6>(step, t, t1)) {
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
      for (int w = 0; w < T2N / WAVE; ++w) {
        s += sRed[0][w];
        m = nanmax(m, sRed[1][w]);
      }
      A.part_sum[b] = s;
      A.part_max[b] = m;
    }
  }
}

void launch_sweep3d_lin2(const Lin3DArgs& A, bool check, int nblocks, hipStream_t s) {
  if (check)
    hipLaunchKernelGGL(sweep3d_lin2_kernel<true>, dim3(nblocks), dim3(T2N), 0, s, A);
  else
    hipLaunchKernelGGL(sweep3d_lin2_kernel<false>, dim3(nblocks), dim3(T2N), 0, s, A);
}

}  // namespace cdiff
