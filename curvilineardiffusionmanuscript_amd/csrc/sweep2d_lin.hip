// sweep2d_lin.hip -- the FP32 correction sweep of the defect-correction step,
// in coefficient form, several pseudo-iterations per launch.
//
// Inside one outer iteration the correction equation is linear: alpha, theta
// and dtau are frozen at the FP64 iterate, and the Newton term of the face
// conductivity is linear in the correction e.  So everything that does not
// depend on e is computed once per outer iteration, in FP64, by k_dc_outer2d
// (dc_outer2d.hip; k_lin_coefs in dc_kernels.hip does the same on its own),
// and stored per face and per cell in FP32:
//
//   face f between cells 0 and 1:  q'_f = P_f e_0 - M_f e_1
//                                   q_f <- q'_f + W_f (q_f - q'_f)
//       P = abar - k0 dubar,  M = abar + k1 dubar  (k: derivative of the face
//       mean times d(alpha)/du; dubar: the FP64 iterate's difference across
//       the face), W = thetabar / (1 + thetabar); faces outside
//       expand_lower(domain, axis) have P = M = W = 0; the ghost of a
//       boundary face is folded into the domain cell's coefficient
//       (e_ghost = s e_domain, s = +1 copy, -1 negated copy, 0 fixed value)
//   cell:  D = m0 x_c + m1 x_w + m2 y_c + m3 y_s
//            + m4 ((x_n + x_nw) - (x_s + x_sw) + (y_e + y_es) - (y_w + y_ws))
//          e <- e + C (r - D - e/dt),   C = dtau / (1 + dtau/dt)
//
// which is the APT update of the frozen-coefficient correction equation
// e/dt + div q'(e) = r, rearranged (the fixed points of q and e do not depend
// on W and C).  A sweep then costs ~40 FP32 operations per cell and the
// kernel is bound by memory, so each launch runs up to K iterations on a
// tile with a K-cell ring held in LDS (temporal blocking): the coefficients
// are read once per launch, and e, q once in and once out.  Two thread
// layouts: 64-column tiles with the region cells spread over the threads
// (sweep2d_lin_kernel), and wave-aligned region rows with tiles of 64 - 2K
// columns (sweep2d_lin_rows_kernel, the default).
#include "cdiff_dev.h"
#include "lin2d_ops.h"
#include "sweep_args.h"

namespace cdiff {

namespace {

constexpr int LTX = 64;

}  // namespace

template <int K, int TY, int LNTHR, bool CHECK>
__global__ __launch_bounds__(LNTHR) void sweep2d_lin_kernel(Lin2DArgs A) {
  constexpr int RX = LTX + 2 * K, RY = TY + 2 * K, NR = RX * RY;
  constexpr int NS = (NR + LNTHR - 1) / LNTHR;
  if (A.ctrl->done) return;
  __shared__ float sE[NR], sX[NR], sY[NR];
  __shared__ float sPX[CHECK ? NR : 1], sPY[CHECK ? NR : 1];
  __shared__ double sRed[2][LNTHR / WAVE];

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
  const int x0 = g.lo[0] + bx * LTX, y0 = g.lo[1] + by * TY;
  const int gx = x0 - K, gy = y0 - K;  // global index of region cell (0, 0)
  const float inv_dt = A.inv_dt;

  // ---- loads: every region cell, once per launch
  float e[NS], qx[NS], qy[NS], rb[NS], cc[NS], m0[NS], m1[NS], m2[NS], m3[NS], m4[NS];
  float px[NS], mx[NS], wx[NS], py[NS], my[NS], wy[NS];
  int code[NS];  // bit 0: domain cell; bits 8..15: lx; bits 16..23: ly (0xff: no cell)
#pragma unroll
  for (int p = 0; p < NS; ++p) {
    const int t = tid + p * LNTHR;
    e[p] = qx[p] = qy[p] = rb[p] = cc[p] = m0[p] = m1[p] = m2[p] = m3[p] = m4[p] = 0.f;
    px[p] = mx[p] = wx[p] = py[p] = my[p] = wy[p] = 0.f;
    code[p] = 0xffff00;
    if (t < NR) {
      const int lx = t % RX, ly = t / RX;
      const int i = gx + lx, r = gy + ly;
      const bool dom = i >= g.lo[0] && i <= g.hi[0] && r >= g.lo[1] && r <= g.hi[1];
      code[p] = (int)dom | (lx << 8) | (ly << 16);
      const bool xon = i >= g.lo[0] - 1 && i <= g.hi[0] && r >= g.lo[1] && r <= g.hi[1];
      const bool yon = i >= g.lo[0] && i <= g.hi[0] && r >= g.lo[1] - 1 && r <= g.hi[1];
      const int64_t c = at(g, i, r, 0);
      if (dom) {
        if (!A.zero_src) e[p] = A.e_src[c];
        rb[p] = A.base[c];
        cc[p] = A.cc[c];
        m0[p] = A.m[0][c];
        m1[p] = A.m[1][c];
        m2[p] = A.m[2][c];
        m3[p] = A.m[3][c];
        m4[p] = A.m[4][c];
      }
      if (xon) {
        if (!A.zero_src) qx[p] = A.qx_src[c];
        px[p] = A.fx[0][c];
        mx[p] = A.fx[1][c];
        wx[p] = A.fx[2][c];
      }
      if (yon) {
        if (!A.zero_src) qy[p] = A.qy_src[c];
        py[p] = A.fy[0][c];
        my[p] = A.fy[1][c];
        wy[p] = A.fy[2][c];
      }
    }
  }
#pragma unroll
  for (int p = 0; p < NS; ++p) {
    const int t = tid + p * LNTHR;
    if (t < NR) {
      sE[t] = e[p];
      sX[t] = qx[p];
      sY[t] = qy[p];
    }
  }
  __syncthreads();

  // ---- iterations s = K - nit .. K-1; iteration s updates the cells
  //      [s+1, RX-1-s) x [s+1, RY-1-s) (the tile at s = K-1)
  double sq = 0.0, smax = 0.0;
  for (int s = K - A.nit; s < K; ++s) {
    const bool last = s == K - 1;
    // faces (xi: between (lx, ly) and (lx+1, ly); eta: (lx, ly) and (lx, ly+1))
#pragma unroll
    for (int p = 0; p < NS; ++p) {
      const int t = tid + p * LNTHR;
      const int lx = (code[p] >> 8) & 0xff, ly = (code[p] >> 16) & 0xff;
      if (lx >= s && ly >= s && ly <= RY - 1 - s) {
        if (lx <= RX - 2 - s) {
          float qp;
          qx[p] = lin_face(px[p], mx[p], wx[p], sE[t], sE[t + 1], qx[p], qp);
          sX[t] = qx[p];
          if (CHECK && last) sPX[t] = qp;
        }
        if (lx <= RX - 1 - s && ly <= RY - 2 - s) {
          float qp;
          qy[p] = lin_face(py[p], my[p], wy[p], sE[t], sE[t + RX], qy[p], qp);
          sY[t] = qy[p];
          if (CHECK && last) sPY[t] = qp;
        }
      }
    }
    __syncthreads();
    // cells
#pragma unroll
    for (int p = 0; p < NS; ++p) {
      const int t = tid + p * LNTHR;
      const int lx = (code[p] >> 8) & 0xff, ly = (code[p] >> 16) & 0xff;
      if ((code[p] & 1) && lx >= s + 1 && lx < RX - 1 - s && ly >= s + 1 && ly < RY - 1 - s) {
        const float cross = lin_cross(sX[t + RX], sX[t + RX - 1], sX[t - RX], sX[t - RX - 1], sY[t + 1],
                                      sY[t - RX + 1], sY[t - 1], sY[t - RX - 1]);
        const float D = lin_div(m0[p], m1[p], m2[p], m3[p], m4[p], sX[t], sX[t - 1], sY[t], sY[t - RX], cross);
        const float en = lin_update(e[p], cc[p], rb[p], D, inv_dt);
        if (CHECK && last) {
          const float crossp = lin_cross(sPX[t + RX], sPX[t + RX - 1], sPX[t - RX], sPX[t - RX - 1], sPY[t + 1],
                                         sPY[t - RX + 1], sPY[t - 1], sPY[t - RX - 1]);
          const float Dp =
              lin_div(m0[p], m1[p], m2[p], m3[p], m4[p], sPX[t], sPX[t - 1], sPY[t], sPY[t - RX], crossp);
          const double res = (double)lin_residual(rb[p], en, inv_dt, Dp);
          sq += res * res;
          smax = nanmax(smax, fabs(res));
          if (res != res) smax = res;
        }
        e[p] = en;
        sE[t] = en;
      }
    }
    __syncthreads();
  }

  // ---- writes: the tile's cells and faces, and the boundary faces below
  //      the domain's first column / row that no cell owns
#pragma unroll
  for (int p = 0; p < NS; ++p) {
    const int lx = (code[p] >> 8) & 0xff, ly = (code[p] >> 16) & 0xff;
    const bool rows = ly >= K && ly < K + TY, cols = lx >= K && lx < K + LTX;
    const int i = gx + lx, r = gy + ly;
    const int64_t c = at(g, i, r, 0);
    if ((code[p] & 1) && rows && cols) {
      A.e_dst[c] = e[p];
      A.qx_dst[c] = qx[p];
      A.qy_dst[c] = qy[p];
    }
    if (i == g.lo[0] - 1 && rows && r <= g.hi[1]) A.qx_dst[c] = qx[p];
    if (r == g.lo[1] - 1 && cols && i <= g.hi[0]) A.qy_dst[c] = qy[p];
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
      for (int w = 0; w < LNTHR / WAVE; ++w) {
        s += sRed[0][w];
        m = nanmax(m, sRed[1][w]);
      }
      A.part_sum[b] = s;
      A.part_max[b] = m;
    }
  }
}

// The same sweep with region rows aligned to wavefronts: the region is 64
// columns wide (a tile of 64 - 2K columns plus the K-cell ring on each
// side), so lane l of every wave holds region column l and wave w holds the
// region rows w, w + NW, w + 2 NW, ...  Row bounds are then wave-uniform
// (scalar branches) and no per-slot index decoding is needed.  A block of NW
// waves owns a tile of (64 - 2K) x (NW NSR - 2K) cells.  The arithmetic of
// every face and cell is the one of sweep2d_lin_kernel, in the same order.
template <int K, int NW, int NSR, bool CHECK>
__global__ __launch_bounds__(NW * WAVE) void sweep2d_lin_rows_kernel(Lin2DArgs A) {
  constexpr int RX = WAVE, RY = NW * NSR, NR = RX * RY;
  constexpr int TX = RX - 2 * K, TY = RY - 2 * K;
  if (A.ctrl->done) return;
  __shared__ float sE[NR], sX[NR], sY[NR];
  __shared__ float sPX[CHECK ? NR : 1], sPY[CHECK ? NR : 1];
  __shared__ double sRed[2][NW];

  const Geo& g = A.g;
  const int tid = threadIdx.x;
  const int lx = tid & (WAVE - 1);
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
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
  const int x0 = g.lo[0] + bx * TX, y0 = g.lo[1] + by * TY;
  const int i = x0 - K + lx;  // global column of this lane
  const int gy = y0 - K;      // global row of region row 0
  const float inv_dt = A.inv_dt;
  const bool icol = i >= g.lo[0] && i <= g.hi[0];
  const bool xcol = i >= g.lo[0] - 1 && i <= g.hi[0];

  float e[NSR], qx[NSR], qy[NSR], rb[NSR], cc[NSR], m0[NSR], m1[NSR], m2[NSR], m3[NSR], m4[NSR];
  float px[NSR], mx[NSR], wx[NSR], py[NSR], my[NSR], wy[NSR];
  unsigned dom = 0;  // bit p: region cell of slot p is a domain cell
#pragma unroll
  for (int p = 0; p < NSR; ++p) {
    const int ly = w + NW * p, r = gy + ly;
    e[p] = qx[p] = qy[p] = rb[p] = cc[p] = m0[p] = m1[p] = m2[p] = m3[p] = m4[p] = 0.f;
    px[p] = mx[p] = wx[p] = py[p] = my[p] = wy[p] = 0.f;
    const bool rin = r >= g.lo[1] && r <= g.hi[1];
    const bool d = icol && rin;
    const bool xon = xcol && rin;
    const bool yon = icol && r >= g.lo[1] - 1 && r <= g.hi[1];
    const int64_t c = at(g, i, r, 0);
    if (d) {
      dom |= 1u << p;
      if (!A.zero_src) e[p] = A.e_src[c];
      rb[p] = A.base[c];
      cc[p] = A.cc[c];
      m0[p] = A.m[0][c];
      m1[p] = A.m[1][c];
      m2[p] = A.m[2][c];
      m3[p] = A.m[3][c];
      m4[p] = A.m[4][c];
    }
    if (xon) {
      if (!A.zero_src) qx[p] = A.qx_src[c];
      px[p] = A.fx[0][c];
      mx[p] = A.fx[1][c];
      wx[p] = A.fx[2][c];
    }
    if (yon) {
      if (!A.zero_src) qy[p] = A.qy_src[c];
      py[p] = A.fy[0][c];
      my[p] = A.fy[1][c];
      wy[p] = A.fy[2][c];
    }
  }
#pragma unroll
  for (int p = 0; p < NSR; ++p) {
    const int t = (w + NW * p) * RX + lx;
    sE[t] = e[p];
    sX[t] = qx[p];
    sY[t] = qy[p];
  }
  __syncthreads();

  double sq = 0.0, smax = 0.0;
  for (int s = K - A.nit; s < K; ++s) {
    const bool last = s == K - 1;
    const bool fx_on = lx >= s && lx <= RX - 2 - s;
    const bool fy_on = lx >= s && lx <= RX - 1 - s;
    const bool c_on = lx >= s + 1 && lx < RX - 1 - s;
#pragma unroll
    for (int p = 0; p < NSR; ++p) {
      const int ly = w + NW * p;
      const int t = ly * RX + lx;
      if (ly >= s && ly <= RY - 1 - s) {
        if (fx_on) {
          float qp;
          qx[p] = lin_face(px[p], mx[p], wx[p], sE[t], sE[t + 1], qx[p], qp);
          sX[t] = qx[p];
          if (CHECK && last) sPX[t] = qp;
        }
        if (fy_on && ly <= RY - 2 - s) {
          float qp;
          qy[p] = lin_face(py[p], my[p], wy[p], sE[t], sE[t + RX], qy[p], qp);
          sY[t] = qy[p];
          if (CHECK && last) sPY[t] = qp;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NSR; ++p) {
      const int ly = w + NW * p;
      const int t = ly * RX + lx;
      if (ly >= s + 1 && ly < RY - 1 - s && c_on && ((dom >> p) & 1u)) {
        const float cross = lin_cross(sX[t + RX], sX[t + RX - 1], sX[t - RX], sX[t - RX - 1], sY[t + 1],
                                      sY[t - RX + 1], sY[t - 1], sY[t - RX - 1]);
        const float D = lin_div(m0[p], m1[p], m2[p], m3[p], m4[p], sX[t], sX[t - 1], sY[t], sY[t - RX], cross);
        const float en = lin_update(e[p], cc[p], rb[p], D, inv_dt);
        if (CHECK && last) {
          const float crossp = lin_cross(sPX[t + RX], sPX[t + RX - 1], sPX[t - RX], sPX[t - RX - 1], sPY[t + 1],
                                         sPY[t - RX + 1], sPY[t - 1], sPY[t - RX - 1]);
          const float Dp =
              lin_div(m0[p], m1[p], m2[p], m3[p], m4[p], sPX[t], sPX[t - 1], sPY[t], sPY[t - RX], crossp);
          const double res = (double)lin_residual(rb[p], en, inv_dt, Dp);
          sq += res * res;
          smax = nanmax(smax, fabs(res));
          if (res != res) smax = res;
        }
        e[p] = en;
        sE[t] = en;
      }
    }
    __syncthreads();
  }

  const bool cols = lx >= K && lx < K + TX;
#pragma unroll
  for (int p = 0; p < NSR; ++p) {
    const int ly = w + NW * p;
    const int r = gy + ly;
    const bool rows = ly >= K && ly < K + TY;
    const int64_t c = at(g, i, r, 0);
    if (((dom >> p) & 1u) && rows && cols) {
      A.e_dst[c] = e[p];
      A.qx_dst[c] = qx[p];
      A.qy_dst[c] = qy[p];
    }
    if (i == g.lo[0] - 1 && rows && r <= g.hi[1]) A.qx_dst[c] = qx[p];
    if (r == g.lo[1] - 1 && cols && i <= g.hi[0]) A.qy_dst[c] = qy[p];
  }
  if (CHECK) {
    sq = wave_sum(sq);
    smax = wave_max(smax);
    if (lx == 0) {
      sRed[0][w] = sq;
      sRed[1][w] = smax;
    }
    __syncthreads();
    if (tid == 0) {
      double s = 0.0, m = 0.0;
      for (int v = 0; v < NW; ++v) {
        s += sRed[0][v];
        m = nanmax(m, sRed[1][v]);
      }
      A.part_sum[b] = s;
      A.part_max[b] = m;
    }
  }
}

template <int K, int TY, int NT>
static void launch_k(const Lin2DArgs& A, bool check, int nblocks, hipStream_t s) {
  if (check)
    hipLaunchKernelGGL((sweep2d_lin_kernel<K, TY, NT, true>), dim3(nblocks), dim3(NT), 0, s, A);
  else
    hipLaunchKernelGGL((sweep2d_lin_kernel<K, TY, NT, false>), dim3(nblocks), dim3(NT), 0, s, A);
}

template <int K, int NW, int NSR>
static void launch_rows(const Lin2DArgs& A, bool check, int nblocks, hipStream_t s) {
  if (check)
    hipLaunchKernelGGL((sweep2d_lin_rows_kernel<K, NW, NSR, true>), dim3(nblocks), dim3(NW * WAVE), 0, s, A);
  else
    hipLaunchKernelGGL((sweep2d_lin_rows_kernel<K, NW, NSR, false>), dim3(nblocks), dim3(NW * WAVE), 0, s, A);
}

// the wave-aligned variant: 8 waves of 4 region rows each (measured at
// 4096^2: 16 waves of 4 rows, or 8 of 6, run fewer blocks per CU and are
// 15-35 % slower)
constexpr int ROWS_NW = 8, ROWS_NSR = 4;

void lin2d_tile(int kmax, int layout, int* cols, int* rows) {
  if (layout == 0) {
    *cols = LTX;
    *rows = kmax >= 4 ? 16 : 8;
    return;
  }
  *cols = WAVE - 2 * kmax;
  *rows = ROWS_NW * ROWS_NSR - 2 * kmax;
}

void launch_sweep2d_lin(const Lin2DArgs& A, int kmax, int layout, bool check, int nblocks, hipStream_t s) {
  if (layout == 0) {
    switch (kmax) {
      case 4: launch_k<4, 16, 512>(A, check, nblocks, s); break;
      case 6: launch_k<6, 16, 1024>(A, check, nblocks, s); break;
      case 8: launch_k<8, 16, 1024>(A, check, nblocks, s); break;
      default: launch_k<2, 8, 512>(A, check, nblocks, s); break;
    }
    return;
  }
  switch (kmax) {
    case 4: launch_rows<4, ROWS_NW, ROWS_NSR>(A, check, nblocks, s); break;
    case 6: launch_rows<6, ROWS_NW, ROWS_NSR>(A, check, nblocks, s); break;
    case 8: launch_rows<8, ROWS_NW, ROWS_NSR>(A, check, nblocks, s); break;
    default: launch_rows<2, ROWS_NW, ROWS_NSR>(A, check, nblocks, s); break;
  }
}

}  // namespace cdiff
