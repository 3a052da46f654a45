// sweep2d_march.hip -- the FP32 correction sweep of sweep2d_lin.hip (same
// coefficient form, same arithmetic per face and per cell), K iterations per
// launch as a pipeline that marches up the rows.
//
// One wavefront owns a strip of 64 - 2K columns (lane l holds column
// x0 - K + l: the strip plus a K-cell ring on each side) and a chunk of rows
// [ya, yb).  It loads rows ya - K .. yb - 1 + K once each, one per step; at
// step t, stage n = 1..K runs iteration n on the row n below the newest:
//   x faces of row t - n + 1   q'_x = P e - M e_east,  q_x <- q' + W (q_x - q')
//   y faces of row t - n       q'_y = P e - M e_north, q_y <- q' + W (q_y - q')
//   cells   of row t - n       e <- e + C (r - D - e/dt)
// from the iteration n - 1 values of the rows around it (kept in registers,
// east and west neighbours by DPP wavefront shifts).  Stage n needs iteration n - 1
// of row t - n + 1, which stage n - 1 produced in the same step, so the
// stages run in order.  Each row's coefficients stay in registers for the K
// steps that use them.  Against the tiled sweep, the ring is re-read only
// along x (64 / (64 - 2K)) and 2K rows per chunk, and there are no barriers;
// the neighbouring strips of a chunk run on one XCD and share the cache lines
// at their edges in its L2.
//
// Values outside the dependency cone of the strip's cells (the ring's edge
// lanes and the first rows of a chunk) are computed from incomplete
// neighbours and never reach a written cell: after stage n, lanes
// n .. 63 - n and rows from ya - K + n up are exact.
#include "cdiff_dev.h"
#include "sweep_args.h"

namespace cdiff {

namespace {

constexpr int MWAVES = 4;  // independent wavefronts per block

// what one lane loads per row
struct RowIn {
  float e, qx, qy;
  float px, mx, wx;                       // x face (lane's column, this row)
  float py, my, wy;                       // y face (between this row and the next)
  float rb, cc, m0, m1, m2, m3, m4;       // cell
};

__device__ __forceinline__ void load_row(const Lin2DArgs& A, int i, int r, bool icol, bool xcol, RowIn& v) {
  const Geo& g = A.g;
  v = RowIn{};
  const bool rin = r >= g.lo[1] && r <= g.hi[1];
  const bool d = icol && rin;
  const bool xon = xcol && rin;
  const bool yon = icol && r >= g.lo[1] - 1 && r <= g.hi[1];
  const int64_t c = at(g, i, r, 0);
  if (d) {
    if (!A.zero_src) v.e = A.e_src[c];
    v.rb = A.base[c];
    v.cc = A.cc[c];
    v.m0 = A.m[0][c];
    v.m1 = A.m[1][c];
    v.m2 = A.m[2][c];
    v.m3 = A.m[3][c];
    v.m4 = A.m[4][c];
  }
  if (xon) {
    if (!A.zero_src) v.qx = A.qx_src[c];
    v.px = A.fx[0][c];
    v.mx = A.fx[1][c];
    v.wx = A.fx[2][c];
  }
  if (yon) {
    if (!A.zero_src) v.qy = A.qy_src[c];
    v.py = A.fy[0][c];
    v.my = A.fy[1][c];
    v.wy = A.fy[2][c];
  }
}

// east / west neighbour by a DPP wavefront shift by one lane (the edge lane
// reads 0; edge lanes lie outside every written cell's dependency cone)
__device__ __forceinline__ float dpp_from_left(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));  // wave_shr:1
}
__device__ __forceinline__ float dpp_from_right(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));  // wave_shl:1
}

}  // namespace

template <int K, bool CHECK>
__global__ __launch_bounds__(MWAVES * WAVE) void sweep2d_lin_march_kernel(Lin2DArgs A) {
  constexpr int TX = WAVE - 2 * K;
  if (A.ctrl->done) return;
  const Geo& g = A.g;
  const int lane = threadIdx.x & (WAVE - 1);
  // XCD-aware split: block b runs on XCD b mod 8; the blocks of one XCD take
  // the chunks c = XCD, XCD + 8, ... and, per chunk, groups of MWAVES
  // neighbouring strips in order, so strips that share the cache lines at
  // their edges share that XCD's L2 (they march up the same rows together)
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = blockIdx.x, k = b >> 3;
  const int groups = (A.strips + MWAVES - 1) / MWAVES;
  const int chunk = (b & 7) + 8 * (k / groups);
  const int strip = (k % groups) * MWAVES + w;
  const int part = b * MWAVES + w;
  if (chunk >= A.chunks || strip >= A.strips) {
    if (CHECK && lane == 0) {
      A.part_sum[part] = 0.0;
      A.part_max[part] = 0.0;
    }
    return;
  }
  const int x0 = g.lo[0] + strip * TX;
  const int i = x0 - K + lane;
  const int ya = g.lo[1] + chunk * A.rows_per_chunk;
  const int yb = min(ya + A.rows_per_chunk, g.hi[1] + 1);
  const bool icol = i >= g.lo[0] && i <= g.hi[0];
  const bool xcol = i >= g.lo[0] - 1 && i <= g.hi[0];
  const bool cols = lane >= K && lane < K + TX;
  const float inv_dt = A.inv_dt;

  // E[n][0]: iteration-n correction of row t - n, E[n][1]: of row t - n - 1
  // QX[n][0..2]: iteration-n x faces of rows t - n + 1, t - n, t - n - 1
  // QY[n][0..1]: iteration-n y faces of rows t - n, t - n - 1
  // (level 0: the loaded values, rows t and t - 1)
  float E[K + 1][2], QX[K + 1][3], QY[K + 1][2];
  // coefficients: FX[n] x faces of row t - n (n < K); CY[n] y faces and cell of row t - n (n <= K)
  float FX[K][3], CY[K + 1][10];
  float QPX[3], QPY[2];  // iteration-K q' (check launches)
#pragma unroll
  for (int n = 0; n <= K; ++n) {
    E[n][0] = E[n][1] = 0.f;
    QX[n][0] = QX[n][1] = QX[n][2] = 0.f;
    QY[n][0] = QY[n][1] = 0.f;
#pragma unroll
    for (int c = 0; c < 10; ++c) CY[n][c] = 0.f;
  }
#pragma unroll
  for (int n = 0; n < K; ++n) FX[n][0] = FX[n][1] = FX[n][2] = 0.f;
  QPX[0] = QPX[1] = QPX[2] = QPY[0] = QPY[1] = 0.f;
  double sq = 0.0, smax = 0.0;

  const int t0 = ya - K, t1 = yb - 1 + K;
  RowIn nxt;
  load_row(A, i, t0, icol, xcol, nxt);
  for (int t = t0; t <= t1; ++t) {
    // ---- shift the histories by one row
#pragma unroll
    for (int n = K; n >= 0; --n) {
      E[n][1] = E[n][0];
      QX[n][2] = QX[n][1];
      QX[n][1] = QX[n][0];
      QY[n][1] = QY[n][0];
    }
#pragma unroll
    for (int n = K; n >= 1; --n) {
#pragma unroll
      for (int c = 0; c < 10; ++c) CY[n][c] = CY[n - 1][c];
    }
#pragma unroll
    for (int n = K - 1; n >= 1; --n) {
      FX[n][0] = FX[n - 1][0];
      FX[n][1] = FX[n - 1][1];
      FX[n][2] = FX[n - 1][2];
    }
    QPX[2] = QPX[1];
    QPX[1] = QPX[0];
    QPY[1] = QPY[0];
    // ---- row t enters at level 0; row t + 1 is fetched while it is used
    E[0][0] = nxt.e;
    QX[0][0] = nxt.qx;
    QY[0][0] = nxt.qy;
    FX[0][0] = nxt.px;
    FX[0][1] = nxt.mx;
    FX[0][2] = nxt.wx;
    CY[0][0] = nxt.py;
    CY[0][1] = nxt.my;
    CY[0][2] = nxt.wy;
    CY[0][3] = nxt.rb;
    CY[0][4] = nxt.cc;
    CY[0][5] = nxt.m0;
    CY[0][6] = nxt.m1;
    CY[0][7] = nxt.m2;
    CY[0][8] = nxt.m3;
    CY[0][9] = nxt.m4;
    if (t < t1) load_row(A, i, t + 1, icol, xcol, nxt);

    // ---- stages
#pragma unroll
    for (int n = 1; n <= K; ++n) {
      const int j = t - n;  // row of this stage's cells and y faces
      // x faces of row j + 1
      {
        const float e0 = E[n - 1][0];
        const float e1 = dpp_from_right(e0);
        const float qold = n == 1 ? QX[0][0] : QX[n - 1][1];
        const float qp = FX[n - 1][0] * e0 - FX[n - 1][1] * e1;
        QX[n][0] = qp + FX[n - 1][2] * (qold - qp);
        if (CHECK && n == K) QPX[0] = qp;
      }
      // y faces of row j
      {
        const float e0 = E[n - 1][1], e1 = E[n - 1][0];
        const float qold = QY[n - 1][1];
        const float qp = CY[n][0] * e0 - CY[n][1] * e1;
        QY[n][0] = qp + CY[n][2] * (qold - qp);
        if (CHECK && n == K) QPY[0] = qp;
      }
      // cells of row j
      {
        const float xc = QX[n][1], xn = QX[n][0], xs = QX[n][2];
        const float yc = QY[n][0], ys = QY[n][1];
        const float xw = dpp_from_left(xc), xnw = dpp_from_left(xn), xsw = dpp_from_left(xs);
        const float ye = dpp_from_right(yc), yes = dpp_from_right(ys), yw = dpp_from_left(yc), yws = dpp_from_left(ys);
        const float cross = ((xn + xnw) - (xs + xsw)) + ((ye + yes) - (yw + yws));
        const float D = CY[n][5] * xc + CY[n][6] * xw + CY[n][7] * yc + CY[n][8] * ys + CY[n][9] * cross;
        const float eo = E[n - 1][1];
        const float en = eo + CY[n][4] * (CY[n][3] - D - eo * inv_dt);
        const bool dom = icol && j >= g.lo[1] && j <= g.hi[1];
        E[n][0] = dom ? en : eo;
        if (CHECK && n == K) {
          // the residual of iteration K with its q' (all lanes shuffle; only
          // the strip's domain cells of the chunk's rows are summed)
          const float pxc = QPX[1], pxn = QPX[0], pxs = QPX[2];
          const float pyc = QPY[0], pys = QPY[1];
          const float pxw = dpp_from_left(pxc), pxnw = dpp_from_left(pxn), pxsw = dpp_from_left(pxs);
          const float pye = dpp_from_right(pyc), pyes = dpp_from_right(pys), pyw = dpp_from_left(pyc), pyws = dpp_from_left(pys);
          const float crossp = ((pxn + pxnw) - (pxs + pxsw)) + ((pye + pyes) - (pyw + pyws));
          const float Dp = CY[n][5] * pxc + CY[n][6] * pxw + CY[n][7] * pyc + CY[n][8] * pys + CY[n][9] * crossp;
          const double res = (dom && cols && j >= ya && j < yb) ? (double)(CY[n][3] - en * inv_dt - Dp) : 0.0;
          sq += res * res;
          smax = nanmax(smax, fabs(res));
          if (res != res) smax = res;
        }
      }
    }

    // ---- writes: iteration K of cells and y faces of row t - K, x faces of row t - K + 1
    const int jc = t - K, jx = t - K + 1;
    if (jc >= ya && jc < yb && cols && icol) {
      const int64_t c = at(g, i, jc, 0);
      A.e_dst[c] = E[K][0];
      A.qy_dst[c] = QY[K][0];
    }
    if (jc == g.lo[1] - 1 && cols && i <= g.hi[0]) A.qy_dst[at(g, i, jc, 0)] = QY[K][0];
    if (jx >= ya && jx < yb && ((cols && icol) || i == g.lo[0] - 1)) A.qx_dst[at(g, i, jx, 0)] = QX[K][0];
  }
  if (CHECK) {
    sq = wave_sum(sq);
    smax = wave_max(smax);
    if (lane == 0) {
      A.part_sum[part] = sq;
      A.part_max[part] = smax;
    }
  }
}

int lin2d_march_blocks(int strips, int chunks) {
  return 8 * ((strips + MWAVES - 1) / MWAVES) * ((chunks + 7) / 8);
}

int lin2d_march_parts(int strips, int chunks) { return MWAVES * lin2d_march_blocks(strips, chunks); }

void launch_sweep2d_march(const Lin2DArgs& A, int kmax, bool check, hipStream_t s) {
  const int nb = lin2d_march_blocks(A.strips, A.chunks);
  switch (kmax) {
    case 4:
      if (check)
        hipLaunchKernelGGL((sweep2d_lin_march_kernel<4, true>), dim3(nb), dim3(MWAVES * WAVE), 0, s, A);
      else
        hipLaunchKernelGGL((sweep2d_lin_march_kernel<4, false>), dim3(nb), dim3(MWAVES * WAVE), 0, s, A);
      break;
    case 8:
      if (check)
        hipLaunchKernelGGL((sweep2d_lin_march_kernel<8, true>), dim3(nb), dim3(MWAVES * WAVE), 0, s, A);
      else
        hipLaunchKernelGGL((sweep2d_lin_march_kernel<8, false>), dim3(nb), dim3(MWAVES * WAVE), 0, s, A);
      break;
    default:
      if (check)
        hipLaunchKernelGGL((sweep2d_lin_march_kernel<6, true>), dim3(nb), dim3(MWAVES * WAVE), 0, s, A);
      else
        hipLaunchKernelGGL((sweep2d_lin_march_kernel<6, false>), dim3(nb), dim3(MWAVES * WAVE), 0, s, A);
      break;
  }
}

}  // namespace cdiff
