// sweep2d_march.hip -- the FP32 correction sweep of sweep2d_lin.hip (same
// coefficient form, same arithmetic per face and per cell: lin2d_ops.h), K
// iterations per launch as a pipeline that marches up the rows.
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
// stages run in order.  Against the tiled sweep, the ring is re-read only
// along x (64 / (64 - 2K)) and 2K rows per chunk, and there are no barriers;
// the neighbouring strips of a chunk run on one XCD and share the cache lines
// at their edges in its L2.
//
// Registers.  A row's coefficients are used by K consecutive steps, so they
// sit in rings of K + 1 row slots (x faces of rows t + 1 .. t - K + 1, y faces
// and cells of rows t .. t - K): the row loop is unrolled K + 1 times, so every
// slot index is a compile-time constant and no coefficient is ever moved.
// The x faces and the correction of row t + 1 are fetched during step t (stage
// 1 needs them first), the y faces and cells of row t during step t (first
// needed at step t + 1).
//
// Loads.  Every coefficient array (P, M, W of both face sets, C, the right-hand
// side, the folded metrics) holds 0 outside its valid range, so a lane reads
// without a mask: from its own cell when that lies in [lo - 1, hi + 1] on
// both axes, else from column (or row) hi + 1, where every array read here
// is 0.  The correction and its fluxes are masked after the load (their
// halo entries are not kept at 0 by every writer).
//
// Bound.  HBM: at 4096^2 (six iterations per launch, 53 us per iteration) a
// copy without the arithmetic takes as long, one without the loads 27 us.  Its
// access pattern -- each wavefront reads 256 bytes of 16 arrays per row, the
// strips of a chunk in step, chunks apart -- moves ~4.1-4.5 TB/s (a
// grid-stride read of the same arrays, 6.2 TB/s); neither more rows in flight
// per wavefront nor interleaved coefficient records change that much, and
// 12 % fewer VALU instructions gain 0.5 %.  A launch reads its 76 bytes per
// cell whatever its depth, so the default runs eight iterations per launch
// (226 VGPRs, two wavefronts per SIMD), and each chunk re-reads 2K rows, so
// the chunks are as tall as one round of resident wavefronts allows (4096^2:
// 23 chunks of 179 rows, 39 us per iteration; 45.5 with chunks of 64 rows in
// three rounds).
//
// Values outside the dependency cone of the strip's cells (the ring's edge
// lanes and the first rows of a chunk) are computed from incomplete
// neighbours and never reach a written cell: after stage n, lanes
// n .. 63 - n and rows from ya - K + n up are exact.
#include <algorithm>
#include <atomic>
#include <type_traits>

#include "cdiff_dev.h"
#include "lin2d_ops.h"
#include "sweep_args.h"

namespace cdiff {

namespace {

constexpr int MWAVES = 4;  // independent wavefronts per block

// east / west neighbour by a DPP wavefront shift by one lane (the edge lane
// reads 0; edge lanes lie outside every written cell's dependency cone)
__device__ __forceinline__ float dpp_from_left(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));  // wave_shr:1
}
__device__ __forceinline__ float dpp_from_right(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));  // wave_shl:1
}

// a float at a byte offset from a (wave-uniform) array base: a 32-bit offset
// register against the base in scalar registers
__device__ __forceinline__ float ldf(const float* __restrict__ base, uint32_t byte_offset) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_offset);
}

__device__ __forceinline__ void stf(float* __restrict__ base, uint32_t byte_offset, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + byte_offset) = v;
}

}  // namespace

template <int K, bool CHECK, int PF>
__global__ __launch_bounds__(MWAVES * WAVE) __attribute__((amdgpu_waves_per_eu(PF == 1 && K <= 6 ? 3 : 2))) void
sweep2d_lin_march_kernel(Lin2DArgs A) {
  constexpr int TX = WAVE - 2 * K;
  constexpr int R = K + PF;  // row slots of the coefficient rings (PF rows fetched ahead)
  if (A.ctrl->done) return;
  const Geo& g = A.g;
  const int lane = threadIdx.x & (WAVE - 1);
  // XCD-aware split: block b runs on XCD b mod 8 and the grid is a multiple
  // of 8 blocks; the blocks of one XCD take one contiguous range of the
  // (chunk, strip) items, chunk by chunk and strips in order, so strips that
  // share the cache lines at their edges share that XCD's L2 (they march up
  // the same rows together)
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = blockIdx.x;
  const int item = ((b & 7) * (int)(gridDim.x >> 3) + (b >> 3)) * MWAVES + w;
  const int chunk = item / A.strips;
  const int strip = item - chunk * A.strips;
  const int part = b * MWAVES + w;
  if (chunk >= A.chunks) {
    if (CHECK && lane == 0) {
      A.part_sum[part] = 0.0;
      A.part_max[part] = 0.0;
    }
    return;
  }
  const int ilo = g.lo[0], ihi = g.hi[0], jlo = g.lo[1], jhi = g.hi[1];
  const int x0 = ilo + strip * TX;
  const int i = x0 - K + lane;
  const int ya = jlo + chunk * A.rows_per_chunk;
  const int yb = min(ya + A.rows_per_chunk, jhi + 1);
  const bool icol = i >= ilo && i <= ihi;
  const bool xcol = i >= ilo - 1 && i <= ihi;
  const bool cols = lane >= K && lane < K + TX;
  const int il = (i >= ilo - 1 && i <= ihi + 1) ? i : ihi + 1;  // the column this lane reads
  const float inv_dt = A.inv_dt;
  const bool zero_src = A.zero_src != 0;
  // byte offset of (il, r) in every FP32 array: rows outside [lo - 1, hi + 1]
  // read row hi + 1
  auto offset = [&](int r) -> uint32_t {
    const int rl = (r >= jlo - 1 && r <= jhi + 1) ? r : jhi + 1;
    return (uint32_t)(g.off + il + g.pitch * (int64_t)rl) * 4u;
  };

  // coefficient rings: FX[s] the x faces (P, M, W) of the row in slot s and
  // its correction and fluxes (e, q_x, q_y; used when the row enters), CY[s]
  // its y faces (P, M, W), right-hand side, C and folded metrics m0..m4; row r
  // sits in slot (r - t0) mod R
  float FX[R][6], CY[R][10];
  // E[n][0]: iteration-n correction of row t - n, E[n][1]: of row t - n - 1
  // QX[n][0..1]: iteration-n x faces of rows t - n + 1, t - n
  // QY[n][0..1]: iteration-n y faces of rows t - n, t - n - 1
  // (level 0: the loaded values, rows t and t - 1)
  // SX[n][0..2]: x_n + x_nw of the x faces of rows t - n + 1 .. t - n - 1 (the
  // cross term's sums, formed once per face row and used again two rows up)
  float E[K + 1][2], QX[K + 1][2], QY[K + 1][2], SX[K + 1][3];
  float QPX[2], QPY[2], SPX[3];  // iteration-K q' (check launches)
#pragma unroll
  for (int s = 0; s < R; ++s) {
#pragma unroll
    for (int c = 0; c < 6; ++c) FX[s][c] = 0.f;
#pragma unroll
    for (int c = 0; c < 10; ++c) CY[s][c] = 0.f;
  }
#pragma unroll
  for (int n = 0; n <= K; ++n) {
    E[n][0] = E[n][1] = 0.f;
    QX[n][0] = QX[n][1] = 0.f;
    QY[n][0] = QY[n][1] = 0.f;
    SX[n][0] = SX[n][1] = SX[n][2] = 0.f;
  }
  QPX[0] = QPX[1] = QPY[0] = QPY[1] = 0.f;
  SPX[0] = SPX[1] = SPX[2] = 0.f;
  double sq = 0.0, smax = 0.0;

  // x faces and the correction of row r, as loaded (the correction and its
  // fluxes are masked when the row enters the pipeline, one step later: a
  // select right after the load would wait for it there)
  auto load_x = [&](int r, float* fx) {
    const uint32_t o = offset(r);
    fx[0] = ldf(A.fx[0], o);
    fx[1] = ldf(A.fx[1], o);
    fx[2] = ldf(A.fx[2], o);
    fx[3] = ldf(A.e_src, o);
    fx[4] = ldf(A.qx_src, o);
    fx[5] = ldf(A.qy_src, o);
  };
  // y faces and cell coefficients of row r
  auto load_y = [&](int r, float* cy) {
    const uint32_t o = offset(r);
    cy[0] = ldf(A.fy[0], o);
    cy[1] = ldf(A.fy[1], o);
    cy[2] = ldf(A.fy[2], o);
    cy[3] = ldf(A.base, o);
    cy[4] = ldf(A.cc, o);
    cy[5] = ldf(A.m[0], o);
    cy[6] = ldf(A.m[1], o);
    cy[7] = ldf(A.m[2], o);
    cy[8] = ldf(A.m[3], o);
    cy[9] = ldf(A.m[4], o);
  };

  const int t0 = ya - K, t1 = yb - 1 + K;
#pragma unroll
  for (int d = 0; d < PF; ++d) {
    if (t0 + d <= t1) load_x(t0 + d, FX[d]);
    if (d >= 1 && t0 + d - 1 < t1) load_y(t0 + d - 1, CY[d - 1]);
  }

  auto step = [&](auto Pc, int t) {
    constexpr int P = decltype(Pc)::value;
    // ---- shift the correction and flux histories by one row
#pragma unroll
    for (int n = K; n >= 0; --n) {
      E[n][1] = E[n][0];
      QX[n][1] = QX[n][0];
      QY[n][1] = QY[n][0];
      SX[n][2] = SX[n][1];
      SX[n][1] = SX[n][0];
    }
    QPX[1] = QPX[0];
    QPY[1] = QPY[0];
    SPX[2] = SPX[1];
    SPX[1] = SPX[0];
    // ---- the loads of the next rows, in flight while the stages run
    if (t + PF - 1 < t1) load_y(t + PF - 1, CY[(P + PF - 1) % R]);
    if (t + PF <= t1) load_x(t + PF, FX[(P + PF) % R]);
    // ---- row t enters at level 0: its correction and fluxes masked to their
    //      valid entries (on the first launch of a solve they are 0)
    {
      const bool rin = !zero_src && t >= jlo && t <= jhi;
      E[0][0] = (icol && rin) ? FX[P][3] : 0.f;
      QX[0][0] = (xcol && rin) ? FX[P][4] : 0.f;
      QY[0][0] = (icol && !zero_src && t >= jlo - 1 && t <= jhi) ? FX[P][5] : 0.f;
    }

    // ---- stages
#pragma unroll
    for (int n = 1; n <= K; ++n) {
      const int j = t - n;  // row of this stage's cells and y faces
      const float* fx = FX[(P - n + 1 + 2 * R) % R];  // x faces of row j + 1
      const float* cy = CY[(P - n + 2 * R) % R];      // y faces and cells of row j
      // x faces of row j + 1
      {
        const float e0 = E[n - 1][0];
        const float e1 = dpp_from_right(e0);
        const float qold = n == 1 ? QX[0][0] : QX[n - 1][1];
        float qp;
        QX[n][0] = lin_face(fx[0], fx[1], fx[2], e0, e1, qold, qp);
        if (CHECK && n == K) QPX[0] = qp;
      }
      // y faces of row j
      {
        float qp;
        QY[n][0] = lin_face(cy[0], cy[1], cy[2], E[n - 1][1], E[n - 1][0], QY[n - 1][1], qp);
        if (CHECK && n == K) QPY[0] = qp;
      }
      // cells of row j
      {
        const float xc = QX[n][1], xn = QX[n][0];
        const float yc = QY[n][0], ys = QY[n][1];
        // the sums of lin_cross formed lane-wise and shifted afterwards: the
        // east pair y_e + y_es is the east lane's y_c + y_s, the west pair the
        // west lane's, and x_s + x_sw is this stage's x_n + x_nw of two rows
        // below -- the same operands in the same order, so the same bits
        SX[n][0] = xn + dpp_from_left(xn);
        const float sy = yc + ys;
        const float cross = lin_cross_sums(SX[n][0], SX[n][2], dpp_from_right(sy), dpp_from_left(sy));
        const float D = lin_div(cy[5], cy[6], cy[7], cy[8], cy[9], xc, dpp_from_left(xc), yc, ys, cross);
        const float eo = E[n - 1][1];
        const float en = lin_update(eo, cy[4], cy[3], D, inv_dt);
        // no mask: outside the domain the correction was loaded as +0 and C,
        // the right-hand side and the metric terms are 0, so the update
        // returns +0 again as long as the fluxes are finite (a solve whose
        // fluxes are not is dropped as diverged)
        E[n][0] = en;
        const bool dom = icol && (unsigned)(j - jlo) <= (unsigned)(jhi - jlo);
        if (CHECK && n == K) {
          // the residual of iteration K with its q' (all lanes shuffle; only
          // the strip's domain cells of the chunk's rows are summed)
          const float pxc = QPX[1], pxn = QPX[0];
          const float pyc = QPY[0], pys = QPY[1];
          SPX[0] = pxn + dpp_from_left(pxn);
          const float spy = pyc + pys;
          const float crossp = lin_cross_sums(SPX[0], SPX[2], dpp_from_right(spy), dpp_from_left(spy));
          const float Dp = lin_div(cy[5], cy[6], cy[7], cy[8], cy[9], pxc, dpp_from_left(pxc), pyc, pys, crossp);
          const double res = (dom && cols && j >= ya && j < yb) ? (double)lin_residual(cy[3], en, inv_dt, Dp) : 0.0;
          sq += res * res;
          smax = nanmax(smax, fabs(res));
          if (res != res) smax = res;
        }
      }
    }

    // ---- writes: iteration K of cells and y faces of row t - K, x faces of row t - K + 1
    //      (32-bit byte offsets from the lane's own column; only lanes of
    //      valid cells and faces store)
    const int jc = t - K, jx = t - K + 1;
    const uint32_t oc = (uint32_t)(g.off + i + g.pitch * (int64_t)jc) * 4u;
    if (jc >= ya && jc < yb && cols && icol) {
      stf(A.e_dst, oc, E[K][0]);
      stf(A.qy_dst, oc, QY[K][0]);
    }
    if (jc == jlo - 1 && cols && i <= ihi) stf(A.qy_dst, oc, QY[K][0]);
    if (jx >= ya && jx < yb && ((cols && icol) || i == ilo - 1))
      stf(A.qx_dst, oc + (uint32_t)g.pitch * 4u, QX[K][0]);
  };

  int t = t0;
  while (march_phases<0, R>(step, t, t1)) {
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
  const int64_t blocks = ((int64_t)strips * chunks + MWAVES - 1) / MWAVES;
  return (int)(8 * ((blocks + 7) / 8));
}

int lin2d_march_resident_waves(int kmax) {
  // the fewer of the check and plain variants (the same register budget);
  // queried once per depth (the kernels' resources do not change)
  static std::atomic<int> cached[3] = {0, 0, 0};
  const int slot = kmax == 4 ? 0 : (kmax == 8 ? 1 : 2);
  if (const int v = cached[slot].load(std::memory_order_relaxed)) return v;
  int a = 0, c = 0;
  hipError_t e1, e2;
  switch (kmax) {
    case 4:
      e1 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, sweep2d_lin_march_kernel<4, true, 1>, MWAVES * WAVE, 0);
      e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, sweep2d_lin_march_kernel<4, false, 1>, MWAVES * WAVE, 0);
      break;
    case 8:
      e1 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, sweep2d_lin_march_kernel<8, true, 1>, MWAVES * WAVE, 0);
      e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, sweep2d_lin_march_kernel<8, false, 1>, MWAVES * WAVE, 0);
      break;
    default:
      e1 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, sweep2d_lin_march_kernel<6, true, 1>, MWAVES * WAVE, 0);
      e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, sweep2d_lin_march_kernel<6, false, 1>, MWAVES * WAVE, 0);
      break;
  }
  if (e1 != hipSuccess || e2 != hipSuccess) return 0;
  const int v = MWAVES * std::min(a, c);
  cached[slot].store(v, std::memory_order_relaxed);
  return v;
}

int lin2d_march_parts(int strips, int chunks) { return MWAVES * lin2d_march_blocks(strips, chunks); }

template <int PF>
void launch_march_pf(const Lin2DArgs& A, int kmax, bool check, hipStream_t s, int nb) {
  switch (kmax) {
    case 4:
      if (check)
        hipLaunchKernelGGL((sweep2d_lin_march_kernel<4, true, PF>), dim3(nb), dim3(MWAVES * WAVE), 0, s, A);
      else
        hipLaunchKernelGGL((sweep2d_lin_march_kernel<4, false, PF>), dim3(nb), dim3(MWAVES * WAVE), 0, s, A);
      break;
    case 8:
      if (check)
        hipLaunchKernelGGL((sweep2d_lin_march_kernel<8, true, PF>), dim3(nb), dim3(MWAVES * WAVE), 0, s, A);
      else
        hipLaunchKernelGGL((sweep2d_lin_march_kernel<8, false, PF>), dim3(nb), dim3(MWAVES * WAVE), 0, s, A);
      break;
    default:
      if (check)
        hipLaunchKernelGGL((sweep2d_lin_march_kernel<6, true, PF>), dim3(nb), dim3(MWAVES * WAVE), 0, s, A);
      else
        hipLaunchKernelGGL((sweep2d_lin_march_kernel<6, false, PF>), dim3(nb), dim3(MWAVES * WAVE), 0, s, A);
      break;
  }
}

void launch_sweep2d_march(const Lin2DArgs& A, int kmax, bool check, hipStream_t s) {
  // rows fetched ahead: 1 (4096^2, six iterations per launch: 53 us per
  // iteration; 2 and 3 rows ahead need 182 and 199 VGPRs, two wavefronts per
  // SIMD, and take 60 us; eight iterations per launch, the default: 39.0 us,
  // two rows ahead 39.3 us at 247 VGPRs)
  launch_march_pf<1>(A, kmax, check, s, lin2d_march_blocks(A.strips, A.chunks));
}

}  // namespace cdiff
