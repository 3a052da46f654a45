// column3d.h -- the column decomposition shared by the k-marching 3D kernels
// (sweep3d_col.hip, sweep3d_lin.hip: 64 x 8; dc_outer3d.hip: 64 x
// DC_OUTER3D_LY): a block of one thread per column cell owns a 64 x LY
// column of cells and marches up a chunk of planes, holding per plane the
// column plus a one-cell ring ((64 + 2) x (LY + 2) cells).
#pragma once
#include "cdiff_dev.h"

namespace cdiff {

// A column of 64 x LY cells (LY * 64 threads) and its ringed plane.
template <int LY>
struct Col3 {
  static constexpr int X = 64, Y = LY;
  static constexpr int RX = X + 2, RY = Y + 2, R = RX * RY;  // ringed plane
  static constexpr int T = X * Y;
  static constexpr int RING = R - T;
  static_assert(RING <= T, "one ring cell per thread at most");
};

constexpr int L3X = Col3<8>::X, L3Y = Col3<8>::Y;
constexpr int R3X = Col3<8>::RX, R3Y = Col3<8>::RY, R3 = Col3<8>::R;
constexpr int L3T = Col3<8>::T;
constexpr int RING = Col3<8>::RING;  // 148

// region cell (lx, ly) of a thread: slot 0 its column cell, slot 1 a ring
// cell (tid < RING)
template <int LY = 8>
__device__ __forceinline__ bool region_cell(int slot, int tid, int& lx, int& ly) {
  using C = Col3<LY>;
  if (slot == 0) {
    lx = tid % C::X + 1;
    ly = tid / C::X + 1;
    return true;
  }
  int k = tid;
  if (k < C::RX) {
    lx = k;
    ly = 0;
    return true;
  }
  k -= C::RX;
  if (k < C::RX) {
    lx = k;
    ly = C::RY - 1;
    return true;
  }
  k -= C::RX;
  if (k < C::Y) {
    lx = 0;
    ly = 1 + k;
    return true;
  }
  k -= C::Y;
  if (k < C::Y) {
    lx = C::RX - 1;
    ly = 1 + k;
    return true;
  }
  return false;
}

// column tile of block b (XCD-aware order: consecutive tiles on one XCD)
struct Column3D {
  int x0, y0, z0, z1;
};
template <int LY = 8>
__device__ __forceinline__ bool column_of(const Geo& g, int ntiles, int chunks, int stripsI, int ch, Column3D& c) {
  const int nb = gridDim.x, per = nb >> 3, b = blockIdx.x;
  const int tile = (b & 7) * per + (b >> 3);
  if (tile >= ntiles) return false;
  const int kc = tile % chunks;
  const int rest = tile / chunks;
  const int si = rest % stripsI, sj = rest / stripsI;
  c.x0 = g.lo[0] + si * Col3<LY>::X;
  c.y0 = g.lo[1] + sj * LY;
  c.z0 = g.lo[2] + kc * ch;
  c.z1 = min(c.z0 + ch - 1, g.hi[2]);
  return true;
}

}  // namespace cdiff
