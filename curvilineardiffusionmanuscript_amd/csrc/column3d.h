// column3d.h -- the 64 x 8 column decomposition shared by the k-marching 3D
// kernels (sweep3d_lin.hip, dc_outer3d.hip): a 512-thread block owns a 64 x 8
// column of cells and marches up a chunk of planes, holding per plane the
// column plus a one-cell ring (66 x 10 cells).
#pragma once
#include "cdiff_dev.h"

namespace cdiff {

constexpr int L3X = 64, L3Y = 8;
constexpr int R3X = L3X + 2, R3Y = L3Y + 2, R3 = R3X * R3Y;  // ringed plane
constexpr int L3T = L3X * L3Y;
constexpr int RING = R3 - L3T;  // 148

// region cell (lx, ly) of a thread: slot 0 its column cell, slot 1 a ring
// cell (tid < RING)
__device__ __forceinline__ bool region_cell(int slot, int tid, int& lx, int& ly) {
  if (slot == 0) {
    lx = tid % L3X + 1;
    ly = tid / L3X + 1;
    return true;
  }
  int k = tid;
  if (k < R3X) {
    lx = k;
    ly = 0;
    return true;
  }
  k -= R3X;
  if (k < R3X) {
    lx = k;
    ly = R3Y - 1;
    return true;
  }
  k -= R3X;
  if (k < L3Y) {
    lx = 0;
    ly = 1 + k;
    return true;
  }
  k -= L3Y;
  if (k < L3Y) {
    lx = R3X - 1;
    ly = 1 + k;
    return true;
  }
  return false;
}

// column tile of block b (XCD-aware order: consecutive tiles on one XCD)
struct Column3D {
  int x0, y0, z0, z1;
};
__device__ __forceinline__ bool column_of(const Geo& g, int ntiles, int chunks, int stripsI, int ch, Column3D& c) {
  const int nb = gridDim.x, per = nb >> 3, b = blockIdx.x;
  const int tile = (b & 7) * per + (b >> 3);
  if (tile >= ntiles) return false;
  const int kc = tile % chunks;
  const int rest = tile / chunks;
  const int si = rest % stripsI, sj = rest / stripsI;
  c.x0 = g.lo[0] + si * L3X;
  c.y0 = g.lo[1] + sj * L3Y;
  c.z0 = g.lo[2] + kc * ch;
  c.z1 = min(c.z0 + ch - 1, g.hi[2]);
  return true;
}

}  // namespace cdiff
