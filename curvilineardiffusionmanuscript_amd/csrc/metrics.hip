// metrics.hip -- the mesh metrics the APT reads, built on the device from the
// node coordinates (the step the reference takes from CurvilinearGrids.jl,
// consumed at src/pseudo_transient/PseudoTransientSolver.jl:122-153,
// mesh_metric_cache.jl:39-250 and flux_divergence.jl:11-14).
//
// Discretization in index space (logical cell width 1), FP64, the arithmetic
// of the host mesh module operation for operation so both give the same bits:
//   cell centre   d(x_c)/d(xi_a) = mean of the node differences of the cell's
//                 edges along a (2 edges in 2D, 4 in 3D); J = det; the inverse
//                 metrics d(xi_a)/d(x_c) = adjugate / J (cofactors in a fixed
//                 order)
//   face b+1/2    normal derivative (along b) = mean of the two adjacent cells'
//                 (the last layer keeps its own), tangential derivatives from
//                 the face's own nodes; J and the conservative metrics
//                 xi_hat = J * inverse
// Node arrays are the full ghost-inclusive node grid, (n0+1) x (n1+1)
// [x (n2+1)], first index fastest; outputs are in the solver's pitched layout.
#include "cdiff_dev.h"
#include "kernels.h"

namespace cdiff {

namespace {

struct Nodes {
  const double* x[3];
  int64_t s1, s2;  // node strides of j and k
};

__device__ __forceinline__ double nd(const Nodes& N, int c, int i, int j, int k) {
  return N.x[c][(int64_t)i + N.s1 * j + N.s2 * k];
}

// 2D forward metrics of full cell (i, j): xxi, xeta, yxi, yeta
__device__ __forceinline__ void fwd2(const Nodes& N, int i, int j, double& xxi, double& xeta, double& yxi,
                                     double& yeta) {
  const double x00 = nd(N, 0, i, j, 0), x10 = nd(N, 0, i + 1, j, 0), x01 = nd(N, 0, i, j + 1, 0),
               x11 = nd(N, 0, i + 1, j + 1, 0);
  const double y00 = nd(N, 1, i, j, 0), y10 = nd(N, 1, i + 1, j, 0), y01 = nd(N, 1, i, j + 1, 0),
               y11 = nd(N, 1, i + 1, j + 1, 0);
  xxi = 0.5 * ((x10 - x00) + (x11 - x01));
  xeta = 0.5 * ((x01 - x00) + (x11 - x10));
  yxi = 0.5 * ((y10 - y00) + (y11 - y01));
  yeta = 0.5 * ((y01 - y00) + (y11 - y10));
}

// 3D forward metric d(x_c)/d(xi_a) of full cell (i, j, k): mean of the four
// node differences along a, summed with the other two axes (ascending) as
// the outer and inner loop
__device__ __forceinline__ double fwd3(const Nodes& N, int c, int a, int i, int j, int k) {
  const int o0 = a == 0 ? 1 : 0, o1 = a == 2 ? 1 : 2;
  double acc = 0.0;
  for (int u = 0; u < 2; ++u)
    for (int v = 0; v < 2; ++v) {
      int hi[3] = {i, j, k}, lo[3] = {i, j, k};
      hi[a] += 1;
      hi[o0] += u;
      lo[o0] += u;
      hi[o1] += v;
      lo[o1] += v;
      acc = acc + (nd(N, c, hi[0], hi[1], hi[2]) - nd(N, c, lo[0], lo[1], lo[2]));
    }
  return 0.25 * acc;
}

// det and inverse of m[c][a] by cofactors, in the host module's order
__device__ __forceinline__ double det_inv3(const double (&m)[3][3], double (&inv)[3][3]) {
  const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
  const double c01 = m[1][0] * m[2][2] - m[1][2] * m[2][0];
  const double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  const double det = m[0][0] * c00 - m[0][1] * c01 + m[0][2] * c02;
  const double adj[3][3] = {
      {c00, m[0][2] * m[2][1] - m[0][1] * m[2][2], m[0][1] * m[1][2] - m[0][2] * m[1][1]},
      {m[1][2] * m[2][0] - m[1][0] * m[2][2], m[0][0] * m[2][2] - m[0][2] * m[2][0],
       m[0][2] * m[1][0] - m[0][0] * m[1][2]},
      {c02, m[0][1] * m[2][0] - m[0][0] * m[2][1], m[0][0] * m[1][1] - m[0][1] * m[1][0]}};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) inv[a][c] = adj[a][c] / det;
  return det;
}

}  // namespace

__global__ void k_node_metrics2d(Geo g, const double* __restrict__ X, const double* __restrict__ Y, int64_t s1,
                                 MetricOut out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int NI = g.n[0], NJ = g.n[1];
  if (e >= (int64_t)NI * NJ) return;
  const int i = (int)(e % NI), j = (int)(e / NI);
  Nodes N{{X, Y, nullptr}, s1, 0};
  const int64_t c = at(g, i, j, 0);
  double xxi, xeta, yxi, yeta;
  fwd2(N, i, j, xxi, xeta, yxi, yeta);
  const double J = xxi * yeta - xeta * yxi;
  out.cell[0][c] = yeta / J;
  out.cell[1][c] = -xeta / J;
  out.cell[2][c] = -yxi / J;
  out.cell[3][c] = xxi / J;
  // i+1/2 face
  {
    double xq = xxi, yq = yxi;
    if (i < NI - 1) {
      double a, b, cc, d;
      fwd2(N, i + 1, j, a, b, cc, d);
      xq = 0.5 * (xxi + a);
      yq = 0.5 * (yxi + cc);
    }
    const double xt = nd(N, 0, i + 1, j + 1, 0) - nd(N, 0, i + 1, j, 0);
    const double yt = nd(N, 1, i + 1, j + 1, 0) - nd(N, 1, i + 1, j, 0);
    const double x_xi = xq, x_eta = xt, y_xi = yq, y_eta = yt;
    out.edge[0][c] = y_eta;
    out.edge[1][c] = -x_eta;
    out.edge[2][c] = -y_xi;
    out.edge[3][c] = x_xi;
    out.edge[4][c] = x_xi * y_eta - x_eta * y_xi;
  }
  // j+1/2 face
  {
    double xq = xeta, yq = yeta;
    if (j < NJ - 1) {
      double a, b, cc, d;
      fwd2(N, i, j + 1, a, b, cc, d);
      xq = 0.5 * (xeta + b);
      yq = 0.5 * (yeta + d);
    }
    const double xt = nd(N, 0, i + 1, j + 1, 0) - nd(N, 0, i, j + 1, 0);
    const double yt = nd(N, 1, i + 1, j + 1, 0) - nd(N, 1, i, j + 1, 0);
    const double x_xi = xt, x_eta = xq, y_xi = yt, y_eta = yq;
    out.edge[5][c] = y_eta;
    out.edge[6][c] = -x_eta;
    out.edge[7][c] = -y_xi;
    out.edge[8][c] = x_xi;
    out.edge[9][c] = x_xi * y_eta - x_eta * y_xi;
  }
}

__global__ void k_node_metrics3d(Geo g, const double* __restrict__ X, const double* __restrict__ Y,
                                 const double* __restrict__ Z, int64_t s1, int64_t s2, MetricOut out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int NI = g.n[0], NJ = g.n[1], NK = g.n[2];
  if (e >= (int64_t)NI * NJ * NK) return;
  const int i = (int)(e % NI);
  const int64_t r = e / NI;
  const int j = (int)(r % NJ), k = (int)(r / NJ);
  Nodes N{{X, Y, Z}, s1, s2};
  const int64_t c = at(g, i, j, k);
  double M[3][3], inv[3][3];
#pragma unroll
  for (int cc = 0; cc < 3; ++cc)
#pragma unroll
    for (int a = 0; a < 3; ++a) M[cc][a] = fwd3(N, cc, a, i, j, k);
  det_inv3(M, inv);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) out.cell[a * 3 + cc][c] = inv[a][cc];
  const int idx[3] = {i, j, k};
  const int lim[3] = {NI, NJ, NK};
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    double Mf[3][3];
    // normal derivative: mean of this cell's and the next cell's along b
    const bool last = idx[b] == lim[b] - 1;
    const int in = i + (b == 0), jn = j + (b == 1), kn = k + (b == 2);
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) Mf[cc][b] = last ? M[cc][b] : 0.5 * (M[cc][b] + fwd3(N, cc, b, in, jn, kn));
    // tangential derivatives from the face's own four nodes (node index +1 along b)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (a == b) continue;
      const int o = 3 - a - b;
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        double acc = 0.0;
        for (int v = 0; v < 2; ++v) {
          int hi[3] = {i, j, k}, lo[3] = {i, j, k};
          hi[b] += 1;
          lo[b] += 1;
          hi[a] += 1;
          hi[o] += v;
          lo[o] += v;
          acc = acc + (nd(N, cc, hi[0], hi[1], hi[2]) - nd(N, cc, lo[0], lo[1], lo[2]));
        }
        Mf[cc][a] = 0.5 * acc;
      }
    }
    double invf[3][3];
    const double Jf = det_inv3(Mf, invf);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) out.edge[b * 10 + a * 3 + cc][c] = invf[a][cc] * Jf;
    out.edge[b * 10 + 9][c] = Jf;
  }
}

}  // namespace cdiff
