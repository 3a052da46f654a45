// sweep3d_ops.h -- per-face and per-cell arithmetic of the 3D APT sweep,
// shared by the 3D sweep kernels.
//
// The 3D cross terms 0.5*g*(0.5*(a+b) - 0.5*(c+d)) of the reference are
// evaluated as (0.25*g)*((a+b) - (c+d)): halving is exact in binary floating
// point, so the two forms give the same number (outside the subnormal range).
#pragma once
#include <limits>

#include "cdiff_dev.h"

namespace cdiff {

// flux_kernel! / fluxprime_kernel! (flux/fluxes.jl:5-26): face between lower
// cell (u0, a0, t0) and upper cell (u1, a1, t1).
template <typename T>
__device__ __forceinline__ void flux3(int harmonic, T q_old, T u1, T a1, T t1, T u0, T a0, T t0, T& qn, T& qp) {
  const T abar = face_mean<T>(harmonic, a0, a1);
  const T tbar = face_mean<T>(harmonic, t0, t1);
  const T du = u1 - u0;
  const T p = -abar * du;
  qn = (q_old * tbar + p) / (T(1) + tbar);
  qp = p;
}

template <typename T>
__device__ __forceinline__ bool isapprox_eps(T x, T y) {
  if (x == y) return true;
  if (!(isfinite(x) && isfinite(y))) return false;
  const T m = fabs(x) > fabs(y) ? fabs(x) : fabs(y);
  return fabs(x - y) <= std::numeric_limits<T>::epsilon() * m;
}

// Per-face values around a cell, gathered for one divergence evaluation.
template <typename T>
struct Around {
  T xc, xW, xJp, xJpW, xJm, xJmW, xKp, xKpW, xKm, xKmW;
  T yc, yJm, yE, yW, yEJm, yWJm, yKp, yKpJm, yKm, yKmJm;
  T zc, zKm, zE, zEKm, zW, zWKm, zJp, zJpKm, zJm, zJmKm;
};

template <typename T>
__device__ __forceinline__ T div3(const T* cf, const Around<T>& q) {
  const T a1 = cf[0] * (q.xc - q.xW);
  const T a2 = cf[1] * (q.yc - q.yJm);
  const T a3 = cf[2] * (q.zc - q.zKm);
  const T qi_eta = cf[3] * ((q.xJp + q.xJpW) - (q.xJm + q.xJmW));
  const T qi_zeta = cf[4] * ((q.xKp + q.xKpW) - (q.xKm + q.xKmW));
  const T qj_xi = cf[3] * ((q.yE + q.yEJm) - (q.yW + q.yWJm));
  const T qj_zeta = cf[5] * ((q.yKp + q.yKpJm) - (q.yKm + q.yKmJm));
  const T qk_eta = cf[5] * ((q.zJp + q.zJpKm) - (q.zJm + q.zJmKm));
  const T qk_xi = cf[4] * ((q.zE + q.zEKm) - (q.zW + q.zWKm));
  const T m1 = cf[6] * (T(0.5) * (q.xc + q.xW));
  const T m2 = cf[7] * (T(0.5) * (q.yc + q.yJm));
  const T m3 = cf[8] * (T(0.5) * (q.zc + q.zKm));
  return a1 + a2 + a3 + qi_eta + qi_zeta + qj_xi + qj_zeta + qk_eta + qk_xi + m1 + m2 + m3;
}

// div3 with the zeta-neighbour face pairs of planes k-1 and k+1 given as their
// sums, (X(0,0,-1) + X(-1,0,-1)) etc., evaluated in the order of div3.
template <typename T>
__device__ __forceinline__ T div3_sums(const T* cf, const Around<T>& q, T xKpS, T xKmS, T yKpS, T yKmS) {
  const T a1 = cf[0] * (q.xc - q.xW);
  const T a2 = cf[1] * (q.yc - q.yJm);
  const T a3 = cf[2] * (q.zc - q.zKm);
  const T qi_eta = cf[3] * ((q.xJp + q.xJpW) - (q.xJm + q.xJmW));
  const T qi_zeta = cf[4] * (xKpS - xKmS);
  const T qj_xi = cf[3] * ((q.yE + q.yEJm) - (q.yW + q.yWJm));
  const T qj_zeta = cf[5] * (yKpS - yKmS);
  const T qk_eta = cf[5] * ((q.zJp + q.zJpKm) - (q.zJm + q.zJmKm));
  const T qk_xi = cf[4] * ((q.zE + q.zEKm) - (q.zW + q.zWKm));
  const T m1 = cf[6] * (T(0.5) * (q.xc + q.xW));
  const T m2 = cf[7] * (T(0.5) * (q.yc + q.yJm));
  const T m3 = cf[8] * (T(0.5) * (q.zc + q.zKm));
  return a1 + a2 + a3 + qi_eta + qi_zeta + qj_xi + qj_zeta + qk_eta + qk_xi + m1 + m2 + m3;
}

}  // namespace cdiff
