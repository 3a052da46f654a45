// lin2d_ops.h -- the arithmetic of one face and one cell of the 2D FP32
// correction sweep in coefficient form.  Every 2D layout of the sweep (64-column
// tiles, wave-aligned tiles, marching strips) calls these, so that they give
// the same bits.  FP32 with fused multiply-adds: the sweep only has to reach
// its own fixed point, which the FP64 outer loop then corrects.
#pragma once
#include <hip/hip_runtime.h>

namespace cdiff {

// q' = P e0 - M e1 and the damped flux q <- q' + W (q_old - q')
__device__ __forceinline__ float lin_face(float P, float M, float W, float e0, float e1, float qold, float& qp) {
  qp = __builtin_fmaf(P, e0, -(M * e1));
  return __builtin_fmaf(W, qold - qp, qp);
}

// the cross terms of the divergence: ((x_n + x_nw) - (x_s + x_sw)) + ((y_e + y_es) - (y_w + y_ws))
__device__ __forceinline__ float lin_cross(float xn, float xnw, float xs, float xsw, float ye, float yes, float yw,
                                           float yws) {
  return ((xn + xnw) - (xs + xsw)) + ((ye + yes) - (yw + yws));
}

// lin_cross from its four pair sums: sx_n = x_n + x_nw, sx_s = x_s + x_sw,
// sy_e = y_e + y_es, sy_w = y_w + y_ws (bit for bit the same value)
__device__ __forceinline__ float lin_cross_sums(float sx_n, float sx_s, float sy_e, float sy_w) {
  return (sx_n - sx_s) + (sy_e - sy_w);
}

// D = m0 x_c + m1 x_w + m2 y_c + m3 y_s + m4 cross
__device__ __forceinline__ float lin_div(float m0, float m1, float m2, float m3, float m4, float xc, float xw, float yc,
                                         float ys, float cross) {
  return __builtin_fmaf(m4, cross, __builtin_fmaf(m3, ys, __builtin_fmaf(m2, yc, __builtin_fmaf(m1, xw, m0 * xc))));
}

// e <- e + C (r - D - e / dt)
__device__ __forceinline__ float lin_update(float e, float C, float r, float D, float inv_dt) {
  return __builtin_fmaf(C, __builtin_fmaf(-e, inv_dt, r - D), e);
}

// residual of the correction equation with the plain flux's divergence D'
__device__ __forceinline__ float lin_residual(float r, float e, float inv_dt, float Dp) {
  return __builtin_fmaf(-e, inv_dt, r) - Dp;
}

}  // namespace cdiff
