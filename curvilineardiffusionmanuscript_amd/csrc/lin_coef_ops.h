// lin_coef_ops.h -- the coefficients of the FP32 correction sweep, formed in
// FP32 by the fused outer passes of the defect-correction step.
//
// These numbers are stored in FP32 and only steer the correction sweep: its
// pseudo-time step C, its flux damping W and the Newton terms of its linear
// operator.  The step's answer is fixed by the FP64 residual that the outer
// pass evaluates, so they are formed with the hardware reciprocal and square
// root (v_rcp_f32, v_sqrt_f32: about 1 ulp) instead of IEEE FP64 division.
// The formulas are those of iteration_parameters.jl:5-14 (Re, dtau/rho,
// theta) and fluxes.jl:126-137 (W = theta_bar / (1 + theta_bar)).
#pragma once
#include <hip/hip_runtime.h>

#include "cdiff_dev.h"

namespace cdiff {

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Uniform constants of the iteration parameters, in FP32.
struct IterConst32 {
  float L2_over_dt;  // L^2 / dt
  float VL;          // V L
  float L_over_V;    // L / V
  float inv_dt;
  float dtau_scale, theta_scale;
};

__device__ __forceinline__ IterConst32 iter_const32(const IterConst& c, double dtau_scale, double theta_scale) {
  IterConst32 k;
  k.L2_over_dt = (float)(c.L * c.L / c.dt);
  k.VL = (float)(c.Vpdtau * c.L);
  k.L_over_V = (float)(c.L / c.Vpdtau);
  k.inv_dt = (float)(1.0 / c.dt);
  k.dtau_scale = (float)dtau_scale;
  k.theta_scale = (float)theta_scale;
  return k;
}

// theta and the cell coefficient C = dtau / (1 + dtau / dt) of the
// correction update, from alpha and rho.  alpha = 0 gives Re = inf, so
// dtau = 0 * inf = NaN and theta = 0, both passed through Julia's strong zero
// as in the FP64 iteration parameters.
__device__ __forceinline__ void iter_params32(const IterConst32& k, float rho, float a, float& C, float& theta) {
  const float pi = 3.14159265358979f;
  const float ra = fast_rcp(a);
  const float Re = pi + fast_sqrt(__builtin_fmaf(k.L2_over_dt * rho, ra, pi * pi));
  const float rRe = fast_rcp(Re);
  const float dtau = strong_zero<float>(k.VL * ra * rRe) * k.dtau_scale;
  theta = strong_zero<float>(k.L_over_V * rRe) * k.theta_scale;
  C = dtau * fast_rcp(__builtin_fmaf(dtau, k.inv_dt, 1.0f));
}

// d(alpha)/du for alpha ~ |u|^t_exp: t_exp alpha / u, 0 where that is not finite
__device__ __forceinline__ float dadu32(float t_exp, float a, float u) {
  const float d = t_exp * a * fast_rcp(u);
  return (u != 0.0f && isfinite(d)) ? d : 0.0f;
}

// face mean of theta and the damping W = theta_bar / (1 + theta_bar)
__device__ __forceinline__ float damping32(int harmonic, float t0, float t1) {
  const float tbar = harmonic ? (2.0f * t0 * t1) * fast_rcp(t0 + t1) : 0.5f * (t0 + t1);
  return tbar * fast_rcp(1.0f + tbar);
}

// Newton weights of the face mean's derivative: d abar / d a0 and d abar / d a1
// (1/2 each for the arithmetic mean, 2 a1^2 / s^2 and 2 a0^2 / s^2 for the
// harmonic one, s = a0 + a1, 0 where s = 0)
__device__ __forceinline__ void mean_weights32(int harmonic, float a0, float a1, float& d0, float& d1) {
  if (!harmonic) {
    d0 = d1 = 0.5f;
    return;
  }
  const float rs = fast_rcp(a0 + a1);
  const float w0 = a1 * rs, w1 = a0 * rs;
  const bool z = (a0 + a1) == 0.0f;
  d0 = z ? 0.0f : 2.0f * w0 * w0;
  d1 = z ? 0.0f : 2.0f * w1 * w1;
}

}  // namespace cdiff
