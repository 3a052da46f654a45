// warm.cpp -- the solver side of the learned warm start: features of the
// resident state, training targets, and the start of an FP32_DC step from the
// network's prediction (with the fallback to T^n).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "dc.h"
#include "kernels.h"
#include "learn.h"
#include "net.h"
#include "solver.h"

using namespace cdiff;

#define HIP_TRY(expr)                                 \
  do {                                                \
    hipError_t _e = (expr);                           \
    if (_e != hipSuccess) return hip_fail(_e, #expr); \
  } while (0)

#define LAUNCH_CHECK(name)                                        \
  do {                                                            \
    hipError_t _e = hipGetLastError();                            \
    if (_e != hipSuccess) return hip_fail(_e, "launch of " name); \
  } while (0)

int cdiff_solver::learn_features(double dt, void* X0, double* ts_dev) {
  if (!metrics_set) return fail(CDIFF_NOT_READY, "cdiff_set_metrics has not been called");
  if (!lin_ok() || !jac) return fail(CDIFF_BAD_ARG, "the learned warm start needs a 2D mesh without periodic faces");
  if (!(dt > 0)) return fail(CDIFF_BAD_ARG, "dt > 0 required");
  int rc = ensure_fp32();
  if (rc) return rc;
  if (!learn_scratch) {
    HIP_TRY(hipMalloc((void**)&learn_scratch, sizeof(double) * 1024));
    owned.push_back(learn_scratch);
  }
  cdiff_opts o;
  cdiff_default_opts(&o);
  double smin = spacing[0];
  for (int a = 1; a < ndim; ++a) smin = std::min(smin, spacing[a]);
  dc_Vpdtau = o.cfl * smin;
  // alpha(T^n) and the FP64 residual of T^n as the start of a step (u = u_prev = T^n)
  if ((rc = dc_outer_pass(dt, Tdev, nullptr, u[1], Tdev, 1, o, false))) return rc;
  hipLaunchKernelGGL(k_alpha_full, dim3((unsigned)(((int64_t)g.n[0] * g.n[1] * g.n[2] + 255) / 256)), dim3(256), 0,
                     stream, g, kap,
                     (const double*)rho, (const double*)nullptr, (const double*)Tdev, a64);
  LAUNCH_CHECK("k_alpha_full");
  launch_domain_maxabs(g, Tdev, learn_scratch, ts_dev, stream);
  launch_domain_maxabs32(g, base32, dt, learn_scratch, ts_dev + 1, stream);
  LAUNCH_CHECK("feature scales");
  FeatIn f{};
  f.g = g;
  f.harmonic = mean == CDIFF_MEAN_HARMONIC;
  f.dt = dt;
  f.dt_ratio = (have_prev && dt_prev > 0) ? dt / dt_prev : 0.0;
  f.j_ref = spacing[0] * spacing[1];
  f.u = Tdev;
  f.um1 = have_prev ? u_prev : Tdev;
  f.a = a64;
  f.g11 = coef[0];
  f.g22 = coef[1];
  f.jac = jac;
  f.r32 = base32;
  f.scales = ts_dev;
  launch_features(f, (__bf16*)X0, stream);
  LAUNCH_CHECK("k_features");
  return CDIFF_OK;
}

int cdiff_solver::learn_target(const double* ts_dev, float* tgt) {
  if (!have_prev) return fail(CDIFF_NOT_READY, "no completed step to take a target from");
  launch_target(g, Tdev, u_prev, ts_dev, tgt, stream);
  LAUNCH_CHECK("k_target");
  return CDIFF_OK;
}

// After the first outer pass of an FP32_DC step (u[1] = T^n with its BC
// ghosts, its residual norm in ctrl_h[2]): evaluate the residual of the
// predicted start and keep whichever start has the smaller residual.
namespace {

// Work in units of one FP32 correction sweep at the same mesh size, from
// measurements at 4096^2 on MI355X (all three scale with the cell count): an
// FP64 APT iteration ~10.5 sweeps (~385 vs 36.8 us), a fused outer pass of the
// defect correction ~10 sweeps (365 us), and one consultation of the network
// (the features with the FP64 residual of T^n, inference ~0.8 ms, the FP64
// residual of the predicted start) ~62 sweeps (~2.27 ms)
constexpr double WORK_FP64_ITER = 10.5, WORK_OUTER_PASS = 10.0, WORK_CONSULT = 62.0;

}  // namespace

// after a consultation that did not pay, the network sits out 15, 31, 63,
// ... (at most 255) steps before it is consulted again: one trial costs ~62
// sweeps, a few percent of the 15 or more steps it then sits out
void cdiff_solver::warm_back_off() {
  ++warm_rejects;
  warm_skip = std::min((16 << std::min(warm_rejects - 1, 4)) - 1, 255);
}

// A used prediction is judged by the work it saved: the step's own work per
// decade of residual reduction times the decades the prediction took off the
// start (log10 r(T^n) / r(prediction)), against the cost of the consultation.
// The APT's iterations per decade are nearly fixed by the mesh and dt, so this
// is the work a start from T^n would have added.  Every step also records its
// work per decade, from which the next step decides whether a consultation
// can pay at all (warm_pays).  Deterministic: iteration counts and residuals
// only.
void cdiff_solver::warm_account(const cdiff_stats& st, bool dc) {
  const double work = dc ? st.niter + WORK_OUTER_PASS * st.outer : WORK_FP64_ITER * st.niter;
  const double r_end = std::max(st.abs_err, 1e-300);
  const double r_start = st.warm_start ? warm_r_pred : step_r_start;
  const double decades = std::log10(std::max(r_start, 1e-300) / r_end);
  // (a step that starts converged, or nearly so, says nothing about the rate)
  if (std::isfinite(decades) && decades >= 0.5 && st.niter > 0) work_per_decade = work / decades;
  if (!st.warm_start) return;
  const double saved = work_per_decade * warm_gain;
  warm_decades = warm_gain;
  if (tune.verbose)
    std::fprintf(stderr, "[warm] saved ~%.1f sweep units (%.2f decades at %.1f per decade), consultation %.0f\n",
                 saved, warm_gain, work_per_decade, WORK_CONSULT);
  if (saved > WORK_CONSULT)
    warm_rejects = 0;
  else
    warm_back_off();
}

// Whether a consultation is expected to pay: the last step's work per decade
// times the decades a prediction is expected to take off (1 before the first
// used prediction; then what the last one took off) against its cost.  On the
// FP64 APT a decade costs ~110-270 sweep units at 4096^2 and a consultation
// pays; the mixed-precision step needs ~13-30 units per decade, so a
// prediction would have to take 2-5 decades off the start, more than the
// BF16 network's ~1.
bool cdiff_solver::warm_pays() const {
  const bool pays = work_per_decade * warm_decades > WORK_CONSULT;
  if (tune.verbose)
    std::fprintf(stderr, "[warm] expected saving %.1f sweep units (%.1f per decade x %.2f decades), consultation %.0f: %s\n",
                 work_per_decade * warm_decades, work_per_decade, warm_decades, WORK_CONSULT,
                 pays ? "consult" : "skip");
  return pays;
}

int cdiff_solver::warm_start_apply(double dt, const cdiff_opts& o, bool* used) {
  *used = false;
  warm_gain = 0.0;
  const double r0 = ctrl_h[2].abs_;
  launch_apply_prediction(g, u[1], net->delta, learn_ts, u[0], stream);
  LAUNCH_CHECK("k_apply_prediction");
  int rc = dc_residual(dt, 0, nullptr, 1, 1, o);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(&ctrl_h[2], ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  if (tune.verbose)
    std::fprintf(stderr, "[dc] residual of the predicted start %.3e (of T^n %.3e)\n", ctrl_h[2].abs_, r0);
  if (ctrl_h[2].abs_ < r0 && ctrl_h[2].status == CDIFF_OK) {
    // the smaller residual wins; whether it paid is judged at the step's end
    *used = true;
    warm_r_pred = ctrl_h[2].abs_;
    warm_gain = std::log10(r0 / ctrl_h[2].abs_);
    return CDIFF_OK;
  }
  warm_back_off();
  // fallback: start from T^n (in u_prev) again
  hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(1), 0, stream, ctrl, o.rtol, o.atol);
  hipLaunchKernelGGL(k_copy, dim3((g.total + 255) / 256), dim3(256), 0, stream, g.total, (const double*)u_prev, u[0]);
  LAUNCH_CHECK("warm start fallback");
  if ((rc = dc_residual(dt, 0, nullptr, 1, 1, o))) return rc;
  HIP_TRY(hipMemcpyAsync(&ctrl_h[2], ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  return CDIFF_OK;
}
